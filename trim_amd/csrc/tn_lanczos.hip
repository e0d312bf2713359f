// Lowest eigenpair of a square matrix by the Lanczos recurrence.
//
// Restates lanczosEV (uni10_lapack_cpu.cpp:416-509), which the reference
// runs on the host only (its GPU entry throws, uni10_lapack_gpu.cu:681-687):
//   v0 = psi / ||psi||;  for it = 0, 1, ...:
//     w = A v_it - beta_{it-1} v_{it-1}          (dgemv, :444)
//     alpha_it = v_it . w;  w -= alpha_it v_it   (ddot, daxpy, :445-447)
//     beta_it = ||w||;  v_{it+1} = w / beta_it   (:449-455)
//     smallest eigenvalue d0 of the tridiagonal (alpha, beta) matrix
//     (dstev, :464-477); stop when |d0 - d0_prev| / max(|d0|, 1) <= tol,
//     when beta_it <= 1e-15 (the Krylov space closed) or at max_iter,
//     never before two steps;
//   psi = sum_k z_k v_k with z the tridiagonal eigenvector of d0 (:480-499).
//
// The O(n^2) and O(n * iter) work runs on the device: the matrix-vector
// product (one wave per row, HBM-bound), the dot/axpy/norm/scale of a step
// (one block, fixed reduction order) and the final combination.  What stays
// on the host is what the reference also does there: the eigenvalue of the
// small tridiagonal matrix after every step (here by Sturm bisection, and
// the final eigenvector by inverse iteration), fed by two doubles copied
// back per step (read in batches of steps).  Every vector lives in one pooled buffer of
// (max_iter + 1) x n doubles, as the reference's Vm.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "tn_internal.h"

namespace {

constexpr int kMvThreads = 128;            // 2 waves = 2 rows per block
constexpr int kMvRows = kMvThreads / 64;
// a matrix at least this large does not stay in the 256 MiB last-level
// cache between steps: its rows are read with streaming (non-temporal) loads
constexpr size_t kStreamBytes = size_t(256) << 20;
constexpr int kStepThreads = 1024;
constexpr double kBetaErr = 1e-15;         // uni10_lapack_cpu.cpp:419
constexpr int kMinIter = 2;                // uni10_lapack_cpu.cpp:418

// w[r] = A[r, :] . v - beta * vprev[r] (beta = *beta_prev, the previous
// step's, 0 at the first step); one wave per row, eight partial
// sums per lane (a fixed order), then a butterfly over the wave.  VEC: the
// rows are 16-byte aligned (n even, A aligned), so every lane loads pairs,
// 4 KiB per wave in flight per iteration; NT: the row loads are streaming
// (8192 x 8192: 94 -> 82 us, 5.7 -> 6.6 TB/s, with two rows per block).
template <bool VEC, bool NT>
__global__ void __launch_bounds__(kMvThreads) lz_matvec(const double* __restrict__ A, int64_t n,
                                                        const double* __restrict__ v,
                                                        const double* __restrict__ vprev,
                                                        const double* __restrict__ beta_prev,
                                                        double* __restrict__ w) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x & 63;
  const double beta = beta_prev ? *beta_prev : 0.0;
  const int64_t row = (int64_t)blockIdx.x * kMvRows + (threadIdx.x >> 6);
  if (row >= n) return;
  const double* __restrict__ a = A + row * n;
  double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int64_t j;
  if constexpr (VEC) {
    // pairs j, j+1 with j = 2 lane + 128 u, four pairs per lane per step
    for (j = 2 * lane; j + 384 + 1 < n; j += 512) {
      d2v x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const d2v* p = reinterpret_cast<const d2v*>(a + j + 128 * u);
        x[u] = NT ? __builtin_nontemporal_load(p) : *p;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double2 y = *reinterpret_cast<const double2*>(v + j + 128 * u);
        s[2 * u] = fma(x[u].x, y.x, s[2 * u]);
        s[2 * u + 1] = fma(x[u].y, y.y, s[2 * u + 1]);
      }
    }
    for (; j + 1 < n; j += 128) {
      const double2 x = *reinterpret_cast<const double2*>(a + j);
      const double2 y = *reinterpret_cast<const double2*>(v + j);
      s[0] = fma(x.x, y.x, s[0]);
      s[1] = fma(x.y, y.y, s[1]);
    }
  } else {
    for (j = lane; j + 448 < n; j += 512) {
#pragma unroll
      for (int u = 0; u < 8; ++u) s[u] = fma(a[j + 64 * u], v[j + 64 * u], s[u]);
    }
    for (; j < n; j += 64) s[0] = fma(a[j], v[j], s[0]);
  }
  double t = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
  if (lane == 0) w[row] = t - beta * vprev[row];
}

// Complex rows (interleaved re, im): w[r] = A[r, :] v - beta * vprev[r], no
// conjugation (zgemv 'T' on the column-major view, uni10_lapack_cpu.cpp:935).
// Pairs are 16-byte aligned; four partial sums per lane and component.
__global__ void __launch_bounds__(kMvThreads) lzc_matvec(const double* __restrict__ A, int64_t n,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ vprev,
                                                         const double* __restrict__ beta_prev,
                                                         double* __restrict__ w) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * kMvRows + (threadIdx.x >> 6);
  if (row >= n) return;
  const double beta = beta_prev ? *beta_prev : 0.0;
  const double2* __restrict__ a = reinterpret_cast<const double2*>(A + row * 2 * n);
  const double2* __restrict__ x = reinterpret_cast<const double2*>(v);
  double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t j = lane;
  for (; j + 192 < n; j += 256) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double2 p = a[j + 64 * u], q = x[j + 64 * u];
      sr[u] = fma(p.x, q.x, sr[u]);
      sr[u] = fma(-p.y, q.y, sr[u]);
      si[u] = fma(p.x, q.y, si[u]);
      si[u] = fma(p.y, q.x, si[u]);
    }
  }
  for (; j < n; j += 64) {
    const double2 p = a[j], q = x[j];
    sr[0] = fma(p.x, q.x, sr[0]);
    sr[0] = fma(-p.y, q.y, sr[0]);
    si[0] = fma(p.x, q.y, si[0]);
    si[0] = fma(p.y, q.x, si[0]);
  }
  double tr = (sr[0] + sr[1]) + (sr[2] + sr[3]);
  double ti = (si[0] + si[1]) + (si[2] + si[3]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    tr += __shfl_xor(tr, off);
    ti += __shfl_xor(ti, off);
  }
  if (lane == 0) {
    w[2 * row] = tr - beta * vprev[2 * row];
    w[2 * row + 1] = ti - beta * vprev[2 * row + 1];
  }
}

// Sum over the block in a fixed order: a butterfly inside each wave, then
// the wave totals in wave order.
__device__ double block_sum(double x, double* red) {
  constexpr int kWaves = kStepThreads / 64;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  const int t = threadIdx.x;
  if ((t & 63) == 0) red[t >> 6] = x;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) r += red[w];
  __syncthreads();
  return r;
}

// One step's vector work on v = v_it and w (holding A v_it - beta v_{it-1}):
// alpha = v.w; w -= alpha v; beta = ||w||; w /= beta when beta > 1e-15.
// ab[0] = alpha, ab[1] = beta.
__global__ void __launch_bounds__(kStepThreads) lz_step(const double* __restrict__ v, double* __restrict__ w,
                                                        int64_t n, double* __restrict__ ab) {
  __shared__ double red[kStepThreads];
  double p = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kStepThreads) p = fma(v[i], w[i], p);
  const double alpha = block_sum(p, red);
  double q = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kStepThreads) {
    const double x = fma(-alpha, v[i], w[i]);
    w[i] = x;
    q = fma(x, x, q);
  }
  const double beta = sqrt(block_sum(q, red));
  if (beta > kBetaErr) {
    const double r = 1.0 / beta;
    for (int64_t i = threadIdx.x; i < n; i += kStepThreads) w[i] *= r;
  }
  if (threadIdx.x == 0) {
    ab[0] = alpha;
    ab[1] = beta;
  }
}

// lz_step with the thread's R elements of v and w held in registers (n <=
// R * kStepThreads): one load and one store pass instead of three of each,
// the same operations in the same order, so the same bits.
template <int R>
__global__ void __launch_bounds__(kStepThreads) lz_step_reg(const double* __restrict__ v, double* __restrict__ w,
                                                            int64_t n, double* __restrict__ ab) {
  __shared__ double red[kStepThreads];
  double vr[R], wr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = threadIdx.x + (int64_t)r * kStepThreads;
    vr[r] = (i < n) ? v[i] : 0.0;
    wr[r] = (i < n) ? w[i] : 0.0;
  }
  double p = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) p = fma(vr[r], wr[r], p);
  const double alpha = block_sum(p, red);
  double q = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    wr[r] = fma(-alpha, vr[r], wr[r]);
    q = fma(wr[r], wr[r], q);
  }
  const double beta = sqrt(block_sum(q, red));
  const double rb = (beta > kBetaErr) ? 1.0 / beta : 1.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = threadIdx.x + (int64_t)r * kStepThreads;
    if (i < n) w[i] = (beta > kBetaErr) ? wr[r] * rb : wr[r];
  }
  if (threadIdx.x == 0) {
    ab[0] = alpha;
    ab[1] = beta;
  }
}

// dst = src * (1 / nrm[0])  (the start vector's normalisation)
__global__ void lz_scale_copy(const double* __restrict__ src, const double* __restrict__ nrm, int64_t n,
                              double* __restrict__ dst) {
  const double r = 1.0 / nrm[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i] * r;
}

__global__ void lz_norm(const double* __restrict__ x, int64_t n, double* __restrict__ out) {
  __shared__ double red[kStepThreads];
  double q = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kStepThreads) q = fma(x[i], x[i], q);
  const double s = block_sum(q, red);
  if (threadIdx.x == 0) out[0] = sqrt(s);
}

// out[j] = sum_{k < m} z[k] V[k][j], accumulated in k order
__global__ void lz_combine(const double* __restrict__ V, const double* __restrict__ z, int64_t m, int64_t n,
                           double* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int64_t k = 0; k < m; ++k) s = fma(z[k], V[k * n + j], s);
    out[j] = s;
  }
}

// ---- host: the small symmetric tridiagonal matrix (diagonal d, off-diagonal e)

// number of eigenvalues below x (Sturm sequence of the LDL^T pivots)
int sturm_below(const std::vector<double>& d, const std::vector<double>& e, int m, double x) {
  int c = 0;
  double q = 1.0;
  for (int i = 0; i < m; ++i) {
    const double off = (i == 0) ? 0.0 : e[i - 1] * e[i - 1] / q;
    q = d[i] - x - off;
    if (q == 0.0) q = -DBL_MIN;
    if (q < 0.0) ++c;
  }
  return c;
}

// smallest eigenvalue by bisection on the Gershgorin interval
double tridiag_min_eig(const std::vector<double>& d, const std::vector<double>& e, int m) {
  double lo = d[0], hi = d[0];
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  const double span = std::max(std::fabs(lo), std::fabs(hi));
  lo -= 2 * DBL_EPSILON * span + DBL_MIN;
  hi += 2 * DBL_EPSILON * span + DBL_MIN;
  for (int it = 0; it < 256; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_below(d, e, m, mid) >= 1)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

// unit eigenvector of lambda by inverse iteration, (T - lambda I) x = b solved
// by Gaussian elimination with partial pivoting; sign fixed so x[0] >= 0
std::vector<double> tridiag_eigvec(const std::vector<double>& d, const std::vector<double>& e, int m,
                                   double lambda) {
  double tn = 0.0;
  for (int i = 0; i < m; ++i) tn = std::max(tn, std::fabs(d[i]) + (i + 1 < m ? std::fabs(e[i]) : 0.0));
  const double tiny = std::max(tn, 1.0) * DBL_EPSILON;
  // LU of the tridiagonal with row interchanges: U has bands u0, u1, u2
  std::vector<double> u0(m), u1(m, 0.0), u2(m, 0.0), l(m, 0.0);
  std::vector<char> swp(m, 0);
  for (int i = 0; i < m; ++i) {
    u0[i] = d[i] - lambda;
    if (i + 1 < m) u1[i] = e[i];
  }
  for (int i = 0; i + 1 < m; ++i) {
    // rows i (u0[i], u1[i], u2[i]) and i+1 (e[i], d[i+1]-lambda, e[i+1])
    double a0 = u0[i], a1 = u1[i], a2 = u2[i];
    double b0 = e[i], b1 = u0[i + 1], b2 = (i + 2 < m) ? e[i + 1] : 0.0;
    if (std::fabs(b0) > std::fabs(a0)) {
      std::swap(a0, b0);
      std::swap(a1, b1);
      std::swap(a2, b2);
      swp[i] = 1;
    }
    if (a0 == 0.0) a0 = tiny;
    const double f = b0 / a0;
    l[i] = f;
    u0[i] = a0;
    u1[i] = a1;
    u2[i] = a2;
    u0[i + 1] = b1 - f * a1;
    u1[i + 1] = b2 - f * a2;
  }
  if (u0[m - 1] == 0.0) u0[m - 1] = tiny;
  std::vector<double> x(m, 1.0);
  for (int pass = 0; pass < 3; ++pass) {
    // forward: apply the row interchanges and multipliers
    for (int i = 0; i + 1 < m; ++i) {
      if (swp[i]) std::swap(x[i], x[i + 1]);
      x[i + 1] -= l[i] * x[i];
    }
    // back substitution with the banded U
    for (int i = m - 1; i >= 0; --i) {
      double s = x[i];
      if (i + 1 < m) s -= u1[i] * x[i + 1];
      if (i + 2 < m) s -= u2[i] * x[i + 2];
      x[i] = s / u0[i];
    }
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) nrm += x[i] * x[i];
    nrm = std::sqrt(nrm);
    for (int i = 0; i < m; ++i) x[i] /= nrm;
  }
  if (x[0] < 0.0)
    for (int i = 0; i < m; ++i) x[i] = -x[i];
  return x;
}

}  // namespace

// The complex form differs only in the product: alpha = Re(v^H w) (zdotc,
// :937-938), the axpy with a real alpha, the norm and the scaling are the
// real operations on the 2n interleaved doubles.
static int lanczos_impl(const double* A, int64_t n, bool cplx, double* psi, int64_t* iters, double err_tol,
                        double* E0, int* converged, void* stream) {
  TN_REQUIRE(A && psi && iters && E0 && converged, TN_EINVAL, "tn_lanczos: NULL argument");
  TN_REQUIRE(n >= 1, TN_EINVAL, "tn_lanczos: n=%lld", (long long)n);
  TN_REQUIRE(*iters > kMinIter, TN_EINVAL, "Maximum iteration number should be set greater than 2.");
  const int64_t M = *iters;
  const int64_t L = cplx ? 2 * n : n;   // doubles per vector
  TN_REQUIRE(M < (int64_t(1) << 31) && (M + 1) <= INT64_MAX / L / 8, TN_EUNSUPP,
             "tn_lanczos: max_iter too large");
  TN_REQUIRE(!cplx || ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(psi)) & 15) == 0,
             TN_EINVAL, "tn_lanczos_c128: A and psi must be 16-byte aligned");
  hipStream_t st = tn::stream_of(stream);
  double* Vm = nullptr;
  double* dsc = nullptr;   // [2] start norm; [3, 3 + M) z; [3 + M, 3 + 3M) alpha, beta per step
  int rc = tn::pool_alloc((void**)&Vm, sizeof(double) * (size_t)(M + 1) * (size_t)L);
  if (rc != TN_OK) return rc;
  rc = tn::pool_alloc((void**)&dsc, sizeof(double) * (size_t)(3 * M + 3));
  if (rc != TN_OK) {
    tn::pool_free(Vm);
    return rc;
  }
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(st);   // queued work may still use the buffers
    tn::pool_free(Vm);
    tn::pool_free(dsc);
    return code;
  };
  const unsigned vb = (unsigned)std::min<int64_t>((L + 255) / 256, 4096);
  // v0 = psi / ||psi||  (uni10_lapack_cpu.cpp:437-438)
  hipLaunchKernelGGL(lz_norm, dim3(1), dim3(kStepThreads), 0, st, psi, L, dsc + 2);
  hipLaunchKernelGGL(lz_scale_copy, dim3(vb), dim3(256), 0, st, psi, dsc + 2, L, Vm);
  if (hipGetLastError() != hipSuccess) return fail((tn::set_error("lanczos: launch failed"), TN_EHIP));

  // rows of A and the Lanczos vectors (row stride n) are 16-byte aligned
  const bool vec = (n % 2 == 0) && ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(Vm)) & 15) == 0;
  const bool stream_a = (size_t)n * (size_t)n * sizeof(double) >= kStreamBytes;
  std::vector<double> As, Bs;
  As.reserve(M);
  Bs.reserve(M);
  // Steps are launched kBatch at a time, each step's alpha and beta kept on
  // the device (dsc[3 + M + 2i ...]), and read back together: the host then
  // applies the stopping rule step by step and ignores steps computed past
  // the stopping point (their vectors are never combined).
  constexpr int64_t kBatch = 8;
  double* d_ab = dsc + 3 + M;
  std::vector<double> hab((size_t)(2 * M));
  int64_t done = 0;
  double beta = 1.0, e_diff = 1.0, e0_old = 0.0;
  bool conv = false;
  int64_t it = 0;
  while ((((e_diff > err_tol) && it < M) || it < kMinIter) && beta > kBetaErr) {
    if (it == done) {
      const int64_t k = std::min<int64_t>(kBatch, M - done);
      for (int64_t i = done; i < done + k; ++i) {
        const double* v = Vm + i * L;
        const double* vprev = (i == 0) ? v : Vm + (i - 1) * L;
        const double* bprev = (i == 0) ? nullptr : d_ab + 2 * (i - 1) + 1;
        double* w = Vm + (i + 1) * L;
        const dim3 grid((unsigned)((n + kMvRows - 1) / kMvRows));
        if (cplx)
          hipLaunchKernelGGL(lzc_matvec, grid, dim3(kMvThreads), 0, st, A, n, v, vprev, bprev, w);
        else if (vec && stream_a)
          hipLaunchKernelGGL((lz_matvec<true, true>), grid, dim3(kMvThreads), 0, st, A, n, v, vprev, bprev, w);
        else if (vec)
          hipLaunchKernelGGL((lz_matvec<true, false>), grid, dim3(kMvThreads), 0, st, A, n, v, vprev, bprev, w);
        else
          hipLaunchKernelGGL((lz_matvec<false, false>), grid, dim3(kMvThreads), 0, st, A, n, v, vprev, bprev, w);
        if (L <= 4 * kStepThreads)
          hipLaunchKernelGGL(lz_step_reg<4>, dim3(1), dim3(kStepThreads), 0, st, v, w, L, d_ab + 2 * i);
        else if (L <= 8 * kStepThreads)
          hipLaunchKernelGGL(lz_step_reg<8>, dim3(1), dim3(kStepThreads), 0, st, v, w, L, d_ab + 2 * i);
        else if (L <= 16 * kStepThreads)
          hipLaunchKernelGGL(lz_step_reg<16>, dim3(1), dim3(kStepThreads), 0, st, v, w, L, d_ab + 2 * i);
        else
          hipLaunchKernelGGL(lz_step, dim3(1), dim3(kStepThreads), 0, st, v, w, L, d_ab + 2 * i);
      }
      if (hipGetLastError() != hipSuccess) return fail((tn::set_error("lanczos: launch failed"), TN_EHIP));
      rc = tn_memcpy_d2h(hab.data() + 2 * done, d_ab + 2 * done, sizeof(double) * 2 * k, st);
      if (rc != TN_OK) return fail(rc);
      done += k;
    }
    const double alpha = hab[2 * it];
    beta = hab[2 * it + 1];
    // an overflow or NaN is no Krylov closure: a NaN beta fails the closure
    // test below and the Sturm count skips NaN entries, so it is caught here
    if (!std::isfinite(alpha) || !std::isfinite(beta)) {
      tn::set_error("lanczos: step %lld gave a non-finite alpha (%g) or beta (%g)", (long long)it, alpha, beta);
      return fail(TN_ENOCONV);
    }
    As.push_back(alpha);
    if (beta > kBetaErr) {
      Bs.push_back(it < M - 1 ? beta : 0.0);
    } else {
      Bs.push_back(0.0);
      conv = true;
    }
    ++it;
    if (it > 1) {
      const double d0 = tridiag_min_eig(As, Bs, (int)it);
      const double base = std::fabs(d0) > 1 ? std::fabs(d0) : 1;
      e_diff = std::fabs(d0 - e0_old) / base;
      e0_old = d0;
      if (e_diff <= err_tol) conv = true;
    }
  }
  if (it > 1) {
    const double lam = tridiag_min_eig(As, Bs, (int)it);
    const std::vector<double> z = tridiag_eigvec(As, Bs, (int)it, lam);
    rc = tn_memcpy_h2d(dsc + 3, z.data(), sizeof(double) * (size_t)it, st);
    if (rc != TN_OK) return fail(rc);
    hipLaunchKernelGGL(lz_combine, dim3(vb), dim3(256), 0, st, Vm, dsc + 3, it, L, psi);
    if (hipGetLastError() != hipSuccess) return fail((tn::set_error("lanczos: launch failed"), TN_EHIP));
    *iters = it;
    *E0 = lam;
  } else {
    // the Krylov space closed after one step: the reference reports 0 and
    // leaves psi as it was (uni10_lapack_cpu.cpp:501-504)
    *iters = 1;
    *E0 = 0.0;
  }
  *converged = conv ? 1 : 0;
  // the pool may hand the buffers to the next call: finish this one first
  rc = tn_stream_sync(stream);
  tn::pool_free(Vm);
  tn::pool_free(dsc);
  return rc;
}

int tn_lanczos_f64(const double* A, int64_t n, double* psi, int64_t* iters, double err_tol, double* E0,
                   int* converged, void* stream) {
  TN_INIT_OR_RETURN();
  return lanczos_impl(A, n, false, psi, iters, err_tol, E0, converged, stream);
}

int tn_lanczos_c128(const double* A, int64_t n, double* psi, int64_t* iters, double err_tol, double* E0,
                    int* converged, void* stream) {
  TN_INIT_OR_RETURN();
  return lanczos_impl(A, n, true, psi, iters, err_tol, E0, converged, stream);
}
