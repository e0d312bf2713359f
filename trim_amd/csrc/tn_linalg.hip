// FP64 dense decompositions and small vector kernels.
//
//  * SVD: one-sided (Hestenes) Jacobi.  The reference calls dgesvd('S','S')
//    on the transposed view (uni10_lapack_cpu.cpp:289-313) or cuSOLVER gesvd
//    on device (uni10_lapack_gpu.cu:418-500).  Jacobi is used here because it
//    is parallel within a round (disjoint column pairs), gives singular values
//    to high relative accuracy, and is deterministic (fixed pairing order,
//    fixed reduction trees).  Matrices that fit one workgroup's LDS are
//    solved by one workgroup each, the working columns (W = A or A^T,
//    column-major) and the accumulated rotations V in LDS; larger ones by
//    the block Jacobi of tn_svd_block.hip (many workgroups per matrix).
//  * eigh: parallel two-sided Jacobi (replaces dsyev, uni10_lapack_cpu.cpp:131-152,
//    which the reference's GPU build also ran on the host,
//    uni10_lapack_gpu.cu:380-416).
//  * norm / scale / diagonal row-column scaling / truncation: single-pass
//    kernels with fixed-order reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "tn_internal.h"
#include "tn_svd_block.h"

namespace {

constexpr int kThreads = 256;
thread_local int g_last_sweeps = 0;   // sweeps of the last SVD batch on this thread
constexpr size_t kLdsBytes = 152 * 1024;   // 160 KiB per CU minus the kernel's static arrays

struct SvdProb {
  int64_t m, n;              // A is m x n row-major
  int64_t a_off, u_off, s_off, vt_off;
  int64_t v0_off;            // warm start: offset of an n x n V0^T (m >= n), -1 = none
  int64_t max_sweeps;        // sweeps before the problem is reported as not converged
  const int* mask;           // with mask_idx >= 0: skipped while mask[mask_idx] == 0
  int64_t mask_idx;
};

// Sum of a, b, c over each aligned group of TP lanes (TP a power of two),
// the three butterflies interleaved so their shuffles overlap.
template <int TP>
__device__ inline void group_sum3(double& a, double& b, double& c) {
#pragma unroll
  for (int o = TP >> 1; o > 0; o >>= 1) {
    const double x = __shfl_xor(a, o, 64), y = __shfl_xor(b, o, 64), z = __shfl_xor(c, o, 64);
    a += x;
    b += y;
    c += z;
  }
}

// Round-robin pairing: np even; round r in [0, np-1); pair k in [0, np/2).
__device__ inline void rr_pair(int r, int k, int np, int& p, int& q) {
  const int L = np - 1;
  if (k == 0) {
    p = r;
    q = L;
  } else {
    p = (r + k) % L;
    q = (r - k + L) % L;
  }
  if (p > q) {
    int t = p;
    p = q;
    q = t;
  }
}

// One workgroup per matrix, W and V in LDS.  TP lanes work on one column
// pair (few lanes per pair: short reductions, which bound the latency of a
// round).
template <int TP>
__global__ void __launch_bounds__(kThreads) svd_jacobi_kernel(const SvdProb* __restrict__ probs,
                                                              const double* __restrict__ A,
                                                              const double* __restrict__ V0T,
                                                              double* __restrict__ U,
                                                              double* __restrict__ S,
                                                              double* __restrict__ VT,
                                                              int* __restrict__ status) {
  extern __shared__ double lds[];
  __shared__ double sig[512];
  __shared__ int rank_of[512];
  __shared__ double red[kThreads];
  const SvdProb pr = probs[blockIdx.x];
  if (pr.mask != nullptr && pr.mask_idx >= 0 && pr.mask[pr.mask_idx] == 0) return;
  const int64_t M = pr.m, N = pr.n;
  const bool tr = (M < N);           // work on A^T
  const int m = (int)(tr ? N : M);   // column length
  const int n = (int)(tr ? M : N);   // number of columns
  if (m == 0 || n == 0) return;
  // Column strides are padded by one double so that the columns handled by
  // different pair groups in a round start on different LDS banks.
  const int ldw = m + 1, ldv = n + 1;
  double* W = lds;
  double* V = W + (size_t)ldw * n;
  const int tid = threadIdx.x;
  const double* Ab = A + pr.a_off;

  // W column-major: W[j*ldw + i]
  // warm start (m >= n only): V = V0, W = A V0; otherwise V = I, W = A
  const bool warm = pr.v0_off >= 0 && !tr;
  for (int64_t e = tid; e < (int64_t)n * n; e += kThreads) {
    int j = (int)(e / n), i = (int)(e % n);
    V[(size_t)j * ldv + i] = warm ? V0T[pr.v0_off + (int64_t)j * n + i] : ((i == j) ? 1.0 : 0.0);
  }
  if (warm) __syncthreads();
  for (int64_t e = tid; e < (int64_t)m * n; e += kThreads) {
    int j = (int)(e / m), i = (int)(e % m);
    if (warm) {
      double x = 0.0;
      for (int k = 0; k < n; ++k) x += Ab[(int64_t)i * N + k] * V[(size_t)j * ldv + k];
      W[(size_t)j * ldw + i] = x;
    } else {
      W[(size_t)j * ldw + i] = tr ? Ab[(int64_t)j * N + i] : Ab[(int64_t)i * N + j];
    }
  }
  // ||A||_F^2 (fixed-order tree), invariant under the rotations
  {
    double s = 0;
    for (int64_t e = tid; e < (int64_t)m * n; e += kThreads) {
      int j = (int)(e / m), i = (int)(e % m);
      double x = W[(size_t)j * ldw + i];
      s += x * x;
    }
    red[tid] = s;
  }
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  const double normF2 = red[0];
  __syncthreads();

  const int np = (n % 2) ? n + 1 : n;
  const int npairs = np / 2;
  constexpr int tp = TP;
  constexpr int groups = kThreads / TP;
  const int g = tid / tp, t = tid % tp;
  // Rotate a pair while |w_p.w_q| > tol |w_p||w_q|.  The computed dot product
  // of orthogonal columns carries rounding noise of order sqrt(m) eps, so the
  // threshold sits above it (10 m eps) or sweeps never stop; singular values
  // are then exact to O(tol^2) relative and vectors to O(tol).
  const double tol = 10.0 * DBL_EPSILON * (double)m;
  // Columns at rounding level (|w| ~ eps ||A||_F, e.g. the null space of a
  // rank-deficient theta) have dot products that are pure noise: a pair is
  // left alone when |w_p.w_q| is below m eps^2 ||A||_F^2.
  const double abs_tol = (double)m * DBL_EPSILON * DBL_EPSILON * normF2;

  int sweep = 0;
  const int max_sweeps = (int)pr.max_sweeps;
  for (; sweep < max_sweeps; ++sweep) {
    int rotated = 0;
    for (int r = 0; r < np - 1; ++r) {
      for (int k0 = 0; k0 < npairs; k0 += groups) {
        const int k = k0 + g;
        // every lane of a group runs the same path; groups past the end idle
        int p = 0, q = 0;
        bool live = (k < npairs);
        if (live) {
          rr_pair(r, k, np, p, q);
          live = (q < n);
        }
        double a = 0, b = 0, c = 0;
        if (live) {
          const double* wp = W + (size_t)p * ldw;
          const double* wq = W + (size_t)q * ldw;
#pragma unroll 4
          for (int i = t; i < m; i += tp) {
            double x = wp[i], y = wq[i];
            a += x * x;
            b += y * y;
            c += x * y;
          }
        }
        group_sum3<TP>(a, b, c);
        if (live && fabs(c) > tol * sqrt(a * b) && fabs(c) > abs_tol) {
          rotated = 1;
          const double zeta = (b - a) / (2.0 * c);
          const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double cs = 1.0 / sqrt(1.0 + tt * tt);
          const double sn = cs * tt;
          double* wp = W + (size_t)p * ldw;
          double* wq = W + (size_t)q * ldw;
#pragma unroll 4
          for (int i = t; i < m; i += tp) {
            double x = wp[i], y = wq[i];
            wp[i] = cs * x - sn * y;
            wq[i] = sn * x + cs * y;
          }
          double* vp = V + (size_t)p * ldv;
          double* vq = V + (size_t)q * ldv;
#pragma unroll 4
          for (int i = t; i < n; i += tp) {
            double x = vp[i], y = vq[i];
            vp[i] = cs * x - sn * y;
            vq[i] = sn * x + cs * y;
          }
        }
      }
      __syncthreads();
    }
    if (!__syncthreads_or(rotated)) break;
  }
  if (tid == 0) {
    if (sweep >= max_sweeps) atomicAdd(status, 1);
    atomicMax(status + 1, sweep + 1);   // sweeps run (diagnostic)
  }

  // singular values = column norms (fixed order per column)
  for (int j = tid; j < n; j += kThreads) {
    const double* w = W + (size_t)j * ldw;
    double s = 0;
    for (int i = 0; i < m; ++i) s += w[i] * w[i];
    sig[j] = sqrt(s);
  }
  __syncthreads();
  for (int j = tid; j < n; j += kThreads) {
    const double sj = sig[j];
    int rk = 0;
    for (int i = 0; i < n; ++i) {
      const double si = sig[i];
      rk += (si > sj) || (si == sj && i < j);
    }
    rank_of[j] = rk;
  }
  __syncthreads();
  double* Ub = U + pr.u_off;
  double* Sb = S + pr.s_off;
  double* VTb = VT + pr.vt_off;
  for (int j = tid; j < n; j += kThreads) Sb[rank_of[j]] = sig[j];
  // normalised W columns -> U (A = U S V^T) or VT (A^T = W S V^T)
  for (int64_t e = tid; e < (int64_t)m * n; e += kThreads) {
    int j = (int)(e / m), i = (int)(e % m);
    double sj = sig[j];
    double v = (sj > 0.0) ? W[(size_t)j * ldw + i] / sj : 0.0;
    int rk = rank_of[j];
    if (!tr) {
      if (pr.u_off >= 0) Ub[(int64_t)i * n + rk] = v;   // U is M x n (n = N)
    } else {
      VTb[(int64_t)rk * N + i] = v;      // VT is n x N (n = M)
    }
  }
  for (int64_t e = tid; e < (int64_t)n * n; e += kThreads) {
    int j = (int)(e / n), i = (int)(e % n);
    int rk = rank_of[j];
    const double vv = V[(size_t)j * ldv + i];
    if (!tr)
      VTb[(int64_t)rk * n + i] = vv;     // VT is n x n
    else if (pr.u_off >= 0)
      Ub[(int64_t)i * n + rk] = vv;      // U is M x M
  }
}

// ---------------------------------------------------------------------------
// Parallel two-sided Jacobi for a symmetric n x n matrix (n <= 90: both
// matrices in LDS), one workgroup per matrix: problem blockIdx.x of probs,
// or `one` when probs is NULL.
struct EighProb {
  int64_t n, a_off, e_off, v_off;
};
__global__ void __launch_bounds__(kThreads) eigh_jacobi_kernel(const EighProb* __restrict__ probs, EighProb one,
                                                               const double* __restrict__ A0,
                                                               double* __restrict__ eig0,
                                                               double* __restrict__ vec0,
                                                               int* __restrict__ status) {
  const EighProb pr = probs ? probs[blockIdx.x] : one;
  const int n = (int)pr.n;
  if (n <= 0) return;
  const double* __restrict__ A = A0 + pr.a_off;
  double* __restrict__ eig = eig0 + pr.e_off;
  double* __restrict__ vec_rows = vec0 + pr.v_off;
  extern __shared__ double lds[];
  double* a = lds;              // n x n row-major (symmetric)
  double* v = lds + n * n;      // n x n, columns = eigenvectors
  __shared__ double cs_sh[128], sn_sh[128];
  __shared__ int p_sh[128], q_sh[128];
  __shared__ double red[kThreads];
  __shared__ double ev[256];
  __shared__ int rk[256];
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += kThreads) {
    a[e] = A[e];
    v[e] = ((e / n) == (e % n)) ? 1.0 : 0.0;
  }
  __syncthreads();
  // symmetrise from the upper triangle (dsyev 'U' reads the upper part)
  for (int e = tid; e < n * n; e += kThreads) {
    int i = e / n, j = e % n;
    if (i > j) a[e] = a[j * n + i];
  }
  __syncthreads();
  double s = 0;
  for (int e = tid; e < n * n; e += kThreads) s += a[e] * a[e];
  red[tid] = s;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  const double thresh = 0.5 * DBL_EPSILON * sqrt(red[0]);
  __syncthreads();
  const int np = (n % 2) ? n + 1 : n;
  const int npairs = np / 2;
  int sweep = 0;
  for (; sweep < 100 && n > 1; ++sweep) {
    int rot = 0;
    for (int r = 0; r < np - 1; ++r) {
      for (int k = tid; k < npairs; k += kThreads) {
        int p, q;
        rr_pair(r, k, np, p, q);
        double c = 1.0, sn = 0.0;
        if (q < n) {
          double apq = a[p * n + q];
          if (fabs(apq) > thresh) {
            double zeta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
            double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = c * t;
            rot = 1;
          }
        }
        cs_sh[k] = c;
        sn_sh[k] = sn;
        p_sh[k] = p;
        q_sh[k] = (q < n) ? q : -1;
      }
      __syncthreads();
      // A <- A J (columns p, q)
      for (int e = tid; e < npairs * n; e += kThreads) {
        int k = e / n, i = e % n;
        int p = p_sh[k], q = q_sh[k];
        if (q < 0) continue;
        double c = cs_sh[k], sn = sn_sh[k];
        double x = a[i * n + p], y = a[i * n + q];
        a[i * n + p] = c * x - sn * y;
        a[i * n + q] = sn * x + c * y;
        double vx = v[i * n + p], vy = v[i * n + q];
        v[i * n + p] = c * vx - sn * vy;
        v[i * n + q] = sn * vx + c * vy;
      }
      __syncthreads();
      // A <- J^T A (rows p, q)
      for (int e = tid; e < npairs * n; e += kThreads) {
        int k = e / n, j = e % n;
        int p = p_sh[k], q = q_sh[k];
        if (q < 0) continue;
        double c = cs_sh[k], sn = sn_sh[k];
        double x = a[p * n + j], y = a[q * n + j];
        a[p * n + j] = c * x - sn * y;
        a[q * n + j] = sn * x + c * y;
      }
      __syncthreads();
    }
    if (!__syncthreads_or(rot)) break;
  }
  if (tid == 0 && sweep >= 100) atomicAdd(status, 1);
  for (int j = tid; j < n; j += kThreads) ev[j] = a[j * n + j];
  __syncthreads();
  for (int j = tid; j < n; j += kThreads) {
    int r = 0;
    for (int i = 0; i < n; ++i) r += (ev[i] < ev[j]) || (ev[i] == ev[j] && i < j);
    rk[j] = r;
  }
  __syncthreads();
  for (int j = tid; j < n; j += kThreads) eig[rk[j]] = ev[j];
  for (int e = tid; e < n * n; e += kThreads) {
    int i = e / n, j = e % n;   // v[i][j]: component i of eigenvector j
    vec_rows[(int64_t)rk[j] * n + i] = v[e];
  }
}

// ---------------------------------------------------------------------------
template <int T>
__device__ inline double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int o = T / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(1024) norm2_kernel(const double* __restrict__ x, int64_t n,
                                                     double* __restrict__ out) {
  __shared__ double red[1024];
  double s = 0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) s += x[i] * x[i];
  double t = block_sum<1024>(s, red);
  if (threadIdx.x == 0) *out = sqrt(t);
}

// x[0], x[inc], ... summed in a fixed order (per-thread strided partials,
// then the block tree), like the norm.
__global__ void __launch_bounds__(1024) sum_kernel(const double* __restrict__ x, int64_t n, int64_t inc,
                                                   double* __restrict__ out) {
  __shared__ double red[1024];
  double s = 0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) s += x[i * inc];
  double t = block_sum<1024>(s, red);
  if (threadIdx.x == 0) *out = t;
}

// SplitMix64 -> uniform[-1, 1): dst[i] = 2 (mix(seed + (i+1) golden) >> 11) 2^-53 - 1
// (every step exact, so the values equal a host evaluation bit for bit)
__global__ void fill_uniform_kernel(double* __restrict__ dst, int64_t n, uint64_t seed) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    dst[i] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
}

__global__ void axpy_kernel(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

__global__ void copy_strided_kernel(double* __restrict__ dst, int64_t dinc, const double* __restrict__ src,
                                    int64_t sinc, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i * dinc] = src[i * sinc];
}

__global__ void scale_kernel(double* __restrict__ x, int64_t n, double a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}

__global__ void scale_inv_dev_kernel(double* __restrict__ x, int64_t n,
                                     const double* __restrict__ d) {
  const double a = 1.0 / d[0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}

__global__ void diag_row_mul_kernel(double* __restrict__ mat, const double* __restrict__ diag,
                                    int64_t M, int64_t N) {
  const int64_t total = M * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    mat[e] *= diag[e / N];
}

__global__ void diag_col_mul_kernel(double* __restrict__ mat, const double* __restrict__ diag,
                                    int64_t M, int64_t N) {
  const int64_t total = M * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    mat[e] *= diag[e % N];
}

__global__ void __launch_bounds__(1024) truncate_kernel(const double* __restrict__ S, int64_t k,
                                                        int64_t dim, double* __restrict__ lam,
                                                        double* __restrict__ norm_out,
                                                        double* __restrict__ err_out) {
  __shared__ double red[1024];
  double a = 0, b = 0;
  for (int64_t i = threadIdx.x; i < k; i += 1024) {
    double s = S[i];
    if (i < dim)
      a += s * s;
    else
      b += s * s;
  }
  const double keep = block_sum<1024>(a, red);
  const double drop = block_sum<1024>(b, red);
  const double nrm = sqrt(keep);
  for (int64_t i = threadIdx.x; i < dim; i += 1024) lam[i] = S[i] / nrm;
  if (threadIdx.x == 0) {
    *norm_out = nrm;
    if (err_out) *err_out = drop;
  }
}

unsigned grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

// ---------------------------------------------------------------------------
// One-sided Jacobi for small matrices (n <= 64 columns of length m <= 8*EMX):
// 32 groups of 8 lanes, one column pair per group and round.  A lane holds a
// contiguous chunk of EMX elements of each of its two columns (and EN = 8 of
// the two rotation columns) in registers for the round; columns live in LDS
// between rounds, each lane chunk padded by 16 bytes so the 8 lanes of a
// group hit disjoint banks.  The squared column norms are kept in LDS and
// updated with each rotation (a' = a - t c, b' = b + t c), so a round needs
// one dot product per pair; they are recomputed exactly at every sweep.
typedef double d4s __attribute__((ext_vector_type(4)));
constexpr int kSmallN = 64;      // columns
constexpr int kSmallTP = 8;      // lanes per pair
constexpr int kSmallEN = 8;      // V chunk per lane (n <= 64)

// DPP moves within a row of 16 lanes: quad_perm [1,0,3,2] (lane ^ 1),
// quad_perm [2,3,0,1] (lane ^ 2) and row_half_mirror (lane i <-> 7 - i in
// each half row), so an 8-lane group sum takes three register moves instead
// of LDS permutes; every lane ends with the same value (each step adds a
// commuted pair).
template <int CTRL>
__device__ inline double dpp_mov(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double group8_sum(double x) {
  x += dpp_mov<0xB1>(x);    // quad_perm [1,0,3,2]
  x += dpp_mov<0x4E>(x);    // quad_perm [2,3,0,1]
  x += dpp_mov<0x141>(x);   // row_half_mirror
  return x;
}

// 1/x and 1/sqrt(x) for x > 0: hardware estimate refined by Newton steps to
// full double precision (each step squares the relative error).
__device__ inline double fast_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}
__device__ inline double fast_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  y = y * fma(-0.5 * x * y, y, 1.5);
  return y;
}

// rr_pair without integer division: p = (r + k) mod L, q = (r - k) mod L for
// 0 <= r, k < L (k > 0), L = np - 1.
__device__ inline void rr_pair_fast(int r, int k, int np, int& p, int& q) {
  const int L = np - 1;
  if (k == 0) {
    p = r;
    q = L;
  } else {
    p = r + k;
    p -= (p >= L) ? L : 0;
    q = r - k;
    q += (q < 0) ? L : 0;
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

// Lane moves of the register-resident sweeps, where lane 8 t + j of a wave
// holds row chunk t of group j's columns.  group_xor<D>: the value held by
// group j ^ D (same t), by DPP inside rows of 16 lanes: quad_perm for xor 1,
// 2 and 3, row_half_mirror (lane i <-> 7 - i) for xor 7, and the two
// composed for 4, 5 and 6.
template <int D>
__device__ inline double group_xor(double x) {
  static_assert(D >= 1 && D <= 7, "group offset");
  if constexpr (D == 1) return dpp_mov<0xB1>(x);
  else if constexpr (D == 2) return dpp_mov<0x4E>(x);
  else if constexpr (D == 3) return dpp_mov<0x1B>(x);
  else if constexpr (D == 7) return dpp_mov<0x141>(x);
  else return group_xor<D ^ 7>(dpp_mov<0x141>(x));
}
template <int D, int E>
__device__ inline void group_xor(double (&x)[E]) {
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = group_xor<D>(x[e]);
}

// x + (x of the lane 16 (S = 16) or 32 (S = 32) away): the permlane swap of
// a value with itself leaves, in every lane, the lower row's (half's) copy
// in the first result and the upper one's in the second, so the sum has one
// operand order everywhere.
template <int S>
__device__ inline double swap_add(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
  unsigned l0, l1, h0, h1;
  if constexpr (S == 16) {
    const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    l0 = l[0], l1 = l[1], h0 = h[0], h1 = h[1];
  } else {
    const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    l0 = l[0], l1 = l[1], h0 = h[0], h1 = h[1];
  }
  const double u = __longlong_as_double((long long)(((unsigned long long)h0 << 32) | l0));
  const double v = __longlong_as_double((long long)(((unsigned long long)h1 << 32) | l1));
  return u + v;
}

// Sum over the eight lanes j, j + 8, ..., j + 56 of a group; every lane of
// the group ends with the same bits (each step adds a commuted pair or one
// operand order everywhere).
__device__ inline double group_sum_strided(double x) {
  x += dpp_mov<0x128>(x);   // row_ror:8, lane ^ 8 inside a row of 16
  x = swap_add<16>(x);
  return swap_add<32>(x);
}

// The Jacobi rotation of a pair with squared norms np, nq and dot product
// c: the tangent from the raw reciprocal and square-root estimates (its
// error only moves the angle), cs = 1 / sqrt(1 + t^2) refined, sn = cs t.
struct JacobiRot {
  double cs, sn, tt;
};
__device__ inline JacobiRot jacobi_rot(double np, double nq, double c) {
  const double zeta = (nq - np) * (0.5 * __builtin_amdgcn_rcp(c));
  const double z2 = fma(zeta, zeta, 1.0);
  const double tt = copysign(1.0, zeta) * __builtin_amdgcn_rcp(fabs(zeta) + z2 * __builtin_amdgcn_rsq(z2));
  const double cs = fast_rsqrt(fma(tt, tt, 1.0));
  return {cs, cs * tt, tt};
}

// When only S and VT are wanted (no U) and M >= N, the kernel first
// factorises W = Q R by Householder reflections in LDS and runs the sweeps on
// X = R^T, whose column Gram matrix is close to diagonal (Drmac and Veselic,
// "New fast and accurate Jacobi SVD algorithm", SIAM J. Matrix Anal. Appl.
// 29 (2008)): on TEBD thetas this halves the sweeps.  From X = U_X S V_X^T,
// R = V_X S U_X^T and A V0 P = Q R give A = (Q V_X) S (V0 P U_X)^T, so the
// right singular vectors are V0 P U_X with U_X the normalised columns of X;
// no rotation is accumulated.  This needs every column of X nonzero, so the
// sweeps run without the absolute floor (pairs are orthogonal relative to
// their own norms, which keeps U_X orthonormal down to the noise columns of
// a rank-deficient A); a column norm at the underflow level (an exactly
// rank-deficient A) sends the problem back through the plain sweeps.
template <int EMX>
__global__ void __launch_bounds__(kThreads) svd_jacobi_small(const SvdProb* __restrict__ probs,
                                                             const double* __restrict__ A,
                                                             const double* __restrict__ V0T,
                                                             double* __restrict__ U,
                                                             double* __restrict__ S,
                                                             double* __restrict__ VT,
                                                             int* __restrict__ status) {
  constexpr int CHW = EMX + 2;                   // W chunk stride (doubles)
  // W column stride: 8 CHW doubles, 16 mod 32, so that two neighbouring
  // columns sit half the LDS banks apart (with the columns in sweep order,
  // the groups of a round read neighbouring columns)
  constexpr int CW = kSmallTP * CHW;
  static_assert(CW % 32 == 16, "column stride must be 16 mod 32 doubles");
  constexpr int CHV = kSmallEN + 2;
  constexpr int CV = kSmallTP * CHV + 2;
  constexpr int kRows = kSmallTP * EMX;          // W rows held (m <= kRows)
  __shared__ __attribute__((aligned(16))) double W[kSmallN * CW];
  __shared__ __attribute__((aligned(16))) double V[kSmallN * CV];
  __shared__ __attribute__((aligned(16))) double nrm[kSmallN];
  __shared__ int rank_of[kThreads];
  __shared__ int col_at[kSmallN];
  __shared__ int qr_col[kSmallN];
  __shared__ double hh_xx[2];
  __shared__ __attribute__((aligned(16))) double hh_v[2 * kRows];
  const SvdProb pr = probs[blockIdx.x];
  if (pr.mask != nullptr && pr.mask_idx >= 0 && pr.mask[pr.mask_idx] == 0) return;
  const int64_t M = pr.m, N = pr.n;
  const bool tr = (M < N);
  const int m = (int)(tr ? N : M);
  const int n = (int)(tr ? M : N);
  if (m == 0 || n == 0) return;
  const int tid = threadIdx.x;
  const int g = tid / kSmallTP, t = tid % kSmallTP;
  const double* Ab = A + pr.a_off;
  const bool want_u = pr.u_off >= 0;
  auto widx = [](int c, int i) { return c * CW + (i / EMX) * CHW + (i % EMX); };
  auto vidx = [](int c, int i) { return c * CV + (i / kSmallEN) * CHV + (i % kSmallEN); };

  // V = I (or V0 for a warm start, m >= n only); W = A or A^T (or A V0),
  // rows m..8*EMX zero
  // (all global loads of a thread are issued before any is used)
  const bool warm = pr.v0_off >= 0 && !tr;
  // W = A (A^T for M < N), read in the source's row order so that
  // neighbouring lanes read neighbouring addresses; for a warm start W = A V0
  // is then formed from LDS, each thread a block of EMX / 2 rows by 4 columns.
  // The first call also fills V, its loads issued with those of A.
  auto load_w = [&](bool with_v) {
    constexpr int kPerV = kSmallN * kSmallTP * kSmallEN / kThreads;
    constexpr int kPerA = kSmallN * kRows / kThreads;
    double vv[kPerV], xa[kPerA];
    if (with_v) {
#pragma unroll
      for (int u = 0; u < kPerV; ++u) {
        const int e = tid + u * kThreads, c = e / (kSmallTP * kSmallEN), i = e % (kSmallTP * kSmallEN);
        vv[u] = (c == i && c < n) ? 1.0 : 0.0;
        if (warm) vv[u] = (c < n && i < n) ? V0T[pr.v0_off + (int64_t)c * n + i] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kPerA; ++u) {
      const int e = tid + u * kThreads;
      const int i = tr ? e % kRows : e / kSmallN;
      const int c = tr ? e / kRows : e % kSmallN;
      xa[u] = (c < n && i < m) ? (tr ? Ab[(int64_t)c * N + i] : Ab[(int64_t)i * N + c]) : 0.0;
    }
    if (with_v) {
#pragma unroll
      for (int u = 0; u < kPerV; ++u) {
        const int e = tid + u * kThreads;
        V[vidx(e / (kSmallTP * kSmallEN), e % (kSmallTP * kSmallEN))] = vv[u];
      }
    }
#pragma unroll
    for (int u = 0; u < kPerA; ++u) {
      const int e = tid + u * kThreads;
      W[widx(tr ? e / kRows : e % kSmallN, tr ? e % kRows : e / kSmallN)] = xa[u];
    }
    __syncthreads();
    if (warm) {
      constexpr int RB = EMX / 2;
      const int r0 = (tid / 16) * RB, c0 = (tid % 16) * 4;
      double acc[RB][4];
#pragma unroll
      for (int u = 0; u < RB; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[u][j] = 0.0;
#pragma unroll 8
      for (int k = 0; k < n; ++k) {
        double a[RB], v[4];
#pragma unroll
        for (int u = 0; u < RB; ++u) a[u] = W[widx(k, r0 + u)];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = V[vidx(c0 + j, k)];
#pragma unroll
        for (int u = 0; u < RB; ++u)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[u][j] = fma(a[u], v[j], acc[u][j]);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < RB; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) W[widx(c0 + j, r0 + u)] = acc[u][j];
      __syncthreads();
    }
  };
  load_w(true);

  // ||A||_F^2 from the exact column norms (group sums, then a fixed-order
  // sum over the columns by one thread)
  auto exact_norms = [&]() {
    for (int c = g; c < n; c += kThreads / kSmallTP) {
      const double* w = W + c * CW + t * CHW;
      double s = 0.0;
#pragma unroll
      for (int e = 0; e < EMX; ++e) s += w[e] * w[e];
      s = group8_sum(s);
      if (t == 0) nrm[c] = s;
    }
    __syncthreads();
  };
  exact_norms();
  double normF2 = 0.0;
  for (int c = 0; c < n; ++c) normF2 += nrm[c];

  const int np = (n % 2) ? n + 1 : n;
  const int npairs = np / 2;

  // Each sweep starts from the columns sorted by decreasing norm (de Rijk):
  // position k of the round-robin schedule holds column col_at[k]; sorting
  // moves no data and cuts the sweeps needed for graded spectra.
  auto sort_positions = [&]() {
    for (int j = tid; j < np; j += kThreads) {
      if (j >= n) {
        col_at[j] = j;   // the padding position of an odd n
        continue;
      }
      const double sj = nrm[j];
      int rk = 0;
      for (int i = 0; i < n; ++i) {
        const double si = nrm[i];
        rk += (si > sj) || (si == sj && i < j);
      }
      col_at[rk] = j;
    }
    __syncthreads();
  };

  // Moves column col_at[j] (and its norm) to position j, so that position =
  // column for the rest of the sweep; ends with col_at = identity.
  auto permute_sorted = [&]() {
    constexpr int kPerW = kSmallN * kRows / kThreads;
    double st[kPerW];
#pragma unroll
    for (int u = 0; u < kPerW; ++u) {
      const int e = tid + u * kThreads, c = e / kRows, i = e % kRows;
      st[u] = (c < n) ? W[widx(col_at[c], i)] : 0.0;
    }
    const double nv = (tid < n) ? nrm[col_at[tid]] : 0.0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPerW; ++u) {
      const int e = tid + u * kThreads, c = e / kRows, i = e % kRows;
      if (c < n) W[widx(c, i)] = st[u];
    }
    if (tid < n) nrm[tid] = nv;
    for (int j = tid; j < np; j += kThreads) col_at[j] = j;
    __syncthreads();
  };

  // W = Q R by Householder reflections, then X = R^T in the first n columns
  // of W.  Cold start: the columns are first moved into decreasing-norm
  // order (a pivot-free stand-in for column pivoting); warm start: A V0 is
  // already ordered by the previous singular values.  qr_col[j] is the
  // column of A V0 at position j.
  auto householder_rt = [&]() {
    if (warm) {
      for (int j = tid; j < n; j += kThreads) qr_col[j] = j;
    } else {
      sort_positions();
      for (int j = tid; j < n; j += kThreads) qr_col[j] = col_at[j];
      permute_sorted();
    }
    __syncthreads();
    // The columns in registers: group G = 8 w + (lane >> 3) of wave w (eight
    // adjacent lanes, so a column sum takes three DPP steps) holds columns G
    // and G + 32, its lane 8 G + t row chunk t of each.  Step j: the owner of
    // column j has published its reflector v = x - alpha e_j (x = rows >= j
    // of the column) and 1 / (v^T v / 2) in LDS; after the barrier every
    // group reflects its columns (a finished column sees a zero coefficient;
    // from step 32 on the first slot holds only finished columns and is
    // skipped), and the owner of column j + 1 publishes the next.
    {
      const int tc = tid & 7;
      const int G = tid >> 3;
      double col[2][EMX];
#pragma unroll
      for (int sl = 0; sl < 2; ++sl) {
        const int k = G + 32 * sl;
        const double* wk = W + (k < n ? k : 0) * CW + tc * CHW;
#pragma unroll
        for (int e = 0; e < EMX; e += 2) {
          const double2 u = *reinterpret_cast<const double2*>(wk + e);
          col[sl][e] = (k < n) ? u.x : 0.0;
          col[sl][e + 1] = (k < n) ? u.y : 0.0;
        }
      }
      // the owner of column jj (slot sl) publishes its reflector and turns
      // the column into column jj of R.  Row jj is element EJ = jj % EMX of
      // lane jj / EMX's chunk: the steps are unrolled by EMX so that EJ is a
      // constant (a runtime element index had the compiler pass the chunk
      // through LDS and set the diagonal through a tree of branches, both on
      // the step's critical path); the column's rows >= jj are cleared by
      // subtracting x (exact) before the diagonal is set.
      auto publish = [&](auto slc, auto ejc, int jj) {
        constexpr int sl = decltype(slc)::value, EJ = decltype(ejc)::value;
        const int tj = jj / EMX;
        double x[EMX];
        double q0 = 0.0, q1 = 0.0;
#pragma unroll
        for (int e = 0; e < EMX; ++e) x[e] = (tc * EMX + e >= jj) ? col[sl][e] : 0.0;
        const double xjl = (tc == tj) ? x[EJ] : 0.0;
#pragma unroll
        for (int e = 0; e < EMX; e += 2) {
          q0 = fma(x[e], x[e], q0);
          q1 = fma(x[e + 1], x[e + 1], q1);
        }
        const double xx = group8_sum(q0 + q1);
        const double xj = group8_sum(xjl);   // one nonzero term
        const bool refl = xx > 0.0;
        const double nrm_x = refl ? xx * fast_rsqrt(refl ? xx : 1.0) : 0.0;
        const double alpha = -copysign(nrm_x, xj);
        const double den = alpha * (alpha - xj);   // v^T v / 2
        const double rden = refl ? fast_rcp(refl ? den : 1.0) : 0.0;
        const bool mine = tc == tj;
        double* vb = hh_v + (jj & 1) * kRows + tc * EMX;
#pragma unroll
        for (int e = 0; e < EMX; ++e) {
          vb[e] = (e == EJ && mine) ? xj - alpha : x[e];
          col[sl][e] = (e == EJ && mine) ? (refl ? alpha : xj) : col[sl][e] - x[e];   // rows >= jj to zero, exactly
        }
        if (tc == 0) hh_xx[jj & 1] = rden;
      };
      if (G == 0) publish(std::integral_constant<int, 0>{}, std::integral_constant<int, 0>{}, 0);
      for (int j0 = 0; j0 < n; j0 += EMX) {
        tn::static_for<EMX>([&](auto uc) {
          constexpr int u = decltype(uc)::value;
          const int j = j0 + u;
          if (j >= n) return;
          __syncthreads();
          double v[EMX];
          const double* vb = hh_v + (j & 1) * kRows + tc * EMX;
#pragma unroll
          for (int e = 0; e < EMX; e += 2) {
            const double2 w2 = *reinterpret_cast<const double2*>(vb + e);
            v[e] = w2.x;
            v[e + 1] = w2.y;
          }
          const double rden = hh_xx[j & 1];
#pragma unroll
          for (int sl = 0; sl < 2; ++sl) {
            if (sl == 0 && j >= 31) continue;   // columns 0..31 are all finished
            const int k = G + 32 * sl;
            double d0 = 0.0, d1 = 0.0;
#pragma unroll
            for (int e = 0; e < EMX; e += 2) {
              d0 = fma(v[e], col[sl][e], d0);
              d1 = fma(v[e + 1], col[sl][e + 1], d1);
            }
            const double coef = (k > j && k < n) ? group8_sum(d0 + d1) * rden : 0.0;
#pragma unroll
            for (int e = 0; e < EMX; ++e) col[sl][e] = fma(-coef, v[e], col[sl][e]);
          }
          const int jn = j + 1;
          constexpr int EJN = (u + 1) % EMX;
          if (jn < n && (jn & 31) == G) {
            if (jn < 32)
              publish(std::integral_constant<int, 0>{}, std::integral_constant<int, EJN>{}, jn);
            else
              publish(std::integral_constant<int, 1>{}, std::integral_constant<int, EJN>{}, jn);
          }
        });
      }
      __syncthreads();
#pragma unroll
      for (int sl = 0; sl < 2; ++sl) {
        const int k = G + 32 * sl;
        if (k < n) {
          double* wk = W + k * CW + tc * CHW;
#pragma unroll
          for (int e = 0; e < EMX; e += 2) *reinterpret_cast<double2*>(wk + e) = double2{col[sl][e], col[sl][e + 1]};
        }
      }
    }
    __syncthreads();
    // X[i][k] = R[k][i] = W[k][i] (k <= i): staged in registers
    constexpr int kPer = (kSmallN * kSmallN + kThreads - 1) / kThreads;
    double xv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = tid + u * kThreads;
      xv[u] = 0.0;
      if (e < n * n) {
        const int i = e / n, k = e % n;
        if (k <= i) xv[u] = W[widx(i, k)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = tid + u * kThreads;
      if (e < n * n) W[widx(e % n, e / n)] = xv[u];
    }
    for (int e = tid; e < n * (kRows - n); e += kThreads) W[widx(e / (kRows - n), n + e % (kRows - n))] = 0.0;
    __syncthreads();
  };

  // a plain sweep run also accumulates the rotations in V
  auto sweeps = [&](bool qr) {
    const int mr = qr ? n : m;
    const double tol = 10.0 * DBL_EPSILON * (double)mr;
    const double abs_tol = qr ? 0.0 : (double)m * DBL_EPSILON * DBL_EPSILON * normF2;
    const int max_sweeps = (int)pr.max_sweeps;
    int sweep = 0;
    for (; sweep < max_sweeps; ++sweep) {
      if (sweep > 0) exact_norms();
      sort_positions();
      if (qr) permute_sorted();   // no V to carry along
      int rotated = 0;
      // the pair of the next round is looked up before the barrier (the
      // positions do not change within a sweep); after permute_sorted (the
      // QR path) position = column, so no table read is needed
      auto pair_of = [&](int r, int& p, int& q) {
        if (g >= npairs) return false;
        int pp, qq;
        rr_pair_fast(r, g, np, pp, qq);
        p = qr ? pp : col_at[pp];
        q = qr ? qq : col_at[qq];
        return p < n && q < n;
      };
      int pn = 0, qn = 0;
      bool ln = pair_of(0, pn, qn);
      for (int r = 0; r < np - 1; ++r) {
        const int p = pn, q = qn;
        const bool live = ln;
        if (live) {
          double* wp = W + p * CW + t * CHW;
          double* wq = W + q * CW + t * CHW;
          double* vp = V + p * CV + t * CHV;
          double* vq = V + q * CV + t * CHV;
          // the rotation columns are read with the working columns, so their
          // LDS latency overlaps the dot product and its reduction
          double yp[kSmallEN], yq[kSmallEN];
          if (!qr) {
#pragma unroll
            for (int e = 0; e < kSmallEN; e += 2) {
              const double2 u = *reinterpret_cast<const double2*>(vp + e);
              const double2 v = *reinterpret_cast<const double2*>(vq + e);
              yp[e] = u.x;
              yp[e + 1] = u.y;
              yq[e] = v.x;
              yq[e + 1] = v.y;
            }
          }
          double xp[EMX], xq[EMX];
#pragma unroll
          for (int e = 0; e < EMX; e += 2) {
            const double2 u = *reinterpret_cast<const double2*>(wp + e);
            const double2 v = *reinterpret_cast<const double2*>(wq + e);
            xp[e] = u.x;
            xp[e + 1] = u.y;
            xq[e] = v.x;
            xq[e + 1] = v.y;
          }
          // the two norms are read with the columns, so their LDS latency
          // overlaps the dot product instead of following its reduction
          const double a = nrm[p], b = nrm[q];
          __builtin_amdgcn_sched_barrier(0);
          double ca = 0.0, cb = 0.0;
#pragma unroll
          for (int e = 0; e < EMX; e += 2) {
            ca = fma(xp[e], xq[e], ca);
            cb = fma(xp[e + 1], xq[e + 1], cb);
          }
          const double c = group8_sum(ca + cb);
          if (c * c > tol * tol * (a * b) && fabs(c) > abs_tol) {
            rotated = 1;
            // the tangent from the raw reciprocal and square-root estimates
            // (its error only moves the angle); c = 1 / sqrt(1 + t^2) refined,
            // so that (c, c t) is a rotation to working accuracy
            const double zeta = (b - a) * (0.5 * __builtin_amdgcn_rcp(c));
            const double z2 = fma(zeta, zeta, 1.0);
            const double tt = copysign(1.0, zeta) * __builtin_amdgcn_rcp(fabs(zeta) + z2 * __builtin_amdgcn_rsq(z2));
            const double cs = fast_rsqrt(fma(tt, tt, 1.0));
            const double sn = cs * tt;
#pragma unroll
            for (int e = 0; e < EMX; e += 2) {
              *reinterpret_cast<double2*>(wp + e) =
                  double2{cs * xp[e] - sn * xq[e], cs * xp[e + 1] - sn * xq[e + 1]};
              *reinterpret_cast<double2*>(wq + e) =
                  double2{sn * xp[e] + cs * xq[e], sn * xp[e + 1] + cs * xq[e + 1]};
            }
            if (!qr) {
#pragma unroll
              for (int e = 0; e < kSmallEN; e += 2) {
                *reinterpret_cast<double2*>(vp + e) =
                    double2{cs * yp[e] - sn * yq[e], cs * yp[e + 1] - sn * yq[e + 1]};
                *reinterpret_cast<double2*>(vq + e) =
                    double2{sn * yp[e] + cs * yq[e], sn * yp[e + 1] + cs * yq[e + 1]};
              }
            }
            if (t == 0) {
              nrm[p] = fmax(a - tt * c, 0.0);
              nrm[q] = fmax(b + tt * c, 0.0);
            }
          }
        }
        if (r + 1 < np - 1) ln = pair_of(r + 1, pn, qn);
        __syncthreads();
      }
      if (!__syncthreads_or(rotated)) break;
    }
    return sweep;
  };

  // Sweeps of the QR path (X = R^T: n columns of n rows) with the columns in
  // registers.  After each sweep's sort the columns form blocks of 8 in
  // decreasing-norm order; the blocks are paired round-robin, one pair
  // (P, Q) per wave per stage, with a block barrier only between stages.
  // In a stage lane 8 t + j of the wave holds row chunk t of A_j = column
  // 8 P + j and of one column of block Q, B_{j ^ g}, and its squared norm.
  // Round g pairs A_j with B_{j ^ g}; the rounds run g = 0, 1, 3, 2, 6, 7,
  // 5, 4, so that each step moves the B chunks by an xor of 1, 2 or 4
  // groups (DPP moves, no LDS).  In the first stage of a sweep every block
  // also meets itself: for d = 1..7 groups j and j ^ d trade one column, the
  // lower group rotates (A_j, A_{j ^ d}) and the higher (B_{j ^ d}, B_j),
  // and the columns trade back.  Every pair of a sweep is rotated once, by
  // one group; its dot product is a sum over the group's eight lanes that
  // each of them holds bit for bit.
  auto sweeps_blocked = [&]() {
    const double tol = 10.0 * DBL_EPSILON * (double)n, tol2 = tol * tol;
    const int max_sweeps = (int)pr.max_sweeps;
    const int wv = tid >> 6, j = tid & 7, tc = (tid >> 3) & 7;
    const int nbk = (((n + 7) >> 3) + 1) & ~1;   // blocks of 8, an even count
    const bool active = wv < nbk / 2;
    int sweep = 0;
    for (; sweep < max_sweeps; ++sweep) {
      // position k of the sweep holds column col_at[k] (decreasing norm,
      // ties by index): wave w counts, for column tid % 64, the larger
      // norms among columns 16 w .. 16 w + 15, and the four counts add up
      {
        const int jc = tid & 63, i0 = 16 * wv;
        const double sj = nrm[jc];
        int rk = 0;
#pragma unroll
        for (int i = 0; i < 16; i += 2) {
          const double2 s2 = *reinterpret_cast<const double2*>(nrm + i0 + i);
          rk += (i0 + i < n) && ((s2.x > sj) || (s2.x == sj && i0 + i < jc));
          rk += (i0 + i + 1 < n) && ((s2.y > sj) || (s2.y == sj && i0 + i + 1 < jc));
        }
        rank_of[tid] = rk;
      }
      __syncthreads();
      if (tid < n) col_at[rank_of[tid] + rank_of[tid + 64] + rank_of[tid + 128] + rank_of[tid + 192]] = tid;
      __syncthreads();
      int rotated = 0;
      for (int stage = 0; stage < nbk - 1; ++stage) {
        int P = 0, Q = 0;
        double a[EMX], b[EMX], na = 0.0, nb = 0.0;
        if (active) {
          rr_pair_fast(stage, wv, nbk, P, Q);
          // the first stage gathers every column from where the previous
          // sweep left it into its sorted position
          const int pa = 8 * P + j, pb = 8 * Q + j;
          const int sa = (pa < n) ? (stage == 0 ? col_at[pa] : pa) : -1;
          const int sb = (pb < n) ? (stage == 0 ? col_at[pb] : pb) : -1;
          const double* wa = W + (sa < 0 ? 0 : sa) * CW + tc * CHW;
          const double* wb = W + (sb < 0 ? 0 : sb) * CW + tc * CHW;
#pragma unroll
          for (int e = 0; e < EMX; e += 2) {
            const double2 u = *reinterpret_cast<const double2*>(wa + e);
            const double2 v = *reinterpret_cast<const double2*>(wb + e);
            a[e] = (sa < 0) ? 0.0 : u.x;
            a[e + 1] = (sa < 0) ? 0.0 : u.y;
            b[e] = (sb < 0) ? 0.0 : v.x;
            b[e + 1] = (sb < 0) ? 0.0 : v.y;
          }
          na = (sa < 0) ? 0.0 : nrm[sa];
          nb = (sb < 0) ? 0.0 : nrm[sb];
        }
        // every column is read before any is stored in its new place
        if (stage == 0) __syncthreads();
        if (active) {
          const int ca = 8 * P + j;

          // one intra-block round: groups j and j ^ D trade a column
          auto intra = [&](auto dc) {
            constexpr int D = decltype(dc)::value;
            constexpr int HB = (D >= 4) ? 4 : (D >= 2) ? 2 : 1;
            const bool low = (j & HB) == 0;
            const int cmax = (low ? 8 * P : 8 * Q) + (j > (j ^ D) ? j : (j ^ D));
            double own[EMX], r[EMX];
#pragma unroll
            for (int e = 0; e < EMX; ++e) {
              own[e] = low ? a[e] : b[e];
              r[e] = low ? b[e] : a[e];
            }
            double n_own = low ? na : nb;
            double nr = group_xor<D>(low ? nb : na);
            group_xor<D>(r);
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int e = 0; e < EMX; e += 2) {
              s0 = fma(own[e], r[e], s0);
              s1 = fma(own[e + 1], r[e + 1], s1);
            }
            const double c = group_sum_strided(s0 + s1);
            // the lower group's pair is (own, r), the higher group's (r, own)
            const double np_ = low ? n_own : nr, nq_ = low ? nr : n_own;
            const bool rot = (cmax < n) & (c * c > tol2 * (np_ * nq_));
            double cs = 1.0, sn = 0.0, tc_ = 0.0;
            if (__builtin_amdgcn_ballot_w64(rot)) {
              rotated |= rot;
              const JacobiRot q = jacobi_rot(np_, nq_, rot ? c : 1.0);
              cs = rot ? q.cs : 1.0;
              sn = rot ? q.sn : 0.0;
              tc_ = rot ? q.tt * c : 0.0;
            }
            const double s2 = low ? -sn : sn;
#pragma unroll
            for (int e = 0; e < EMX; ++e) {
              const double u = own[e], v = r[e];
              own[e] = cs * u + s2 * v;
              r[e] = cs * v - s2 * u;
            }
            n_own = fmax(low ? n_own - tc_ : n_own + tc_, 0.0);
            nr = fmax(low ? nr + tc_ : nr - tc_, 0.0);
            group_xor<D>(r);
            nr = group_xor<D>(nr);
#pragma unroll
            for (int e = 0; e < EMX; ++e) {
              a[e] = low ? own[e] : r[e];
              b[e] = low ? r[e] : own[e];
            }
            na = low ? n_own : nr;
            nb = low ? nr : n_own;
          };
          // one cross round: (A_j, B_{j ^ g}), A the first of the pair
          auto cross = [&](int g) {
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int e = 0; e < EMX; e += 2) {
              s0 = fma(a[e], b[e], s0);
              s1 = fma(a[e + 1], b[e + 1], s1);
            }
            const double c = group_sum_strided(s0 + s1);
            const bool rot = (ca < n) & (8 * Q + (j ^ g) < n) & (c * c > tol2 * (na * nb));
            // a skipped pair applies (1, 0), which leaves its columns and
            // norms unchanged bit for bit; a wave with no pair to rotate
            // skips the rotation's square roots and reciprocals
            double cs = 1.0, sn = 0.0, tc_ = 0.0;
            if (__builtin_amdgcn_ballot_w64(rot)) {
              rotated |= rot;
              const JacobiRot q = jacobi_rot(na, nb, rot ? c : 1.0);
              cs = rot ? q.cs : 1.0;
              sn = rot ? q.sn : 0.0;
              tc_ = rot ? q.tt * c : 0.0;
            }
#pragma unroll
            for (int e = 0; e < EMX; ++e) {
              const double u = a[e], v = b[e];
              a[e] = cs * u - sn * v;
              b[e] = sn * u + cs * v;
            }
            na = fmax(na - tc_, 0.0);
            nb = fmax(nb + tc_, 0.0);
          };
          auto move = [&](auto dc) {
            constexpr int D = decltype(dc)::value;
            group_xor<D>(b);
            nb = group_xor<D>(nb);
          };
          using I1 = std::integral_constant<int, 1>;
          using I2 = std::integral_constant<int, 2>;
          using I4 = std::integral_constant<int, 4>;
          if (stage == 0) {
            intra(I1{});
            intra(I2{});
            intra(std::integral_constant<int, 3>{});
            intra(I4{});
            intra(std::integral_constant<int, 5>{});
            intra(std::integral_constant<int, 6>{});
            intra(std::integral_constant<int, 7>{});
          }
          cross(0);
          move(I1{});
          cross(1);
          move(I2{});
          cross(3);
          move(I1{});
          cross(2);
          move(I4{});
          cross(6);
          move(I1{});
          cross(7);
          move(I2{});
          cross(5);
          move(I1{});
          cross(4);
          // B_{j ^ 4} is now in group j
          const int cb = 8 * Q + (j ^ 4);
          double* pa = W + ca * CW + tc * CHW;
          double* pb = W + cb * CW + tc * CHW;
#pragma unroll
          for (int e = 0; e < EMX; e += 2) {
            *reinterpret_cast<double2*>(pa + e) = double2{a[e], a[e + 1]};
            *reinterpret_cast<double2*>(pb + e) = double2{b[e], b[e + 1]};
          }
          // the last stage holds every block: it leaves the exact squared
          // norms for the next sweep's ranking
          if (stage == nbk - 2) {
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int e = 0; e < EMX; ++e) {
              s0 = fma(a[e], a[e], s0);
              s1 = fma(b[e], b[e], s1);
            }
            na = group_sum_strided(s0);
            nb = group_sum_strided(s1);
          }
          if (tc == 0) {
            if (ca < n) nrm[ca] = na;
            if (cb < n) nrm[cb] = nb;
          }
        }
        __syncthreads();
      }
      if (!__syncthreads_or(rotated)) break;
    }
    return sweep;
  };

  bool qr = !tr && !want_u;
  int total = 0;
  bool nonconv = false;
  if (qr) {
    householder_rt();
    exact_norms();
    const int sw = sweeps_blocked();
    total = sw + 1;
    nonconv = sw >= (int)pr.max_sweeps;
    exact_norms();
    // a column at the underflow level cannot be normalised: plain sweeps
    int tiny = 0;
    for (int c = tid; c < n; c += kThreads) tiny |= !(nrm[c] >= 1e-270);
    if (__syncthreads_or(tiny)) {
      qr = false;
      load_w(false);
      exact_norms();
    }
  }
  if (!qr) {
    const int sw = sweeps(false);
    total += sw + 1;
    nonconv = sw >= (int)pr.max_sweeps;
    exact_norms();
  }
  if (tid == 0) {
    if (nonconv) atomicAdd(status, 1);
    atomicMax(status + 1, total);
  }

  // singular values = exact column norms; descending rank (ties by index)
  for (int j = tid; j < n; j += kThreads) {
    const double sj = nrm[j];
    int rk = 0;
    for (int i = 0; i < n; ++i) {
      const double si = nrm[i];
      rk += (si > sj) || (si == sj && i < j);
    }
    rank_of[j] = rk;
  }
  __syncthreads();
  double* Ub = U + pr.u_off;
  double* Sb = S + pr.s_off;
  double* VTb = VT + pr.vt_off;
  for (int j = tid; j < n; j += kThreads) Sb[rank_of[j]] = sqrt(nrm[j]);
  if (qr) {
    // 1 / sigma_k once per column (hh_v is free here), read by the scaling
    // of the first product below instead of a division per element
    for (int j = tid; j < n; j += kThreads) hh_v[j] = 1.0 / sqrt(nrm[j]);
    __syncthreads();
    // P = U_X^T (V0 Pi)^T, row k = the sum over QR positions i of
    // X[i][k] / sigma_k times V0[r][qr_col[i]]; then one Newton-Schulz step
    // VT = (3 P - (P P^T) P) / 2, which makes the rows orthonormal to
    // rounding (U_X is orthogonal only to the sweeps' tolerance, and a warm
    // start would carry that error from one call to the next).  The three
    // D x D products run on the FP64 matrix cores from LDS: wave w owns a
    // (D/2)-square quadrant of 16 x 16 tiles; lane l supplies A[l%16][l/16]
    // and B[l/16][l%16], accumulator element v is C[l/16 + 4v][l%16].
    constexpr int D = (kRows >= 64) ? 64 : 32;
    constexpr int NT = (D == 64) ? 4 : 1;
    const int w = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
    const int kmax = (n + 3) & ~3;
    auto tile_r = [&](int tt) { return (D == 64) ? ((w >> 1) * 32 + 16 * (tt >> 1)) : 16 * (w >> 1); };
    auto tile_c = [&](int tt) { return (D == 64) ? ((w & 1) * 32 + 16 * (tt & 1)) : 16 * (w & 1); };
    auto mm = [&](auto a_at, auto b_at, d4s(&acc)[NT]) {
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) acc[tt] = d4s{0.0, 0.0, 0.0, 0.0};
      // four k-steps' operand reads issued together
#pragma unroll 4
      for (int kk = 0; kk < kmax; kk += 4) {
        const int k = kk + lk;
#pragma unroll
        for (int tt = 0; tt < NT; ++tt)
          acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_at(tile_r(tt) + lr, k), b_at(k, tile_c(tt) + lr), acc[tt],
                                                         0, 0, 0);
      }
    };
    d4s pk[NT], hk[NT];
    mm([&](int k, int i) { return W[widx(k, i)]; },
       [&](int i, int r) { return (i < n && r < n) ? V[vidx(warm ? i : qr_col[i], r)] : 0.0; }, pk);   // warm: qr_col = identity
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int k = tile_r(tt) + lk + 4 * v;
        pk[tt][v] = (k < n) ? pk[tt][v] * hh_v[k] : 0.0;
      }
    __syncthreads();
    // P row k -> W column k
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
#pragma unroll
      for (int v = 0; v < 4; ++v) W[widx(tile_r(tt) + lk + 4 * v, tile_c(tt) + lr)] = pk[tt][v];
    __syncthreads();
    // G = P P^T -> V column k
    mm([&](int k, int r) { return W[widx(k, r)]; }, [&](int r, int c) { return W[widx(c, r)]; }, hk);
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
#pragma unroll
      for (int v = 0; v < 4; ++v) V[vidx(tile_r(tt) + lk + 4 * v, tile_c(tt) + lr)] = hk[tt][v];
    __syncthreads();
    // H = G P; VT row rank(k) = (3 P - H) / 2
    mm([&](int k, int c) { return V[vidx(k, c)]; }, [&](int c, int r) { return W[widx(c, r)]; }, hk);
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int k = tile_r(tt) + lk + 4 * v, r = tile_c(tt) + lr;
        if (k < n && r < n) VTb[(int64_t)rank_of[k] * n + r] = 1.5 * pk[tt][v] - 0.5 * hk[tt][v];
      }
    return;
  }
  for (int e = tid; e < m * n; e += kThreads) {
    const int j = e / m, i = e % m;
    const double sj = sqrt(nrm[j]);
    const double v = (sj > 0.0) ? W[widx(j, i)] / sj : 0.0;
    const int rk = rank_of[j];
    if (!tr) {
      if (want_u) Ub[(int64_t)i * n + rk] = v;
    } else {
      VTb[(int64_t)rk * N + i] = v;
    }
  }
  for (int e = tid; e < n * n; e += kThreads) {
    const int j = e / n, i = e % n;
    const int rk = rank_of[j];
    const double vv = V[vidx(j, i)];
    if (!tr)
      VTb[(int64_t)rk * n + i] = vv;
    else if (want_u)
      Ub[(int64_t)i * n + rk] = vv;
  }
}

// A batch of SVD problems with its device-side description, scratch and
// status uploaded once: executing it only launches kernels (no allocation,
// no host round trip), so it can be replayed and captured into a graph.
}  // namespace

struct tn_svd_plan_s {
  std::vector<SvdProb> probs;   // small-kernel problems, then LDS ones
  size_t n_small = 0, n_lds = 0;
  int64_t small_mm = 0;
  size_t lds_need = 0;
  int tp_lds = 8;
  SvdProb* d_p = nullptr;
  tn::bj::Plan* blk = nullptr;  // block Jacobi: problems too large for LDS
  int* d_status = nullptr;      // [0] problems not converged, [1] max sweeps
  bool v0_any = false;          // some problem has a warm start
  bool u_any = false;           // some problem writes U (u_off >= 0)
  std::vector<int64_t> orig;    // caller's index of each entry of probs
  std::vector<int64_t> blk_orig;   // caller's index of each block-Jacobi problem
};

namespace {

void svd_plan_free(tn_svd_plan_s* P) {
  tn::pool_free(P->d_status);
  tn::pool_free(P->d_p);
  tn::bj::destroy(P->blk);
  delete P;
}

typedef void (*SvdKern)(const SvdProb*, const double*, const double*, double*, double*, double*, int*);
typedef void (*SvdSmallKern)(const SvdProb*, const double*, const double*, double*, double*, double*, int*);

SvdKern pick_svd_kernel(int tp) {
  return (tp == 8) ? svd_jacobi_kernel<8> : (tp == 4) ? svd_jacobi_kernel<4> : svd_jacobi_kernel<2>;
}

void raise_lds_limit(int tp, size_t bytes);

// Problems beyond the small kernel with min(M, N) >= this go to the block
// path; smaller ones that fit one workgroup's LDS to the LDS kernel.
constexpr int64_t kBlockMinDefault = kSmallN + 1;
int64_t g_block_min = kBlockMinDefault;
// Sweep cap of the one-workgroup kernels (60) and of the block path (40);
// tn_svd_set_max_sweeps lowers it for plans created afterwards.
constexpr int kOneWgMaxSweeps = 60;
std::atomic<int> g_max_sweeps{0};

int svd_plan_build(tn_svd_plan_s** out, int64_t batch, const int64_t* m, const int64_t* n,
                   const int64_t* a_off, const int64_t* v0_off, const int64_t* u_off,
                   const int64_t* s_off, const int64_t* vt_off) {
  std::vector<SvdProb> small_probs, lds_probs;
  std::vector<int64_t> small_orig, lds_orig;
  std::vector<tn::bj::Spec> blk;
  auto* P = new tn_svd_plan_s();
  for (int64_t i = 0; i < batch; ++i) {
    if (m[i] < 0 || n[i] < 0) {
      delete P;
      tn::set_error("svd: negative shape");
      return TN_EINVAL;
    }
    const int64_t mm = std::max(m[i], n[i]), nn = std::min(m[i], n[i]);
    const int cap = g_max_sweeps.load();
    SvdProb p{m[i], n[i], a_off ? a_off[i] : 0, u_off ? u_off[i] : 0, s_off ? s_off[i] : 0,
              vt_off ? vt_off[i] : 0, v0_off ? v0_off[i] : -1, cap > 0 ? std::min(cap, kOneWgMaxSweeps) : kOneWgMaxSweeps,
              nullptr, -1};
    P->v0_any |= p.v0_off >= 0;
    P->u_any |= p.u_off >= 0;
    const size_t bytes = sizeof(double) * (size_t)((mm + 1) * nn + (nn + 1) * nn);
    if (nn <= kSmallN && mm <= kSmallTP * 16) {
      small_probs.push_back(p);
      small_orig.push_back(i);
      P->small_mm = std::max(P->small_mm, mm);
    } else if (bytes <= kLdsBytes && nn < g_block_min) {
      P->lds_need = std::max(P->lds_need, bytes);
      lds_probs.push_back(p);
      lds_orig.push_back(i);
    } else if (mm > 0 && nn > 0) {
      blk.push_back(tn::bj::Spec{p.m, p.n, p.a_off, p.u_off, p.s_off, p.vt_off, p.v0_off});
      P->blk_orig.push_back(i);
    }
  }
  // lanes per column pair: as few as the widest problem allows with one
  // pass over the pairs of a round (8 lanes up to 64 columns)
  auto pick_tp = [](const std::vector<SvdProb>& v) {
    int64_t mx = 0;
    for (auto& p : v) mx = std::max(mx, std::min(p.m, p.n));
    const int64_t pairs = (mx + 1) / 2;
    return pairs <= 32 ? 8 : (pairs <= 64 ? 4 : 2);
  };
  P->tp_lds = pick_tp(lds_probs);
  if (!lds_probs.empty()) raise_lds_limit(P->tp_lds, P->lds_need);
  P->n_small = small_probs.size();
  P->n_lds = lds_probs.size();
  P->probs = small_probs;
  P->probs.insert(P->probs.end(), lds_probs.begin(), lds_probs.end());
  P->orig = small_orig;
  P->orig.insert(P->orig.end(), lds_orig.begin(), lds_orig.end());
  int rc = tn::bj::build(&P->blk, blk, g_max_sweeps.load());
  if (rc == TN_OK) rc = tn::pool_alloc((void**)&P->d_status, 2 * sizeof(int));
  if (rc == TN_OK) {
    const int zero[2] = {0, 0};   // counters start clear (synchronous copy)
    rc = tn_memcpy_h2d(P->d_status, zero, sizeof(zero), nullptr);
  }
  if (rc == TN_OK && !P->probs.empty()) rc = tn::pool_alloc((void**)&P->d_p, sizeof(SvdProb) * P->probs.size());
  if (rc == TN_OK && !P->probs.empty())
    rc = tn_memcpy_h2d(P->d_p, P->probs.data(), sizeof(SvdProb) * P->probs.size(), nullptr);
  if (rc != TN_OK) {
    svd_plan_free(P);
    return rc;
  }
  *out = P;
  return TN_OK;
}

// The dynamic LDS limit of each LDS-resident kernel only ever grows (set at
// plan creation, so that launches, which may be captured into a graph, make
// no attribute calls); a failure surfaces as a launch error.
void raise_lds_limit(int tp, size_t bytes) {
  static std::mutex mu;
  static size_t limit[3] = {0, 0, 0};
  const int slot = (tp == 8) ? 0 : (tp == 4) ? 1 : 2;
  std::lock_guard<std::mutex> g(mu);
  if (bytes <= limit[slot]) return;
  (void)hipFuncSetAttribute((const void*)pick_svd_kernel(tp), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)bytes);
  (void)hipGetLastError();
  limit[slot] = bytes;
}

bool svd_plan_empty(const tn_svd_plan_s* P) { return P->probs.empty() && tn::bj::empty(P->blk); }

// The status counters (problems not converged, most sweeps) accumulate over
// executions; svd_plan_status reads and clears them, so a non-converged
// problem in any execution since the last check is reported.
int svd_plan_launch(const tn_svd_plan_s* P, const double* A, const double* V0T, double* U, double* S,
                    double* VT, hipStream_t st) {
  if (P->n_small > 0) {
    SvdSmallKern k = (P->small_mm <= 32) ? svd_jacobi_small<4>
                     : (P->small_mm <= 64) ? svd_jacobi_small<8> : svd_jacobi_small<16>;
    hipLaunchKernelGGL(k, dim3((unsigned)P->n_small), dim3(kThreads), 0, st, P->d_p, A, V0T, U, S, VT, P->d_status);
    TN_LAUNCH_CHECK();
  }
  if (P->n_lds > 0) {
    SvdKern k = pick_svd_kernel(P->tp_lds);
    hipLaunchKernelGGL(k, dim3((unsigned)P->n_lds), dim3(kThreads), P->lds_need, st, P->d_p + P->n_small, A, V0T, U,
                       S, VT, P->d_status);
    TN_LAUNCH_CHECK();
  }
  return tn::bj::launch(P->blk, A, V0T, U, S, VT, P->d_status, st);
}

int svd_plan_status(const tn_svd_plan_s* P, int* nonconv, int* sweeps, hipStream_t st) {
  int h[2] = {0, 0};
  int rc = tn_memcpy_d2h(h, P->d_status, 2 * sizeof(int), st);   // synchronises
  if (rc != TN_OK) return rc;
  rc = tn_zero(P->d_status, 2 * sizeof(int), st);                // fill kernel, ordered before later runs
  if (rc != TN_OK) return rc;
  g_last_sweeps = h[1];
  if (nonconv) *nonconv = h[0];
  if (sweeps) *sweeps = h[1];
  return TN_OK;
}

// One-shot batch: plan, launch, wait, check.
int svd_batched(int64_t batch, const int64_t* m, const int64_t* n, const double* A, const int64_t* a_off,
                const double* V0T, const int64_t* v0_off, double* U, const int64_t* u_off, double* S,
                const int64_t* s_off, double* VT, const int64_t* vt_off, hipStream_t st) {
  if (batch == 0) return TN_OK;
  tn_svd_plan_s* P = nullptr;
  int rc = svd_plan_build(&P, batch, m, n, a_off, v0_off, u_off, s_off, vt_off);
  if (rc != TN_OK) return rc;
  rc = svd_plan_launch(P, A, V0T, U, S, VT, st);
  int nonconv = 0;
  if (rc == TN_OK) rc = svd_plan_status(P, &nonconv, nullptr, st);
  svd_plan_free(P);
  if (rc != TN_OK) return rc;
  TN_REQUIRE(nonconv == 0, TN_ENOCONV, "svd: %d problem(s) did not converge in 60 sweeps", nonconv);
  return TN_OK;
}

// ---------------------------------------------------------------------------
// The allocation-free SVD (tn_svd_f64_noalloc): the one-workgroup kernel's
// problem descriptor and convergence counters live in this file-scope device
// array instead of a pool-allocated plan.  256 bytes per device, held until
// the process ends; every call on a device shares them, so calls that
// overlap on one device race.
struct NoAllocScratch {
  SvdProb prob;
  int status[2];
};
static_assert(sizeof(NoAllocScratch) <= 256, "no-alloc scratch too small");

static __device__ __align__(256) unsigned char scratch[256];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
NoAllocScratch* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch`, looked up once per
// device with that device current (the array's name is not a device address
// in host code).
int scratch_ptr(NoAllocScratch** out) {
  int dev = 0;
  TN_HIP_CHECK(hipGetDevice(&dev));
  TN_REQUIRE(dev >= 0 && dev < kMaxDevices, TN_EUNSUPP, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    TN_REQUIRE(e == hipSuccess && ptr != nullptr, TN_EHIP, "hipGetSymbolAddress(scratch): %s",
               hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<NoAllocScratch*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return TN_OK;
}

// Writes the problem descriptor and clears the counters (a kernel, so no
// memset node or host staging is involved).
__global__ void noalloc_setup_kernel(NoAllocScratch* __restrict__ d, SvdProb p) {
  if (threadIdx.x == 0) {
    d->prob = p;
    d->status[0] = 0;
    d->status[1] = 0;
  }
}

}  // namespace


extern "C" {

int tn_svd_f64_noalloc(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(A && S && VT, TN_EINVAL, "tn_svd_f64_noalloc: NULL buffer");
  TN_REQUIRE(M >= 0 && N >= 0, TN_EINVAL, "tn_svd_f64_noalloc: negative shape");
  const int64_t mm = std::max(M, N), nn = std::min(M, N);
  TN_REQUIRE(nn <= kSmallN && mm <= kSmallTP * 16, TN_EUNSUPP,
             "tn_svd_f64_noalloc: %lld x %lld exceeds the maximum min(M, N) <= %d, max(M, N) <= %d",
             (long long)M, (long long)N, kSmallN, kSmallTP * 16);
  NoAllocScratch* d = nullptr;
  int rc = scratch_ptr(&d);
  if (rc != TN_OK) return rc;
  hipStream_t st = tn::stream_of(stream);
  const int cap = g_max_sweeps.load();
  const SvdProb p{M, N, 0, U ? 0 : -1, 0, 0, -1, cap > 0 ? std::min(cap, kOneWgMaxSweeps) : kOneWgMaxSweeps,
                  nullptr, -1};
  hipLaunchKernelGGL(noalloc_setup_kernel, dim3(1), dim3(64), 0, st, d, p);
  TN_LAUNCH_CHECK();
  // the same kernel and template choice as a plan holding this one problem
  SvdSmallKern k = (mm <= 32) ? svd_jacobi_small<4> : (mm <= 64) ? svd_jacobi_small<8> : svd_jacobi_small<16>;
  hipLaunchKernelGGL(k, dim3(1), dim3(kThreads), 0, st, &d->prob, A, nullptr, U, S, VT, d->status);
  TN_LAUNCH_CHECK();
  int h[2] = {0, 0};
  rc = tn_memcpy_d2h(h, d->status, sizeof(h), stream);   // synchronises the stream
  if (rc != TN_OK) return rc;
  g_last_sweeps = h[1];
  TN_REQUIRE(h[0] == 0, TN_ENOCONV, "svd: %d problem(s) did not converge in %d sweeps", h[0], (int)p.max_sweeps);
  return TN_OK;
}

int tn_svd_set_max_sweeps(int n) {
  TN_REQUIRE(n >= 0, TN_EINVAL, "tn_svd_set_max_sweeps: negative count");
  g_max_sweeps.store(n);
  return TN_OK;
}

int tn_svd_set_block_min(int64_t n) {
  TN_REQUIRE(n >= 0, TN_EINVAL, "tn_svd_set_block_min: negative size");
  g_block_min = (n == 0) ? kBlockMinDefault : n;
  return TN_OK;
}

int tn_svd_last_sweeps(int* sweeps) {
  TN_REQUIRE(sweeps != nullptr, TN_EINVAL, "tn_svd_last_sweeps: NULL out pointer");
  *sweeps = g_last_sweeps;
  return TN_OK;
}

int tn_svd_f64(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT,
               void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(A && S && VT, TN_EINVAL, "tn_svd_f64: NULL buffer");
  int64_t m = M, n = N, z = 0, uo = U ? 0 : -1;
  return svd_batched(1, &m, &n, A, &z, nullptr, nullptr, U, &uo, S, &z, VT, &z, tn::stream_of(stream));
}

int tn_svd_f64_warm(const double* A, int64_t M, int64_t N, const double* V0T, double* U, double* S,
                    double* VT, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(A && S && VT && V0T, TN_EINVAL, "tn_svd_f64_warm: NULL buffer");
  TN_REQUIRE(M >= N, TN_EINVAL, "tn_svd_f64_warm: needs M >= N (got %lld x %lld)", (long long)M, (long long)N);
  int64_t m = M, n = N, z = 0, uo = U ? 0 : -1;
  return svd_batched(1, &m, &n, A, &z, V0T, &z, U, &uo, S, &z, VT, &z, tn::stream_of(stream));
}

int tn_svd_f64_batched(int64_t batch, const int64_t* m, const int64_t* n, const double* A,
                       const int64_t* a_off, double* U, const int64_t* u_off, double* S,
                       const int64_t* s_off, double* VT, const int64_t* vt_off, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(batch >= 0 && (batch == 0 || (m && n && A && U && S && VT)), TN_EINVAL,
             "tn_svd_f64_batched: NULL argument");
  return svd_batched(batch, m, n, A, a_off, nullptr, nullptr, U, u_off, S, s_off, VT, vt_off,
                     tn::stream_of(stream));
}

int tn_svd_plan_create(tn_svd_plan* plan, int64_t batch, const int64_t* m, const int64_t* n,
                       const int64_t* a_off, const int64_t* v0_off, const int64_t* u_off,
                       const int64_t* s_off, const int64_t* vt_off) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(plan != nullptr, TN_EINVAL, "tn_svd_plan_create: NULL out pointer");
  TN_REQUIRE(batch >= 0 && (batch == 0 || (m && n)), TN_EINVAL, "tn_svd_plan_create: NULL shapes");
  if (v0_off)
    for (int64_t i = 0; i < batch; ++i)
      TN_REQUIRE(v0_off[i] < 0 || m[i] >= n[i], TN_EINVAL, "tn_svd_plan_create: warm start needs M >= N");
  return svd_plan_build(plan, batch, m, n, a_off, v0_off, u_off, s_off, vt_off);
}

int tn_svd_plan_exec(tn_svd_plan P, const double* A, const double* V0T, double* U, double* S, double* VT,
                     void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_svd_plan_exec: NULL plan");
  if (svd_plan_empty(P)) return TN_OK;
  TN_REQUIRE(A && S && VT && (U || !P->u_any), TN_EINVAL, "tn_svd_plan_exec: NULL buffer");
  bool needs_v0 = P->v0_any;
  TN_REQUIRE(!needs_v0 || V0T, TN_EINVAL, "tn_svd_plan_exec: the plan has warm starts but V0T is NULL");
  return svd_plan_launch(P, A, V0T, U, S, VT, tn::stream_of(stream));
}

int tn_svd_plan_set_mask(tn_svd_plan P, const int* mask, const int32_t* prob_idx) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_svd_plan_set_mask: NULL plan");
  TN_REQUIRE(mask == nullptr || prob_idx != nullptr || (P->probs.empty() && P->blk_orig.empty()), TN_EINVAL,
             "tn_svd_plan_set_mask: NULL problem indices");
  int rc = tn_device_sync();   // launches queued with the old descriptors finish first
  if (rc != TN_OK) return rc;
  for (size_t k = 0; k < P->probs.size(); ++k) {
    P->probs[k].mask = mask;
    P->probs[k].mask_idx = mask ? (int64_t)prob_idx[P->orig[k]] : -1;
  }
  if (!P->probs.empty()) rc = tn_memcpy_h2d(P->d_p, P->probs.data(), sizeof(SvdProb) * P->probs.size(), nullptr);
  if (rc != TN_OK) return rc;
  std::vector<int64_t> bidx(P->blk_orig.size(), -1);
  for (size_t k = 0; k < bidx.size(); ++k) bidx[k] = mask ? (int64_t)prob_idx[P->blk_orig[k]] : -1;
  return tn::bj::set_mask(P->blk, mask, bidx);
}

int tn_svd_plan_check(tn_svd_plan P, int* nonconverged, int* sweeps, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_svd_plan_check: NULL plan");
  return svd_plan_status(P, nonconverged, sweeps, tn::stream_of(stream));
}

int tn_svd_plan_destroy(tn_svd_plan P) {
  if (!P) return TN_OK;
  const hipError_t e = hipDeviceSynchronize();   // launches may still read the plan
  svd_plan_free(P);
  TN_REQUIRE(e == hipSuccess, TN_EHIP, "hipDeviceSynchronize: %s", hipGetErrorString(e));
  return TN_OK;
}

// Launches the eigh kernel for `batch` problems (probs on the device, or
// `one` for a single problem) and reports how many did not converge.
static int eigh_launch(int64_t batch, const EighProb* d_probs, EighProb one, int64_t nmax, const double* A,
                       double* eig, double* vec_rows, hipStream_t st) {
  int* d_status = nullptr;
  int rc = tn::pool_alloc((void**)&d_status, sizeof(int));
  if (rc != TN_OK) return rc;
  rc = tn_zero(d_status, sizeof(int), st);
  if (rc == TN_OK) {
    size_t lds = sizeof(double) * 2 * nmax * nmax;
    static std::once_flag attr;
    std::call_once(attr, [] {
      (void)hipFuncSetAttribute((const void*)eigh_jacobi_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                140 * 1024);
      (void)hipGetLastError();
    });
    hipLaunchKernelGGL(eigh_jacobi_kernel, dim3((unsigned)batch), dim3(kThreads), lds, st, d_probs, one, A, eig,
                       vec_rows, d_status);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      tn::set_error("eigh kernel launch: %s", hipGetErrorString(e));
      rc = TN_EHIP;
    }
  }
  int h = 0;
  if (rc == TN_OK) rc = tn_memcpy_d2h(&h, d_status, sizeof(int), st);
  tn::pool_free(d_status);
  if (rc != TN_OK) return rc;
  TN_REQUIRE(h == 0, TN_ENOCONV, "eigh: Jacobi did not converge for %d matri%s", h, h == 1 ? "x" : "ces");
  return TN_OK;
}

int tn_eigh_f64(const double* A, int64_t N, double* eig, double* vec_rows, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(A && eig && vec_rows, TN_EINVAL, "tn_eigh_f64: NULL buffer");
  TN_REQUIRE(N >= 1 && N <= 90, TN_EUNSUPP, "tn_eigh_f64: N=%lld outside [1, 90]", (long long)N);
  return eigh_launch(1, nullptr, EighProb{N, 0, 0, 0}, N, A, eig, vec_rows, tn::stream_of(stream));
}

int tn_eigh_f64_batched(int64_t batch, const int64_t* n, const double* A, const int64_t* a_off, double* eig,
                        const int64_t* e_off, double* vec_rows, const int64_t* v_off, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(batch >= 0, TN_EINVAL, "tn_eigh_f64_batched: negative batch");
  if (batch == 0) return TN_OK;
  TN_REQUIRE(n && A && a_off && eig && e_off && vec_rows && v_off, TN_EINVAL, "tn_eigh_f64_batched: NULL argument");
  TN_REQUIRE(batch < (int64_t(1) << 31), TN_EUNSUPP, "tn_eigh_f64_batched: batch too large");
  std::vector<EighProb> h((size_t)batch);
  int64_t nmax = 0;
  for (int64_t i = 0; i < batch; ++i) {
    TN_REQUIRE(n[i] >= 0 && n[i] <= 90, TN_EUNSUPP, "tn_eigh_f64_batched: n[%lld]=%lld outside [0, 90]",
               (long long)i, (long long)n[i]);
    TN_REQUIRE(a_off[i] >= 0 && e_off[i] >= 0 && v_off[i] >= 0, TN_EINVAL, "tn_eigh_f64_batched: negative offset");
    h[i] = EighProb{n[i], a_off[i], e_off[i], v_off[i]};
    nmax = std::max(nmax, n[i]);
  }
  if (nmax == 0) return TN_OK;
  hipStream_t st = tn::stream_of(stream);
  EighProb* d = nullptr;
  int rc = tn::pool_alloc((void**)&d, sizeof(EighProb) * (size_t)batch);
  if (rc != TN_OK) return rc;
  rc = tn_memcpy_h2d(d, h.data(), sizeof(EighProb) * (size_t)batch, st);
  if (rc == TN_OK) rc = eigh_launch(batch, d, EighProb{0, 0, 0, 0}, nmax, A, eig, vec_rows, st);
  tn::pool_free(d);
  return rc;
}

int tn_sum_f64(const double* x, int64_t n, int64_t inc, double* out_dev, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(out_dev && (n == 0 || x) && n >= 0 && inc >= 1, TN_EINVAL, "tn_sum_f64: bad argument");
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(1024), 0, tn::stream_of(stream), x, n, inc, out_dev);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_fill_uniform_f64(double* dst, int64_t n, uint64_t seed, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(n >= 0 && (n == 0 || dst), TN_EINVAL, "tn_fill_uniform_f64: bad argument");
  if (n == 0) return TN_OK;
  hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(n)), dim3(256), 0, tn::stream_of(stream), dst, n, seed);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_axpy_f64(double a, const double* x, double* y, int64_t n, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(n >= 0 && (n == 0 || (x && y)), TN_EINVAL, "tn_axpy_f64: bad argument");
  if (n == 0) return TN_OK;
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(256), 0, tn::stream_of(stream), a, x, y, n);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_copy_strided_f64(double* dst, int64_t dst_inc, const double* src, int64_t src_inc, int64_t n,
                        void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(n >= 0 && dst_inc >= 1 && src_inc >= 1 && (n == 0 || (dst && src)), TN_EINVAL,
             "tn_copy_strided_f64: bad argument");
  if (n == 0) return TN_OK;
  hipLaunchKernelGGL(copy_strided_kernel, dim3(grid_for(n)), dim3(256), 0, tn::stream_of(stream), dst, dst_inc,
                     src, src_inc, n);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_norm2_f64(const double* x, int64_t n, double* out_dev, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(out_dev && (n == 0 || x) && n >= 0, TN_EINVAL, "tn_norm2_f64: bad argument");
  hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, tn::stream_of(stream), x, n, out_dev);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_scale_f64(double* x, int64_t n, double a, void* stream) {
  TN_INIT_OR_RETURN();
  if (n == 0) return TN_OK;
  TN_REQUIRE(x && n > 0, TN_EINVAL, "tn_scale_f64: bad argument");
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, tn::stream_of(stream), x, n, a);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_scale_inv_dev_f64(double* x, int64_t n, const double* d, void* stream) {
  TN_INIT_OR_RETURN();
  if (n == 0) return TN_OK;
  TN_REQUIRE(x && d && n > 0, TN_EINVAL, "tn_scale_inv_dev_f64: bad argument");
  hipLaunchKernelGGL(scale_inv_dev_kernel, dim3(grid_for(n)), dim3(256), 0, tn::stream_of(stream),
                     x, n, d);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_diag_row_mul_f64(double* mat, const double* diag, int64_t M, int64_t N, void* stream) {
  TN_INIT_OR_RETURN();
  if (M == 0 || N == 0) return TN_OK;
  TN_REQUIRE(mat && diag && M > 0 && N > 0, TN_EINVAL, "tn_diag_row_mul_f64: bad argument");
  hipLaunchKernelGGL(diag_row_mul_kernel, dim3(grid_for(M * N)), dim3(256), 0,
                     tn::stream_of(stream), mat, diag, M, N);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_diag_col_mul_f64(double* mat, const double* diag, int64_t M, int64_t N, void* stream) {
  TN_INIT_OR_RETURN();
  if (M == 0 || N == 0) return TN_OK;
  TN_REQUIRE(mat && diag && M > 0 && N > 0, TN_EINVAL, "tn_diag_col_mul_f64: bad argument");
  hipLaunchKernelGGL(diag_col_mul_kernel, dim3(grid_for(M * N)), dim3(256), 0,
                     tn::stream_of(stream), mat, diag, M, N);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_truncate_f64(const double* S, int64_t k, int64_t chi, double* lam, double* norm_out,
                    double* err_out, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(S && lam && norm_out && k >= 1 && chi >= 1, TN_EINVAL, "tn_truncate_f64: bad argument");
  int64_t dim = std::min(chi, k);
  hipLaunchKernelGGL(truncate_kernel, dim3(1), dim3(1024), 0, tn::stream_of(stream), S, k, dim, lam,
                     norm_out, err_out);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

}  // extern "C"
