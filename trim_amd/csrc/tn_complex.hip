// Complex FP64 (interleaved re, im pairs) dense permute and matrix product,
// the first pieces of the CTYPE path (UniTensorComplex.cpp permute,
// matrixMul on std::complex<double> in uni10_lapack_cpu.cpp).
//
// Both reuse the real kernels.  A complex tensor of dims (d0..d_{r-1}) is
// the real tensor (d0..d_{r-1}, 2): its permute is the real permute with the
// pair axis kept last (a bit-exact copy).  C = A B with A (M x K), B (K x N)
// complex is one real product of M x 2K by the 2K x 2N expansion
//   [ Br  Bi ]
//   [ -Bi Br ]   (per element: rows 2k, 2k+1; columns 2n, 2n+1),
// whose output rows are already the interleaved complex rows of C:
// C[m][2n] = sum_k Ar Br - Ai Bi, C[m][2n+1] = sum_k Ar Bi + Ai Br.  That is
// 8 M N K real flops, the cost of the complex product itself.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "tn_internal.h"

namespace {

__global__ void expand_b_kernel(const double* __restrict__ B, int64_t K, int64_t N, double* __restrict__ X) {
  const int64_t total = K * N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / N, n = e - (e / N) * N;
    const double br = B[2 * e], bi = B[2 * e + 1];
    double* r0 = X + (2 * k) * (2 * N) + 2 * n;
    double* r1 = r0 + 2 * N;
    r0[0] = br;
    r0[1] = bi;
    r1[0] = -bi;
    r1[1] = br;
  }
}

// c[i] = conj(a) b over the m rows of two columns (strided by the lanes of
// one wave, then a butterfly sum), and the squared norms of both columns.
__device__ inline void wave_sum3(double& x, double& y, double& z, double& w) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    x += __shfl_xor(x, o);
    y += __shfl_xor(y, o);
    z += __shfl_xor(z, o);
    w += __shfl_xor(w, o);
  }
}

__device__ inline void rr_pair_c(int r, int k, int np, int& p, int& q) {
  // round-robin pairing of np positions (np even): round r, pair k
  const int L = np - 1;
  if (k == 0) {
    p = r;
    q = L;
  } else {
    p = (r + k) % L;
    q = (r - k + L) % L;
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

// One-sided (Hestenes) Jacobi SVD of a complex m x n working matrix W
// (m >= n), columns contiguous (column j at W + 2 m j, interleaved re/im),
// with the rotations accumulated in V (n x n, column j at V + 2 n j).  For
// the pair (p, q): alpha = |w_p|^2, beta = |w_q|^2, gamma = w_p^H w_q =
// |gamma| e^{i phi}; with w~ = e^{-i phi} w_q the pair is a real one, and the
// real rotation (c, s) of the real Jacobi step orthogonalises it:
//   w_p' = c w_p - s e^{-i phi} w_q,  w_q' = s e^{i phi} w_p + c w_q.
// One workgroup per matrix; each wave takes pairs of the round in turn.
// Load: W = A (row-major M x N) or W = A^H (ctr, when M < N); `shift`
// (device scalar or NULL) is added to the diagonal (the Hermitian eigen
// solver's positive shift).
__global__ void __launch_bounds__(256) svd_c128_kernel(const double* __restrict__ A, int64_t M, int64_t N, int ctr,
                                                       const double* __restrict__ shift, double* __restrict__ W,
                                                       double* __restrict__ V, int max_sweeps,
                                                       int* __restrict__ status) {
  const int64_t m = ctr ? N : M, n = ctr ? M : N;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double sh = shift ? *shift : 0.0;
  for (int64_t e = tid; e < m * n; e += 256) {
    const int64_t i = e % m, j = e / m;            // W(i, j)
    double re, im;
    if (ctr) {                                      // A^H(i, j) = conj(A[j][i])
      re = A[2 * (j * N + i)];
      im = -A[2 * (j * N + i) + 1];
    } else {
      re = A[2 * (i * N + j)];
      im = A[2 * (i * N + j) + 1];
    }
    if (i == j) re += sh;
    W[2 * (j * m + i)] = re;
    W[2 * (j * m + i) + 1] = im;
  }
  for (int64_t e = tid; e < n * n; e += 256) {
    V[2 * e] = (e % n == e / n) ? 1.0 : 0.0;
    V[2 * e + 1] = 0.0;
  }
  __syncthreads();
  const double tol = 10.0 * DBL_EPSILON * (double)m;
  const int np = (int)((n % 2) ? n + 1 : n);
  const int npairs = np / 2;
  int sweep = 0;
  for (; sweep < max_sweeps && n > 1; ++sweep) {
    int rotated = 0;
    for (int r = 0; r < np - 1; ++r) {
      for (int k = wave; k < npairs; k += 4) {
        int p, q;
        rr_pair_c(r, k, np, p, q);
        if (q >= n) continue;
        double* wp = W + 2 * (int64_t)p * m;
        double* wq = W + 2 * (int64_t)q * m;
        double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;
        for (int64_t i = lane; i < m; i += 64) {
          const double xr = wp[2 * i], xi = wp[2 * i + 1], yr = wq[2 * i], yi = wq[2 * i + 1];
          a = fma(xr, xr, fma(xi, xi, a));
          b = fma(yr, yr, fma(yi, yi, b));
          gr = fma(xr, yr, fma(xi, yi, gr));        // Re(conj(x) y)
          gi = fma(xr, yi, fma(-xi, yr, gi));       // Im(conj(x) y)
        }
        wave_sum3(a, b, gr, gi);
        const double g = sqrt(gr * gr + gi * gi);
        if (g == 0.0 || !(g > tol * sqrt(a * b))) continue;
        rotated = 1;
        const double zeta = (b - a) / (2.0 * g);
        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        const double er = gr / g, ei = gi / g;     // e^{i phi}
        for (int64_t i = lane; i < m; i += 64) {
          const double xr = wp[2 * i], xi = wp[2 * i + 1], yr = wq[2 * i], yi = wq[2 * i + 1];
          // e^{-i phi} y = (er yr + ei yi) + i (er yi - ei yr)
          const double ur = er * yr + ei * yi, ui = er * yi - ei * yr;
          // e^{i phi} x = (er xr - ei xi) + i (er xi + ei xr)
          const double vr = er * xr - ei * xi, vi = er * xi + ei * xr;
          wp[2 * i] = c * xr - s * ur;
          wp[2 * i + 1] = c * xi - s * ui;
          wq[2 * i] = s * vr + c * yr;
          wq[2 * i + 1] = s * vi + c * yi;
        }
        double* vp = V + 2 * (int64_t)p * n;
        double* vq = V + 2 * (int64_t)q * n;
        for (int64_t i = lane; i < n; i += 64) {
          const double xr = vp[2 * i], xi = vp[2 * i + 1], yr = vq[2 * i], yi = vq[2 * i + 1];
          const double ur = er * yr + ei * yi, ui = er * yi - ei * yr;
          const double vr = er * xr - ei * xi, vi = er * xi + ei * xr;
          vp[2 * i] = c * xr - s * ur;
          vp[2 * i + 1] = c * xi - s * ui;
          vq[2 * i] = s * vr + c * yr;
          vq[2 * i + 1] = s * vi + c * yi;
        }
      }
      __syncthreads();
    }
    if (!__syncthreads_or(rotated)) break;
  }
  if (tid == 0 && sweep >= max_sweeps) atomicAdd(status, 1);
}

// Column norms of W, descending rank (ties by index), and the outputs.
// SVD (A = U diag(S) VT, k = min(M, N)): without ctr U[:, r] = w_j / s_j and
// VT[r, :] = v_j^H; with ctr (W = A^H) U[:, r] = v_j and VT[r, :] = (w_j /
// s_j)^H.  A zero column leaves zero vectors.  Hermitian eigen solver
// (eig != NULL): eigenvalue s_j - shift, ascending, and eigenvector row
// conj(v_j) (the rows E with H = E^H diag(eig) E, as zheev's column-major
// vectors read back row-major).
__global__ void __launch_bounds__(256) svd_c128_finish(const double* __restrict__ W, const double* __restrict__ V,
                                                       int64_t M, int64_t N, int ctr, double* __restrict__ U,
                                                       double* __restrict__ S, double* __restrict__ VT,
                                                       const double* __restrict__ shift, double* __restrict__ eig,
                                                       double* __restrict__ vec_rows, double* __restrict__ nrm,
                                                       int* __restrict__ rank) {
  const int64_t m = ctr ? N : M, n = ctr ? M : N;
  const int tid = threadIdx.x;
  for (int64_t j = tid; j < n; j += 256) {
    double s = 0.0;
    const double* w = W + 2 * j * m;
    for (int64_t i = 0; i < m; ++i) s = fma(w[2 * i], w[2 * i], fma(w[2 * i + 1], w[2 * i + 1], s));
    nrm[j] = sqrt(s);
  }
  __syncthreads();
  for (int64_t j = tid; j < n; j += 256) {
    int64_t rk = 0;
    for (int64_t i = 0; i < n; ++i) rk += (nrm[i] > nrm[j]) || (nrm[i] == nrm[j] && i < j);
    rank[j] = (int)rk;
  }
  __syncthreads();
  if (eig) {
    const double sh = *shift;
    for (int64_t j = tid; j < n; j += 256) {
      const int64_t r = n - 1 - rank[j];                // ascending
      eig[r] = nrm[j] - sh;
    }
    for (int64_t e = tid; e < n * n; e += 256) {
      const int64_t j = e / n, i = e % n;
      const int64_t r = n - 1 - rank[j];
      vec_rows[2 * (r * n + i)] = V[2 * (j * n + i)];
      vec_rows[2 * (r * n + i) + 1] = -V[2 * (j * n + i) + 1];
    }
    return;
  }
  for (int64_t j = tid; j < n; j += 256) S[rank[j]] = nrm[j];
  const int64_t k = n;   // = min(M, N)
  // the normalised columns of W: U (no ctr, M x k) or VT rows conjugated (ctr, k x N)
  for (int64_t e = tid; e < m * n; e += 256) {
    const int64_t j = e / m, i = e % m;
    const double s = nrm[j];
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    const double wr = W[2 * (j * m + i)] * inv, wi = W[2 * (j * m + i) + 1] * inv;
    const int64_t r = rank[j];
    if (!ctr) {
      if (U) {
        U[2 * (i * k + r)] = wr;
        U[2 * (i * k + r) + 1] = wi;
      }
    } else {
      VT[2 * (r * N + i)] = wr;
      VT[2 * (r * N + i) + 1] = -wi;
    }
  }
  for (int64_t e = tid; e < n * n; e += 256) {
    const int64_t j = e / n, i = e % n;
    const int64_t r = rank[j];
    const double vr = V[2 * (j * n + i)], vi = V[2 * (j * n + i) + 1];
    if (!ctr) {
      VT[2 * (r * N + i)] = vr;                         // VT[r][i] = conj(v_j[i])
      VT[2 * (r * N + i) + 1] = -vi;
    } else if (U) {
      U[2 * (i * k + r)] = vr;                          // U[i][r] = v_j[i]
      U[2 * (i * k + r) + 1] = vi;
    }
  }
}

__global__ void fro_shift_kernel(const double* __restrict__ A, int64_t n2, double* __restrict__ out) {
  // out = 2 ||A||_F (+ 1 when A = 0): a shift that makes A + shift I positive definite
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < n2; e += 256) s = fma(A[e], A[e], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0] > 0.0 ? 2.0 * sqrt(red[0]) : 1.0;
}

__global__ void expand_c128_kernel(const double* __restrict__ B, int64_t n, double* __restrict__ X) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double br = B[2 * e], bi = B[2 * e + 1];
    X[4 * e] = br;
    X[4 * e + 1] = bi;
    X[4 * e + 2] = -bi;
    X[4 * e + 3] = br;
  }
}

}  // namespace

// SVD and Hermitian eigen solver share the workspace and kernels.
static int svd_c128_run(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT, bool herm,
                        double* eig, double* vec_rows, hipStream_t st) {
  const int ctr = (!herm && M < N) ? 1 : 0;
  const int64_t m = ctr ? N : M, n = ctr ? M : N;
  const size_t wbytes = sizeof(double) * 2 * (size_t)(m * n + n * n) + sizeof(double) * (size_t)(n + 1) +
                        sizeof(int) * (size_t)(n + 2);
  char* ws = nullptr;
  int rc = tn::pool_alloc((void**)&ws, wbytes);
  if (rc != TN_OK) return rc;
  double* W = reinterpret_cast<double*>(ws);
  double* Vw = W + 2 * m * n;
  double* nrm = Vw + 2 * n * n;
  double* shift = nrm + n;
  int* status = reinterpret_cast<int*>(shift + 1);
  int* rank = status + 2;
  rc = tn_zero(status, sizeof(int), st);
  if (rc == TN_OK && herm) {
    hipLaunchKernelGGL(fro_shift_kernel, dim3(1), dim3(256), 0, st, A, 2 * M * N, shift);
    if (hipGetLastError() != hipSuccess) rc = TN_EHIP;
  }
  if (rc == TN_OK) {
    hipLaunchKernelGGL(svd_c128_kernel, dim3(1), dim3(256), 0, st, A, M, N, ctr, herm ? shift : nullptr, W, Vw, 60,
                       status);
    if (hipGetLastError() != hipSuccess) rc = TN_EHIP;
  }
  if (rc == TN_OK) {
    hipLaunchKernelGGL(svd_c128_finish, dim3(1), dim3(256), 0, st, W, Vw, M, N, ctr, U, S, VT, shift,
                       herm ? eig : nullptr, vec_rows, nrm, rank);
    if (hipGetLastError() != hipSuccess) rc = TN_EHIP;
  }
  if (rc == TN_EHIP) tn::set_error("complex SVD: kernel launch failed");
  int h = 0;
  if (rc == TN_OK) rc = tn_memcpy_d2h(&h, status, sizeof(int), st);
  tn::pool_free(ws);
  if (rc != TN_OK) return rc;
  TN_REQUIRE(h == 0, TN_ENOCONV, "complex %s: Jacobi did not converge in 60 sweeps", herm ? "eigh" : "svd");
  return TN_OK;
}

int tn_svd_c128(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(M >= 0 && N >= 0, TN_EINVAL, "tn_svd_c128: negative size");
  if (M == 0 || N == 0) return TN_OK;
  TN_REQUIRE(A && S && VT, TN_EINVAL, "tn_svd_c128: NULL buffer");
  TN_REQUIRE(M * N <= (int64_t(1) << 26), TN_EUNSUPP, "tn_svd_c128: %lld x %lld too large", (long long)M,
             (long long)N);
  return svd_c128_run(A, M, N, U, S, VT, false, nullptr, nullptr, tn::stream_of(stream));
}

int tn_eigh_c128(const double* A, int64_t N, double* eig, double* vec_rows, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(N >= 0, TN_EINVAL, "tn_eigh_c128: negative size");
  if (N == 0) return TN_OK;
  TN_REQUIRE(A && eig && vec_rows, TN_EINVAL, "tn_eigh_c128: NULL buffer");
  TN_REQUIRE(N <= 4096, TN_EUNSUPP, "tn_eigh_c128: N=%lld too large", (long long)N);
  return svd_c128_run(A, N, N, nullptr, nullptr, nullptr, true, eig, vec_rows, tn::stream_of(stream));
}

int tn_expand_c128(const double* B, int64_t n, double* X, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(n >= 0 && (n == 0 || (B && X)), TN_EINVAL, "tn_expand_c128: bad argument");
  if (n == 0) return TN_OK;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(expand_c128_kernel, dim3(blocks), dim3(256), 0, tn::stream_of(stream), B, n, X);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_permute_dense_c128(const double* src, double* dst, int rank, const int64_t* dims, const int* perm,
                          void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(rank >= 0 && rank < 15 && (rank == 0 || (dims && perm)), TN_EINVAL,
             "tn_permute_dense_c128: rank %d", rank);
  int64_t d2[16];
  int p2[16];
  for (int i = 0; i < rank; ++i) {
    d2[i] = dims[i];
    p2[i] = perm[i];
  }
  d2[rank] = 2;
  p2[rank] = rank;
  return tn_permute_dense_f64(src, dst, rank + 1, d2, p2, stream);
}

int tn_matmul_c128(const double* A, const double* B, int64_t M, int64_t N, int64_t K, double* C, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(M >= 0 && N >= 0 && K >= 0, TN_EINVAL, "tn_matmul_c128: negative size");
  if (M == 0 || N == 0) return TN_OK;
  TN_REQUIRE(C != nullptr, TN_EINVAL, "tn_matmul_c128: NULL C");
  if (K == 0) {
    const int rc = tn_zero(C, sizeof(double) * 2 * (size_t)M * (size_t)N, stream);
    if (rc != TN_OK) return rc;
    return tn_stream_sync(stream);
  }
  TN_REQUIRE(A && B, TN_EINVAL, "tn_matmul_c128: NULL operand");
  double* X = nullptr;
  int rc = tn::pool_alloc((void**)&X, sizeof(double) * 4 * (size_t)K * (size_t)N);
  if (rc != TN_OK) return rc;
  hipStream_t st = tn::stream_of(stream);
  const unsigned blocks = (unsigned)std::min<int64_t>((K * N + 255) / 256, 8192);
  hipLaunchKernelGGL(expand_b_kernel, dim3(blocks), dim3(256), 0, st, B, K, N, X);
  if (hipGetLastError() != hipSuccess) {
    tn::pool_free(X);
    tn::set_error("tn_matmul_c128: launch failed");
    return TN_EHIP;
  }
  rc = tn_matmul_f64(A, X, M, 2 * N, 2 * K, C, TN_GEMM_F64, stream);   // synchronises the stream
  if (rc != TN_OK) (void)hipStreamSynchronize(st);
  tn::pool_free(X);
  return rc;
}
