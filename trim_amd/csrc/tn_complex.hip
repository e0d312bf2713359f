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

}  // namespace

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
