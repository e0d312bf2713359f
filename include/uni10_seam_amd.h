/*
 * uni10_seam_amd.h - the backend seam of Uni10 with the reference's own names
 * and signatures, implemented over the MI355X kernel library (tn_hip.h).
 *
 * A maintainer who keeps the reference's C++ classes (UniTensor, Block,
 * Matrix, Network) and replaces only their backend links
 * trim_amd/libuni10_seam_amd.so in place of the reference's
 * uni10_tools_gpu.cu / uni10_lapack_gpu.cu objects.  Declarations follow
 * uni10/src/uni10/tools/uni10_tools.h:56-135 (memory, reshapeElem, diagonal
 * copies, element access) and uni10/src/uni10/numeric/lapack/uni10_lapack.h
 * :54-83 (matrixMul, vector operations, diagonal scaling, eigSyDecompose,
 * matrixSVD, setTranspose, setIdentity, lanczosEV).
 *
 * Semantics:
 *   - `ongpu` flags keep the reference's meaning: a pointer flagged true is
 *     device memory, false is host memory.  elemAlloc always returns device
 *     memory and sets ongpu = true (the reference's GPU build does the same
 *     when the device has room).  Host operands are staged through device
 *     memory: every computation runs on the device, never on the host.
 *   - Errors throw std::runtime_error with the library's message (the
 *     reference throws runtime_error/logic_error and its CUDA helpers exit()).
 *   - Work is enqueued on the library's default stream; functions that
 *     return a value or touch host memory synchronise.
 */
#ifndef UNI10_SEAM_AMD_H
#define UNI10_SEAM_AMD_H

#include <cstddef>

namespace uni10 {

// ---- memory (uni10_tools.h:91-101) ----------------------------------------
void* elemAlloc(size_t memsize, bool& ongpu);
void* elemAllocForce(size_t memsize, bool ongpu);
void* elemCopy(void* des, const void* src, size_t memsize, bool des_ongpu, bool src_ongpu);
void elemFree(void* ptr, size_t memsize, bool ongpu);
void elemBzero(void* ptr, size_t memsize, bool ongpu);

// ---- diagonal copies, reshape, element access (uni10_tools.h:107-119) ------
void setDiag(double* elem, double* diag_elem, size_t M, size_t N, size_t diag_N, bool ongpu, bool diag_ongpu);
void getDiag(double* elem, double* diag_elem, size_t M, size_t N, size_t diag_N, bool ongpu, bool diag_ongpu);
void reshapeElem(double* oldElem, int bondNum, size_t elemNum, size_t* offset, double* newElem);
double getElemAt(size_t idx, double* elem, bool ongpu);
void setElemAt(size_t idx, double val, double* elem, bool ongpu);

// ---- numerics (uni10_lapack.h:54-83) ---------------------------------------
void matrixMul(double* A, double* B, int M, int N, int K, double* C, bool ongpuA, bool ongpuB, bool ongpuC);
void vectorAdd(double* Y, double* X, size_t N, bool y_ongpu, bool x_ongpu);
void vectorScal(double a, double* X, size_t N, bool ongpu);
void vectorMul(double* Y, double* X, size_t N, bool y_ongpu, bool x_ongpu);
double vectorSum(double* X, size_t N, int inc, bool ongpu);
double vectorNorm(double* X, size_t N, int inc, bool ongpu);
void diagRowMul(double* mat, double* diag, size_t M, size_t N, bool mat_ongpu, bool diag_ongpu);
void diagColMul(double* mat, double* diag, size_t M, size_t N, bool mat_ongpu, bool diag_ongpu);
void eigSyDecompose(double* Kij, int N, double* Eig, double* EigVec, bool ongpu);
void matrixSVD(double* Mij_ori, int M, int N, double* U, double* S, double* vT, bool ongpu);
void setTranspose(double* A, size_t M, size_t N, double* AT, bool ongpu, bool ongpuT);
void setTranspose(double* A, size_t M, size_t N, bool ongpu);
void setIdentity(double* elem, size_t M, size_t N, bool ongpu);
bool lanczosEV(double* A, double* psi, size_t dim, size_t& max_iter, double err_tol, double& eigVal, double* eigVec,
               bool ongpu);

}  // namespace uni10

#endif  // UNI10_SEAM_AMD_H
