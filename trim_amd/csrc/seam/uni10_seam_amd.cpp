// The reference's backend seam (uni10_tools.h, uni10_lapack.h) over tn_*:
// every function keeps the reference's signature and argument meaning, and
// runs its arithmetic on the device; host operands (ongpu == false) are
// staged through device buffers.
#include "uni10_seam_amd.h"

#include <cstdlib>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tn_hip.h"

namespace uni10 {
namespace {

void check(int rc, const char* what) {
  if (rc != TN_OK) {
    std::ostringstream os;
    os << what << ": " << tn_last_error() << " (code " << rc << ")";
    throw std::runtime_error(os.str());
  }
}

// A device view of `count` doubles at p: p itself when it is device memory,
// else a pooled device copy (copied in when `in`, back to p when `out`).
class Dev {
 public:
  Dev(double* p, size_t count, bool ongpu, bool in, bool out) : host_(ongpu ? nullptr : p), n_(count), out_(out) {
    if (ongpu || count == 0) {
      d_ = p;
      return;
    }
    check(tn_malloc(reinterpret_cast<void**>(&d_), count * sizeof(double)), "device staging");
    own_ = true;
    if (in) check(tn_memcpy_h2d(d_, p, count * sizeof(double), nullptr), "device staging");
  }
  ~Dev() {
    if (own_) tn_free(d_);
  }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  // copies a host operand back (call before the destructor; may throw)
  void finish() {
    if (own_ && out_) check(tn_memcpy_d2h(host_, d_, n_ * sizeof(double), nullptr), "device staging");
  }
  double* get() const { return d_; }

 private:
  double* d_ = nullptr;
  double* host_;
  size_t n_;
  bool out_;
  bool own_ = false;
};

double read_scalar(const double* dev) {
  double v = 0.0;
  check(tn_memcpy_d2h(&v, dev, sizeof(double), nullptr), "scalar read");
  return v;
}

}  // namespace

// ---- memory -----------------------------------------------------------------
void* elemAlloc(size_t memsize, bool& ongpu) {
  void* p = nullptr;
  check(tn_malloc(&p, memsize ? memsize : 8), "elemAlloc");
  ongpu = true;
  return p;
}

void* elemAllocForce(size_t memsize, bool ongpu) {
  if (!ongpu) {
    void* p = std::malloc(memsize ? memsize : 8);
    if (!p) throw std::runtime_error("elemAllocForce: host allocation failed");
    return p;
  }
  void* p = nullptr;
  check(tn_malloc(&p, memsize ? memsize : 8), "elemAllocForce");
  return p;
}

void* elemCopy(void* des, const void* src, size_t memsize, bool des_ongpu, bool src_ongpu) {
  if (memsize == 0) return des;
  if (des_ongpu && src_ongpu)
    check(tn_memcpy_d2d(des, src, memsize, nullptr), "elemCopy");
  else if (des_ongpu)
    check(tn_memcpy_h2d(des, src, memsize, nullptr), "elemCopy");
  else if (src_ongpu)
    check(tn_memcpy_d2h(des, src, memsize, nullptr), "elemCopy");
  else
    std::memcpy(des, src, memsize);
  return des;
}

void elemFree(void* ptr, size_t /*memsize*/, bool ongpu) {
  if (!ptr) return;
  if (ongpu) {
    check(tn_device_sync(), "elemFree");   // queued work may still use the block
    check(tn_free(ptr), "elemFree");
  } else {
    std::free(ptr);
  }
}

void elemBzero(void* ptr, size_t memsize, bool ongpu) {
  if (memsize == 0) return;
  if (ongpu)
    check(tn_zero(ptr, memsize, nullptr), "elemBzero");
  else
    std::memset(ptr, 0, memsize);
}

// ---- diagonal copies, reshape, element access ----------------------------------
void setDiag(double* elem, double* diag_elem, size_t M, size_t N, size_t diag_N, bool ongpu, bool diag_ongpu) {
  size_t n = M < N ? M : N;
  n = n < diag_N ? n : diag_N;
  if (n == 0) return;
  Dev e(elem, (n - 1) * N + n, ongpu, true, true), d(diag_elem, n, diag_ongpu, true, false);
  check(tn_copy_strided_f64(e.get(), (int64_t)N + 1, d.get(), 1, (int64_t)n, nullptr), "setDiag");
  e.finish();
}

void getDiag(double* elem, double* diag_elem, size_t M, size_t N, size_t diag_N, bool ongpu, bool diag_ongpu) {
  size_t n = M < N ? M : N;
  n = n < diag_N ? n : diag_N;
  if (n == 0) return;
  Dev e(elem, (n - 1) * N + n, ongpu, true, false), d(diag_elem, n, diag_ongpu, false, true);
  check(tn_copy_strided_f64(d.get(), 1, e.get(), (int64_t)N + 1, (int64_t)n, nullptr), "getDiag");
  d.finish();
}

void reshapeElem(double* oldElem, int bondNum, size_t elemNum, size_t* offset, double* newElem) {
  // the reference's GPU entry: both buffers are device memory
  check(tn_reshape_elem_f64(oldElem, bondNum, elemNum, offset, newElem, nullptr), "reshapeElem");
}

double getElemAt(size_t idx, double* elem, bool ongpu) {
  if (!ongpu) return elem[idx];
  return read_scalar(elem + idx);
}

void setElemAt(size_t idx, double val, double* elem, bool ongpu) {
  if (!ongpu) {
    elem[idx] = val;
    return;
  }
  check(tn_memcpy_h2d(elem + idx, &val, sizeof(double), nullptr), "setElemAt");
}

// ---- numerics -----------------------------------------------------------------
void matrixMul(double* A, double* B, int M, int N, int K, double* C, bool ongpuA, bool ongpuB, bool ongpuC) {
  Dev a(A, (size_t)M * K, ongpuA, true, false), b(B, (size_t)K * N, ongpuB, true, false);
  Dev c(C, (size_t)M * N, ongpuC, false, true);
  if ((size_t)M * N == 0) return;
  check(tn_matmul_f64(a.get(), b.get(), M, N, K, c.get(), TN_GEMM_F64, nullptr), "matrixMul");
  c.finish();
}

void vectorAdd(double* Y, double* X, size_t N, bool y_ongpu, bool x_ongpu) {
  Dev y(Y, N, y_ongpu, true, true), x(X, N, x_ongpu, true, false);
  check(tn_axpy_f64(1.0, x.get(), y.get(), (int64_t)N, nullptr), "vectorAdd");
  y.finish();
}

void vectorScal(double a, double* X, size_t N, bool ongpu) {
  Dev x(X, N, ongpu, true, true);
  if (N) check(tn_scale_f64(x.get(), (int64_t)N, a, nullptr), "vectorScal");
  x.finish();
}

void vectorMul(double* Y, double* X, size_t N, bool y_ongpu, bool x_ongpu) {
  Dev y(Y, N, y_ongpu, true, true), x(X, N, x_ongpu, true, false);
  // row i of the N x 1 matrix Y times X_i: the elementwise product
  if (N) check(tn_diag_row_mul_f64(y.get(), x.get(), (int64_t)N, 1, nullptr), "vectorMul");
  y.finish();
}

double vectorSum(double* X, size_t N, int inc, bool ongpu) {
  if (N == 0) return 0.0;
  Dev x(X, (N - 1) * (size_t)inc + 1, ongpu, true, false);
  double* out = nullptr;
  check(tn_malloc(reinterpret_cast<void**>(&out), sizeof(double)), "vectorSum");
  int rc = tn_sum_f64(x.get(), (int64_t)N, inc, out, nullptr);
  double v = 0.0;
  if (rc == TN_OK) rc = tn_memcpy_d2h(&v, out, sizeof(double), nullptr);
  tn_free(out);
  check(rc, "vectorSum");
  return v;
}

double vectorNorm(double* X, size_t N, int inc, bool ongpu) {
  if (N == 0) return 0.0;
  Dev x(X, (N - 1) * (size_t)inc + 1, ongpu, true, false);
  double* buf = nullptr;
  const double* src = x.get();
  check(tn_malloc(reinterpret_cast<void**>(&buf), sizeof(double) * (inc == 1 ? 1 : N + 1)), "vectorNorm");
  int rc = TN_OK;
  if (inc != 1) {   // gather the strided entries first
    rc = tn_copy_strided_f64(buf + 1, 1, x.get(), inc, (int64_t)N, nullptr);
    src = buf + 1;
  }
  double v = 0.0;
  if (rc == TN_OK) rc = tn_norm2_f64(src, (int64_t)N, buf, nullptr);
  if (rc == TN_OK) rc = tn_memcpy_d2h(&v, buf, sizeof(double), nullptr);
  tn_free(buf);
  check(rc, "vectorNorm");
  return v;
}

void diagRowMul(double* mat, double* diag, size_t M, size_t N, bool mat_ongpu, bool diag_ongpu) {
  Dev m(mat, M * N, mat_ongpu, true, true), d(diag, M, diag_ongpu, true, false);
  if (M * N) check(tn_diag_row_mul_f64(m.get(), d.get(), (int64_t)M, (int64_t)N, nullptr), "diagRowMul");
  m.finish();
}

void diagColMul(double* mat, double* diag, size_t M, size_t N, bool mat_ongpu, bool diag_ongpu) {
  Dev m(mat, M * N, mat_ongpu, true, true), d(diag, N, diag_ongpu, true, false);
  if (M * N) check(tn_diag_col_mul_f64(m.get(), d.get(), (int64_t)M, (int64_t)N, nullptr), "diagColMul");
  m.finish();
}

void eigSyDecompose(double* Kij, int N, double* Eig, double* EigVec, bool ongpu) {
  Dev k(Kij, (size_t)N * N, ongpu, true, false), e(Eig, (size_t)N, ongpu, false, true);
  Dev v(EigVec, (size_t)N * N, ongpu, false, true);
  check(tn_eigh_f64(k.get(), N, e.get(), v.get(), nullptr), "eigSyDecompose");
  e.finish();
  v.finish();
}

void matrixSVD(double* Mij_ori, int M, int N, double* U, double* S, double* vT, bool ongpu) {
  const size_t k = (size_t)(M < N ? M : N);
  Dev a(Mij_ori, (size_t)M * N, ongpu, true, false), s(S, k, ongpu, false, true);
  // U and vT may be NULL when the caller wants only some factors
  Dev u(U, U ? (size_t)M * k : 0, ongpu, false, true);
  double* vt_tmp = nullptr;
  if (!vT) check(tn_malloc(reinterpret_cast<void**>(&vt_tmp), sizeof(double) * (k * N + 1)), "matrixSVD");
  {
    Dev vt(vT ? vT : vt_tmp, vT ? k * N : 0, vT ? ongpu : true, false, true);
    const int rc = tn_svd_f64(a.get(), M, N, U ? u.get() : nullptr, s.get(), vt.get(), nullptr);
    if (rc != TN_OK) tn_free(vt_tmp);
    check(rc, "matrixSVD");
    vt.finish();
  }
  tn_free(vt_tmp);
  s.finish();
  u.finish();
}

void setTranspose(double* A, size_t M, size_t N, double* AT, bool ongpu, bool ongpuT) {
  Dev a(A, M * N, ongpu, true, false), at(AT, M * N, ongpuT, false, true);
  if (M * N) check(tn_transpose_f64(a.get(), (int64_t)M, (int64_t)N, at.get(), nullptr), "setTranspose");
  at.finish();
}

void setTranspose(double* A, size_t M, size_t N, bool ongpu) {
  if (M * N == 0) return;
  Dev a(A, M * N, ongpu, true, true);
  double* t = nullptr;
  check(tn_malloc(reinterpret_cast<void**>(&t), sizeof(double) * M * N), "setTranspose");
  int rc = tn_transpose_f64(a.get(), (int64_t)M, (int64_t)N, t, nullptr);
  if (rc == TN_OK) rc = tn_memcpy_d2d(a.get(), t, sizeof(double) * M * N, nullptr);
  if (rc == TN_OK) rc = tn_device_sync();
  tn_free(t);
  check(rc, "setTranspose");
  a.finish();
}

void setIdentity(double* elem, size_t M, size_t N, bool ongpu) {
  if (M * N == 0) return;
  Dev e(elem, M * N, ongpu, false, true);
  const size_t n = M < N ? M : N;
  std::vector<double> ones(n, 1.0);
  double* d = nullptr;
  check(tn_malloc(reinterpret_cast<void**>(&d), sizeof(double) * n), "setIdentity");
  int rc = tn_zero(e.get(), sizeof(double) * M * N, nullptr);
  if (rc == TN_OK) rc = tn_memcpy_h2d(d, ones.data(), sizeof(double) * n, nullptr);
  if (rc == TN_OK) rc = tn_copy_strided_f64(e.get(), (int64_t)N + 1, d, 1, (int64_t)n, nullptr);
  if (rc == TN_OK) rc = tn_device_sync();
  tn_free(d);
  check(rc, "setIdentity");
  e.finish();
}

bool lanczosEV(double* A, double* psi, size_t dim, size_t& max_iter, double err_tol, double& eigVal, double* eigVec,
               bool ongpu) {
  if (max_iter <= 2) throw std::runtime_error("Maximum iteration number should be set greater than 2.");
  Dev a(A, dim * dim, ongpu, true, false), p(psi, dim, ongpu, true, false), v(eigVec, dim, ongpu, false, true);
  // the solver works in place on its start vector: run it on the output
  check(tn_memcpy_d2d(v.get(), p.get(), dim * sizeof(double), nullptr), "lanczosEV");
  int64_t iters = (int64_t)max_iter;
  int converged = 0;
  check(tn_lanczos_f64(a.get(), (int64_t)dim, v.get(), &iters, err_tol, &eigVal, &converged, nullptr), "lanczosEV");
  max_iter = (size_t)iters;
  v.finish();
  return converged != 0;
}

}  // namespace uni10
