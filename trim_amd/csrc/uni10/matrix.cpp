// Matrix, device buffers and the thread-local execution state.
// Follows uni10/src/uni10/tensor-network/lib/MatrixReal.cpp (init :35-67,
// resize :289-359, transpose :257-272), Matrix.cpp (operator*= :71-97),
// MatrixTools.cpp (exph :109-121), data-structure/lib/BlockReal.cpp (svd
// :226-250, norm :252-261, eigh :359-391) and BlockTools.cpp (RDotR :38-76),
// with every element operation done by the kernels behind tn_hip.h.
#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>

#include "common.h"

namespace uni10 {

namespace {
thread_local int t_gemm_mode = GEMM_F64;
thread_local void* t_stream = nullptr;
}  // namespace

void setGemmMode(GemmMode m) { t_gemm_mode = (int)m; }
GemmMode gemmMode() { return (GemmMode)t_gemm_mode; }
void setStream(void* s) {
  t_stream = s;
  tn_pool_set_stream(s);   // temporaries freed by this thread are ordered on it
}
void* currentStream() { return t_stream; }
void synchronize() { detail::check(tn_stream_sync(t_stream), "synchronize"); }

namespace detail {
void* stream() { return t_stream; }
int gemm_mode() { return t_gemm_mode; }
void sync() { check(tn_stream_sync(t_stream), "tn_stream_sync"); }
}  // namespace detail

// ---------------------------------------------------------------------------
DevBuf::DevBuf(size_t n, bool zero) { reset(n, zero); }

DevBuf::~DevBuf() {
  if (p_) tn_free(p_);
}

DevBuf::DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
  o.p_ = nullptr;
  o.n_ = 0;
}

DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    if (p_) tn_free(p_);
    p_ = o.p_;
    n_ = o.n_;
    o.p_ = nullptr;
    o.n_ = 0;
  }
  return *this;
}

void DevBuf::reset(size_t n, bool zero) {
  if (p_) tn_free(p_);
  p_ = nullptr;
  n_ = n;
  void* q = nullptr;
  detail::check(tn_malloc(&q, std::max<size_t>(n, 1) * sizeof(double)), "tn_malloc");
  p_ = (double*)q;
  if (zero) detail::check(tn_zero(p_, std::max<size_t>(n, 1) * sizeof(double), detail::stream()), "tn_zero");
}

// ---------------------------------------------------------------------------
Matrix::Matrix() : Rnum(0), Cnum(0), diag(false) {}

Matrix::Matrix(size_t R, size_t C, bool d, bool) : Rnum(R), Cnum(C), diag(d) {
  buf.reset(elemNum(), true);
}

Matrix::Matrix(size_t R, size_t C, const double* elem, bool d, bool src_ongpu)
    : Rnum(R), Cnum(C), diag(d) {
  buf.reset(elemNum(), false);
  setElem(elem, src_ongpu);
}

Matrix::Matrix(uninit_t, size_t R, size_t C, bool d) : Rnum(R), Cnum(C), diag(d) {
  buf.reset(elemNum(), false);
}

Matrix::Matrix(cflag, size_t R, size_t C, bool d) : Rnum(R), Cnum(C), diag(d), ctype(true) {
  buf.reset(nstore(), true);
}

Matrix::Matrix(cflag, size_t R, size_t C, const std::complex<double>* elem, bool d)
    : Rnum(R), Cnum(C), diag(d), ctype(true) {
  buf.reset(nstore(), false);
  setElem(reinterpret_cast<const double*>(elem), false);
}

Matrix::Matrix(const Matrix& m) : Rnum(m.Rnum), Cnum(m.Cnum), diag(m.diag), ctype(m.ctype) {
  buf.reset(nstore(), false);
  if (nstore())
    detail::check(tn_memcpy_d2d(buf.get(), m.buf.get(), nstore() * sizeof(double), detail::stream()),
                  "Matrix copy");
}

Matrix::Matrix(Matrix&& m) noexcept
    : Rnum(m.Rnum), Cnum(m.Cnum), diag(m.diag), ctype(m.ctype), buf(std::move(m.buf)) {
  m.Rnum = m.Cnum = 0;
}

Matrix& Matrix::operator=(const Matrix& m) {
  if (this == &m) return *this;
  Rnum = m.Rnum;
  Cnum = m.Cnum;
  diag = m.diag;
  ctype = m.ctype;
  buf.reset(nstore(), false);
  if (nstore())
    detail::check(tn_memcpy_d2d(buf.get(), m.buf.get(), nstore() * sizeof(double), detail::stream()),
                  "Matrix assign");
  return *this;
}

Matrix& Matrix::operator=(Matrix&& m) noexcept {
  Rnum = m.Rnum;
  Cnum = m.Cnum;
  diag = m.diag;
  ctype = m.ctype;
  buf = std::move(m.buf);
  return *this;
}

void Matrix::real_only(const char* what) const {
  if (ctype) detail::fail(std::string(what) + ": not provided for complex matrices");
}

std::vector<std::complex<double>> Matrix::getHostElemC() const {
  if (!ctype) detail::fail("Matrix::getHostElemC: the matrix is real");
  std::vector<std::complex<double>> h(elemNum());
  if (!h.empty())
    detail::check(tn_memcpy_d2h(h.data(), buf.get(), h.size() * sizeof(std::complex<double>), detail::stream()),
                  "Matrix::getHostElemC");
  return h;
}

std::vector<double> Matrix::getHostElem() const {
  std::vector<double> h(nstore());
  if (!h.empty())
    detail::check(tn_memcpy_d2h(h.data(), buf.get(), h.size() * sizeof(double), detail::stream()),
                  "Matrix::getHostElem");
  return h;
}

void Matrix::setElem(const double* elem, bool src_ongpu) {
  if (!elemNum()) return;
  int rc = src_ongpu ? tn_memcpy_d2d(buf.get(), elem, nstore() * sizeof(double), detail::stream())
                     : tn_memcpy_h2d(buf.get(), elem, nstore() * sizeof(double), detail::stream());
  detail::check(rc, "Matrix::setElem");
}

void Matrix::setElem(const std::vector<double>& elem, bool src_ongpu) {
  if (elem.size() < nstore()) detail::fail("Matrix::setElem: not enough elements");
  setElem(elem.data(), src_ongpu);
}

double Matrix::operator[](size_t idx) const {
  real_only("Matrix::operator[]");
  if (idx >= elemNum()) detail::fail("Matrix::operator[]: index out of range");
  double v = 0;
  detail::check(tn_memcpy_d2h(&v, buf.get() + idx, sizeof(double), detail::stream()), "Matrix::[]");
  return v;
}

double Matrix::at(size_t r, size_t c) const {
  if (r >= Rnum || c >= Cnum) detail::fail("The input indices are out of range.");
  if (diag) return (r == c) ? (*this)[r] : 0.0;
  return (*this)[r * Cnum + c];
}

void Matrix::set_zero() {
  if (elemNum()) detail::check(tn_zero(buf.get(), nstore() * sizeof(double), detail::stream()), "set_zero");
}

void Matrix::identity() {
  real_only("Matrix::identity");
  std::vector<double> h(elemNum(), 0.0);
  if (diag)
    std::fill(h.begin(), h.end(), 1.0);
  else
    for (size_t i = 0; i < std::min(Rnum, Cnum); ++i) h[i * Cnum + i] = 1.0;
  setElem(h);
}

void Matrix::randomize(uint64_t seed) {
  detail::check(tn_fill_uniform_f64(buf.get(), (int64_t)nstore(), seed, detail::stream()), "randomize");
}

Matrix& Matrix::resize(size_t row, size_t col) {
  // MatrixReal.cpp:289-359
  real_only("Matrix::resize");
  if (diag) {
    size_t n_new = std::min(row, col), n_old = elemNum();
    if (n_new != n_old) {
      DevBuf nb(n_new, true);
      size_t keep = std::min(n_new, n_old);
      if (keep)
        detail::check(tn_memcpy_d2d(nb.get(), buf.get(), keep * sizeof(double), detail::stream()), "resize");
      buf = std::move(nb);
    }
    Rnum = row;
    Cnum = col;
    return *this;
  }
  DevBuf nb(row * col, true);
  size_t dr = std::min(row, Rnum), dc = std::min(col, Cnum);
  if (col == Cnum) {
    if (dr * dc)
      detail::check(tn_memcpy_d2d(nb.get(), buf.get(), dr * dc * sizeof(double), detail::stream()), "resize");
  } else {
    for (size_t r = 0; r < dr; ++r)
      detail::check(tn_memcpy_d2d(nb.get() + r * col, buf.get() + r * Cnum, dc * sizeof(double),
                                  detail::stream()),
                    "resize");
  }
  buf = std::move(nb);
  Rnum = row;
  Cnum = col;
  return *this;
}

Matrix& Matrix::transpose() {
  real_only("Matrix::transpose");
  if (!diag && Rnum * Cnum) {
    DevBuf nb(Rnum * Cnum);
    detail::check(tn_transpose_f64(buf.get(), (int64_t)Rnum, (int64_t)Cnum, nb.get(), detail::stream()),
                  "Matrix::transpose");
    buf = std::move(nb);
  }
  std::swap(Rnum, Cnum);
  return *this;
}

double Matrix::norm() const {
  if (!elemNum()) return 0.0;
  DevBuf out(1);
  detail::check(tn_norm2_f64(buf.get(), (int64_t)nstore(), out.get(), detail::stream()), "Matrix::norm");
  double v = 0;
  detail::check(tn_memcpy_d2h(&v, out.get(), sizeof(double), detail::stream()), "Matrix::norm");
  return v;
}

double Matrix::sum() const {
  real_only("Matrix::sum");
  auto h = getHostElem();
  double s = 0;
  for (double v : h) s += v;
  return s;
}

double Matrix::trace() const {
  real_only("Matrix::trace");
  auto h = getHostElem();
  double s = 0;
  if (diag)
    for (double v : h) s += v;
  else
    for (size_t i = 0; i < std::min(Rnum, Cnum); ++i) s += h[i * Cnum + i];
  return s;
}

Matrix& Matrix::operator*=(double a) {
  if (elemNum())
    detail::check(tn_scale_f64(buf.get(), (int64_t)nstore(), a, detail::stream()), "Matrix::operator*=");
  return *this;
}

Matrix& Matrix::operator+=(const Matrix& m) {
  if (m.Rnum != Rnum || m.Cnum != Cnum || m.diag != diag || m.ctype != ctype)
    detail::fail("The dimensions of the two matrices do not match for matrix addition.");
  auto a = getHostElem();
  auto b = m.getHostElem();
  for (size_t i = 0; i < a.size(); ++i) a[i] += b[i];
  setElem(a);
  return *this;
}

Matrix Matrix::getDiag() const {
  real_only("Matrix::getDiag");
  if (diag) return *this;
  Matrix d(Rnum, Cnum, true);
  auto h = getHostElem();
  std::vector<double> dv(d.elemNum());
  for (size_t i = 0; i < dv.size(); ++i) dv[i] = h[i * Cnum + i];
  d.setElem(dv);
  return d;
}

static Matrix dense_of_diag(const Matrix& m) {
  Matrix d(m.row(), m.col(), false);
  auto h = m.getHostElem();
  std::vector<double> full(m.row() * m.col(), 0.0);
  for (size_t i = 0; i < h.size(); ++i) full[i * m.col() + i] = h[i];
  d.setElem(full);
  return d;
}

std::vector<Matrix> Matrix::svd() const {
  // BlockReal.cpp:226-250 -> matrixSVD: U (R x k), S (k x k diagonal), VT (k x C)
  if (ctype) {
    // BlockComplex.cpp svd -> matrixSVD on std::complex<double> (zgesvd)
    if (diag) detail::fail("Matrix::svd: a complex diagonal matrix is its own SVD");
    size_t k = std::min(Rnum, Cnum);
    std::vector<Matrix> outs;
    outs.emplace_back(CTYPE, Rnum, k, false);
    outs.emplace_back(k, k, true);
    outs.emplace_back(CTYPE, k, Cnum, false);
    if (k)
      detail::check(tn_svd_c128(buf.get(), (int64_t)Rnum, (int64_t)Cnum, outs[0].buf.get(), outs[1].buf.get(),
                                outs[2].buf.get(), detail::stream()),
                    "Matrix::svd");
    return outs;
  }
  if (diag) return dense_of_diag(*this).svd();
  size_t k = std::min(Rnum, Cnum);
  std::vector<Matrix> outs;
  outs.emplace_back(Rnum, k, false);
  outs.emplace_back(k, k, true);
  outs.emplace_back(k, Cnum, false);
  if (k)
    detail::check(tn_svd_f64(buf.get(), (int64_t)Rnum, (int64_t)Cnum, outs[0].buf.get(), outs[1].buf.get(),
                             outs[2].buf.get(), detail::stream()),
                  "Matrix::svd");
  return outs;
}

std::vector<Matrix> Matrix::eigh() const {
  // BlockReal.cpp:359-391 -> eigSyDecompose: [eigenvalues (diagonal), eigenvectors as rows]
  if (Rnum != Cnum) detail::fail("Cannot perform eigenvalue decomposition on a non-square matrix.");
  if (diag) detail::fail("Cannot perform eigenvalue decomposition on a diagonal matrix. Need not to do so.");
  std::vector<Matrix> outs;
  if (ctype) {
    // eigSyDecompose on std::complex<double> (zheev): real eigenvalues
    outs.emplace_back(Rnum, Cnum, true);
    outs.emplace_back(CTYPE, Rnum, Cnum, false);
    detail::check(tn_eigh_c128(buf.get(), (int64_t)Rnum, outs[0].buf.get(), outs[1].buf.get(), detail::stream()),
                  "Matrix::eigh");
    return outs;
  }
  outs.emplace_back(Rnum, Cnum, true);
  outs.emplace_back(Rnum, Cnum, false);
  detail::check(tn_eigh_f64(buf.get(), (int64_t)Rnum, outs[0].buf.get(), outs[1].buf.get(), detail::stream()),
                "Matrix::eigh");
  return outs;
}

Matrix RDotR(const Matrix& a, const Matrix& b) {
  if (a.typeID() != 1 || b.typeID() != 1) detail::fail("matrix product: complex matrices are not provided");
  if (a.col() != b.row())
    detail::fail("The dimensions of the two matrices do not match for matrix multiplication.");
  if (!a.isDiag() && !b.isDiag()) {
    Matrix c(a.row(), b.col(), false);
    if (a.row() && b.col())
      detail::check(tn_matmul_f64(a.getElem(), b.getElem(), (int64_t)a.row(), (int64_t)b.col(),
                                  (int64_t)a.col(), c.getElem(), detail::gemm_mode(), detail::stream()),
                    "matrixMul");
    return c;
  }
  if (a.isDiag() && !b.isDiag()) {
    Matrix c(b);
    c.resize(a.row(), b.col());
    size_t n = std::min(a.elemNum(), c.row());
    if (n < c.row()) detail::fail("diagonal x dense: diagonal shorter than the row count");
    detail::check(tn_diag_row_mul_f64(c.getElem(), a.getElem(), (int64_t)c.row(), (int64_t)c.col(),
                                      detail::stream()),
                  "diagRowMul");
    return c;
  }
  if (!a.isDiag() && b.isDiag()) {
    Matrix c(a);
    c.resize(a.row(), b.col());
    if (b.elemNum() < c.col()) detail::fail("dense x diagonal: diagonal shorter than the column count");
    detail::check(tn_diag_col_mul_f64(c.getElem(), b.getElem(), (int64_t)c.row(), (int64_t)c.col(),
                                      detail::stream()),
                  "diagColMul");
    return c;
  }
  // diagonal x diagonal (BlockTools.cpp:62-69): the first min(|a|, |b|)
  // entries of a times those of b (vectorMul), the rest zero.
  Matrix c(a.row(), b.col(), true);
  c.set_zero();
  const size_t n = std::min(std::min(a.elemNum(), b.elemNum()), c.elemNum());
  if (n) {
    detail::check(tn_memcpy_d2d(c.getElem(), a.getElem(), n * sizeof(double), detail::stream()), "RDotR");
    // row i of the n x 1 matrix c times b_i: the elementwise product
    detail::check(tn_diag_row_mul_f64(c.getElem(), b.getElem(), (int64_t)n, 1, detail::stream()), "vectorMul");
  }
  return c;
}

Matrix operator*(const Matrix& a, const Matrix& b) { return RDotR(a, b); }

Matrix operator*(const Matrix& a, double s) {
  Matrix c(a);
  c *= s;
  return c;
}

Matrix operator*(double s, const Matrix& a) { return a * s; }

Matrix operator+(const Matrix& a, const Matrix& b) {
  Matrix c(a);
  c += b;
  return c;
}

bool operator==(const Matrix& a, const Matrix& b) {
  // Block.cpp:110-131: same shape/diag flag, elementwise |a-b| < 1e-12
  if (a.Rnum != b.Rnum || a.Cnum != b.Cnum || a.diag != b.diag || a.ctype != b.ctype) return false;
  auto ha = a.getHostElem(), hb = b.getHostElem();
  for (size_t i = 0; i < ha.size(); ++i)
    if (std::fabs(ha[i] - hb[i]) > 1e-12) return false;
  return true;
}

std::ostream& operator<<(std::ostream& os, const Matrix& m) {
  os << m.Rnum << " x " << m.Cnum << " = " << m.elemNum() << (m.diag ? ", Diagonal" : "")
     << (m.ctype ? ", Complex" : "") << "\n\n";
  auto h = m.getHostElem();
  if (m.ctype) {
    for (size_t r = 0; r < m.Rnum; ++r) {
      for (size_t c = 0; c < m.Cnum; ++c) {
        const size_t e = m.diag ? r : r * m.Cnum + c;
        const bool z = m.diag && r != c;
        os << " (" << std::fixed << std::setprecision(3) << (z ? 0.0 : h[2 * e]) << "," << (z ? 0.0 : h[2 * e + 1])
           << ")";
      }
      os << "\n\n";
    }
    return os;
  }
  for (size_t r = 0; r < m.Rnum; ++r) {
    for (size_t c = 0; c < m.Cnum; ++c) {
      double v = m.diag ? (r == c ? h[r] : 0.0) : h[r * m.Cnum + c];
      os << std::setw(7) << std::fixed << std::setprecision(3) << v;
    }
    os << "\n\n";
  }
  return os;
}

Matrix exph(double a, const Matrix& mat) {
  // MatrixTools.cpp:109-121: U^T diag(exp(a*lambda)) U, U = eigenvector rows
  std::vector<Matrix> r = mat.eigh();
  Matrix UT(r[1]);
  UT.transpose();
  auto ev = r[0].getHostElem();
  for (auto& v : ev) v = std::exp(a * v);
  r[0].setElem(ev);
  return UT * (r[0] * r[1]);
}

Matrix takeExp(double a, const Matrix& mat) { return exph(a, mat); }

size_t lanczosEigh(Matrix& ori_mat, double& E0, Matrix& psi, size_t max_iter, double err_tol) {
  // Block.cpp:474-506 -> lanczosEV (uni10_lapack_cpu.cpp:416-509), on the device
  if (ori_mat.typeID() != 1) detail::fail("lanczosEigh: the complex form is tn_lanczos_c128");
  if (ori_mat.row() != ori_mat.col())
    detail::fail("Cannot perform Lanczos algorithm to find the lowest eigen value and eigen vector on a "
                 "non-square matrix.");
  if (ori_mat.row() != psi.elemNum())
    detail::fail("Error in Lanczos initial vector psi. The vector dimension does not match with the number of "
                 "the columns.");
  if (ori_mat.isDiag() || psi.isDiag()) detail::fail("lanczosEigh: diagonal matrices are not supported");
  int64_t it = (int64_t)max_iter;
  double e0 = 0.0;
  int conv = 0;
  detail::check(tn_lanczos_f64(ori_mat.getElem(), (int64_t)ori_mat.row(), psi.getElem(), &it, err_tol, &e0,
                               &conv, detail::stream()),
                "lanczosEigh");
  if (!conv) detail::fail("Lanczos algorithm fails in converging.");
  E0 = e0;
  return (size_t)it;
}

}  // namespace uni10
