// UniTensor: block-sparse symmetric FP64 tensor with its elements in HBM.
//
// The block plan ("grouping") is host metadata, built once per distinct bond
// list and cached.  Element work runs on the device through tn_hip.h:
//   permute (no symmetry)  -> tn_permute_dense_f64 (tiled transpose / row copy)
//   permute (symmetric)    -> one tn_blockperm plan per (layout, permutation),
//                             cached; one launch moves every sub-block
//   contract               -> one grouped GEMM plan over all common charges
//   transpose              -> tn_transpose_f64 per block
// Reference: uni10/src/uni10/tensor-network/lib/UniTensorReal.cpp
// (grouping :1169-1321, initUniT :1323-1385, setRawElem :96-174,
// getRawElem :241-280, at :1098-1165, permute :532-768, transpose :448-506,
// putBlock/getBlock :197-344), UniTensorTools.cpp (contract :124-244,
// otimes :246-273), UniTensor.cpp (operator= :236-305, save :922-986,
// file constructor :486-582, operator* :224-234).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <mutex>
#include <numeric>
#include <set>
#include <unordered_map>

#include "common.h"
#include "layout.h"

namespace uni10 {

int64_t UniTensor::COUNTER = 0;
int64_t UniTensor::ELEMNUM = 0;
int64_t UniTensor::MAXELEMNUM = 0;
int64_t UniTensor::MAXELEMTEN = 0;

// ---------------------------------------------------------------------------
// Layout (grouping) with a cache keyed by the bond list.
std::string bonds_key(const std::vector<Bond>& bonds) {
  std::string k;
  k.reserve(bonds.size() * 24);
  char tmp[64];
  for (const Bond& b : bonds) {
    k += (b.type() == BD_IN) ? 'I' : 'O';
    for (size_t i = 0; i < b.qnums().size(); ++i) {
      const Qnum& q = b.qnums()[i];
      std::snprintf(tmp, sizeof(tmp), "%d.%d.%d*%d,", q.U1(), (int)q.prt(), (int)q.prtF(), b.qdegs()[i]);
      k += tmp;
    }
    k += ';';
  }
  return k;
}

static LayoutPtr build_layout(const std::vector<Bond>& bonds) {
  auto L = std::make_shared<Layout>();
  L->bonds = bonds;
  L->key = bonds_key(bonds);
  const int nb = (int)bonds.size();
  if (nb == 0) {
    L->RBondNum = 0;
    L->RQdim = L->CQdim = 0;
    L->blocks[Qnum(0)] = BlockInfo{1, 1, 0};
    L->elemNum = 1;
    return L;
  }
  int nrow = 0, ncol = 0;
  bool seen_out = false;
  L->RQdim = L->CQdim = 1;
  for (const Bond& b : bonds) {
    if (b.type() == BD_IN) {
      if (seen_out) detail::fail("Error in the input bond array: BD_OUT bonds must be placed after all BD_IN bonds.");
      L->RQdim *= b.qnums().size();
      nrow++;
    } else {
      L->CQdim *= b.qnums().size();
      ncol++;
      seen_out = true;
    }
  }
  L->RBondNum = nrow;
  // rows: odometer over run indices of the in-bonds, last bond fastest
  std::map<Qnum, size_t> row_mdim, col_mdim;
  std::map<Qnum, std::vector<int>> row_q2idx, col_q2idx;
  std::vector<size_t> tR_dim(L->RQdim, 1), tR_off(L->RQdim, 0), tC_dim(L->CQdim, 1), tC_off(L->CQdim, 0);
  if (nrow) {
    std::vector<int> offs(nrow, 0);
    for (int boff = 0;; ++boff) {
      Qnum q(0);
      size_t dim = 1;
      for (int b = 0; b < nrow; ++b) {
        q = q * bonds[b].qnums()[offs[b]];
        dim *= bonds[b].qdegs()[offs[b]];
      }
      auto it = row_mdim.find(q);
      tR_off[boff] = (it == row_mdim.end()) ? 0 : it->second;
      tR_dim[boff] = dim;
      row_mdim[q] = tR_off[boff] + dim;
      row_q2idx[q].push_back(boff);
      int bidx;
      for (bidx = nrow - 1; bidx >= 0; --bidx) {
        if (++offs[bidx] < (int)bonds[bidx].qnums().size()) break;
        offs[bidx] = 0;
      }
      if (bidx < 0) break;
    }
  } else {
    row_mdim[Qnum(0)] = 1;
    row_q2idx[Qnum(0)].push_back(0);
  }
  if (ncol) {
    std::vector<int> offs(ncol, 0);
    for (int boff = 0;; ++boff) {
      Qnum q(0);
      size_t dim = 1;
      for (int b = 0; b < ncol; ++b) {
        q = q * bonds[nrow + b].qnums()[offs[b]];
        dim *= bonds[nrow + b].qdegs()[offs[b]];
      }
      if (row_mdim.count(q)) {
        auto it = col_mdim.find(q);
        tC_off[boff] = (it == col_mdim.end()) ? 0 : it->second;
        tC_dim[boff] = dim;
        col_mdim[q] = tC_off[boff] + dim;
        col_q2idx[q].push_back(boff);
      }
      int bidx;
      for (bidx = ncol - 1; bidx >= 0; --bidx) {
        if (++offs[bidx] < (int)bonds[nrow + bidx].qnums().size()) break;
        offs[bidx] = 0;
      }
      if (bidx < 0) break;
    }
  } else if (row_mdim.count(Qnum(0))) {
    col_mdim[Qnum(0)] = 1;
    col_q2idx[Qnum(0)].push_back(0);
  }
  std::set<int64_t> qidx_set;
  size_t off = 0;
  for (auto& kv : col_mdim) {
    const Qnum& q = kv.first;
    BlockInfo bi{row_mdim[q], kv.second, off};
    off += bi.Rnum * bi.Cnum;
    L->blocks[q] = bi;
    for (int r : row_q2idx[q]) {
      L->RQidx2Blk[r] = q;
      for (int c : col_q2idx[q]) {
        L->RQidx2Dim[r] = tR_dim[r];
        L->RQidx2Off[r] = tR_off[r];
        L->CQidx2Dim[c] = tC_dim[c];
        L->CQidx2Off[c] = tC_off[c];
        qidx_set.insert((int64_t)r * L->CQdim + c);
      }
    }
  }
  size_t enc = 0;
  for (auto& r : L->RQidx2Dim)
    for (auto& c : L->CQidx2Dim) {
      int64_t qidx = (int64_t)r.first * L->CQdim + c.first;
      if (qidx_set.count(qidx)) {
        L->QidxEnc.emplace_back(qidx, enc);
        enc += r.second * c.second;
      }
    }
  L->elemNum = off;
  if (L->blocks.empty()) {
    std::ostringstream err;
    err << "There is no symmetry block with the given bonds:\n";
    for (const Bond& b : bonds) err << "    " << b;
    detail::fail(err.str());
  }
  return L;
}

LayoutPtr get_layout(const std::vector<Bond>& bonds) {
  static std::mutex mu;
  static std::unordered_map<std::string, LayoutPtr> cache;
  std::string key = bonds_key(bonds);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  LayoutPtr L = build_layout(bonds);
  std::lock_guard<std::mutex> g(mu);
  if (cache.size() > 20000) cache.clear();
  cache.emplace(key, L);
  return L;
}

void Layout::decode(int64_t qidx, std::vector<int>& idxs) const {
  idxs.resize(bonds.size());
  for (int b = (int)bonds.size() - 1; b >= 0; --b) {
    int n = (int)bonds[b].qnums().size();
    idxs[b] = (int)(qidx % n);
    qidx /= n;
  }
}

void Layout::subblock(int64_t qidx, size_t& eoff, size_t& bcols, size_t& sr, size_t& sc) const {
  int r = (int)(qidx / CQdim), c = (int)(qidx % CQdim);
  const BlockInfo& bi = blocks.at(RQidx2Blk.at(r));
  eoff = bi.offset + RQidx2Off.at(r) * bi.Cnum + CQidx2Off.at(c);
  bcols = bi.Cnum;
  sr = RQidx2Dim.at(r);
  sc = CQidx2Dim.at(c);
}

// ---------------------------------------------------------------------------
UniTensor::UniTensor() : m_name("") {
  initUniT();
}

UniTensor::UniTensor(double val) : m_name("") {
  initUniT();
  detail::check(tn_memcpy_h2d(buf.get(), &val, sizeof(double), detail::stream()), "UniTensor(double)");
}

UniTensor::UniTensor(const std::vector<Bond>& b, const std::string& name) : m_name(name), bonds(b) {
  initUniT();
}

UniTensor::UniTensor(const std::vector<Bond>& b, const std::string& name, InitFill fill)
    : m_name(name), bonds(b) {
  initUniT(fill == Zero);
}

UniTensor::UniTensor(const std::vector<Bond>& b, std::vector<int>& lab, const std::string& name)
    : m_name(name), bonds(b) {
  initUniT();
  setLabel(lab);
}

// A complex tensor is held as the real tensor with one more bond, an
// outgoing bond of dimension 2 and charge 0 last: it doubles every block's
// column count and keeps (re, im) of each element adjacent, so the block
// layout is the complex layout with interleaved pairs, and every layout
// operation (permute with the pair bond kept last, setRawElem, getBlock,
// block contraction) is the real one on that tensor.
namespace {
constexpr int kPairLabel = -1073741823;        // the pair bond's label (no caller uses it)
constexpr int kPairLabel2 = -1073741822;       // the second pair axis of a contraction's expansion
}  // namespace

void UniTensor::make_complex(const std::vector<Bond>& b, bool zero) {
  bonds = b;
  bonds.push_back(Bond(BD_OUT, 2));
  initUniT(zero);
  labels.back() = kPairLabel;
  m_ctype = true;
}

UniTensor::UniTensor(cflag, const std::vector<Bond>& b, const std::string& name) : m_name(name) {
  make_complex(b, true);
}

UniTensor::UniTensor(cflag, const std::vector<Bond>& b, std::vector<int>& lab, const std::string& name)
    : m_name(name) {
  make_complex(b, true);
  setLabel(lab);
}

void UniTensor::real_only(const char* what) const {
  if (m_ctype) detail::fail(std::string(what) + ": not provided for complex tensors");
}

void UniTensor::initUniT(bool zero) {
  m_layout = get_layout(bonds);
  status = 0;
  if (!bonds.empty()) {
    m_elemNum = m_layout->elemNum;
    RBondNum = m_layout->RBondNum;
    labels.resize(bonds.size());
    std::iota(labels.begin(), labels.end(), 0);
    status |= HAVEBOND;
  } else {
    m_elemNum = 1;
    RBondNum = 0;
    labels.clear();
    status |= HAVEELEM;
  }
  buf.reset(m_elemNum, zero);
  ELEMNUM += m_elemNum;
  COUNTER++;
  MAXELEMNUM = std::max(MAXELEMNUM, ELEMNUM);
  MAXELEMTEN = std::max<int64_t>(MAXELEMTEN, (int64_t)m_elemNum);
}

UniTensor::UniTensor(const UniTensor& T)
    : m_ctype(T.m_ctype), m_name(T.m_name), bonds(T.bonds), labels(T.labels), m_layout(T.m_layout),
      m_elemNum(T.m_elemNum), RBondNum(T.RBondNum), status(T.status) {
  buf.reset(m_elemNum, false);
  detail::check(tn_memcpy_d2d(buf.get(), T.buf.get(), m_elemNum * sizeof(double), detail::stream()),
                "UniTensor copy");
  ELEMNUM += m_elemNum;
  COUNTER++;
  MAXELEMNUM = std::max(MAXELEMNUM, ELEMNUM);
}

UniTensor::UniTensor(UniTensor&& T) noexcept
    : m_ctype(T.m_ctype), m_name(std::move(T.m_name)), bonds(std::move(T.bonds)), labels(std::move(T.labels)),
      m_layout(std::move(T.m_layout)), buf(std::move(T.buf)), m_elemNum(T.m_elemNum), RBondNum(T.RBondNum),
      status(T.status) {
  T.m_elemNum = 0;
  T.status = 0;
}

UniTensor& UniTensor::operator=(const UniTensor& T) {
  if (this == &T) return *this;
  ELEMNUM -= m_elemNum;
  m_ctype = T.m_ctype;
  m_name = T.m_name;
  bonds = T.bonds;
  labels = T.labels;
  m_layout = T.m_layout;
  m_elemNum = T.m_elemNum;
  RBondNum = T.RBondNum;
  status = T.status;
  if (buf.size() != m_elemNum) buf.reset(m_elemNum, false);
  detail::check(tn_memcpy_d2d(buf.get(), T.buf.get(), m_elemNum * sizeof(double), detail::stream()),
                "UniTensor assign");
  ELEMNUM += m_elemNum;
  MAXELEMNUM = std::max(MAXELEMNUM, ELEMNUM);
  return *this;
}

UniTensor& UniTensor::operator=(UniTensor&& T) noexcept {
  if (this == &T) return *this;
  ELEMNUM -= m_elemNum;
  m_ctype = T.m_ctype;
  m_name = std::move(T.m_name);
  bonds = std::move(T.bonds);
  labels = std::move(T.labels);
  m_layout = std::move(T.m_layout);
  buf = std::move(T.buf);
  m_elemNum = T.m_elemNum;
  RBondNum = T.RBondNum;
  status = T.status;
  T.m_elemNum = 0;
  T.status = 0;
  return *this;
}

UniTensor::~UniTensor() {
  ELEMNUM -= m_elemNum;
  if (m_elemNum || buf.get()) COUNTER--;
}

UniTensor& UniTensor::assign(const std::vector<Bond>& b) {
  if (m_ctype) {
    UniTensor T(CTYPE, b, m_name);
    *this = std::move(T);
    return *this;
  }
  UniTensor T(b, m_name);
  *this = std::move(T);
  return *this;
}

void UniTensor::setLabel(const std::vector<int>& newLabels) {
  if (newLabels.size() != bondNum())
    detail::fail("The size of input vector(labels) does not match for the number of bonds.");
  std::set<int> s(newLabels.begin(), newLabels.end());
  if (s.size() != newLabels.size()) detail::fail("Labels must be distinct.");
  if (m_ctype && s.count(kPairLabel)) detail::fail("Label -1073741823 is reserved for complex tensors.");
  labels = newLabels;
  if (m_ctype) labels.push_back(kPairLabel);
}

void UniTensor::setLabel(int* newLabels) { setLabel(std::vector<int>(newLabels, newLabels + bondNum())); }

size_t UniTensor::blockNum() const { return m_layout->blocks.size(); }

std::vector<Qnum> UniTensor::blockQnum() const {
  std::vector<Qnum> q;
  for (auto& kv : m_layout->blocks) q.push_back(kv.first);
  return q;
}

Qnum UniTensor::blockQnum(size_t idx) const {
  if (idx >= m_layout->blocks.size()) detail::fail("Index exceeds the number of the blocks.");
  auto it = m_layout->blocks.begin();
  std::advance(it, idx);
  return it->first;
}

const std::map<Qnum, BlockInfo>& UniTensor::blockInfo() const { return m_layout->blocks; }

// ---------------------------------------------------------------------------
void UniTensor::setRawElem(const double* raw) {
  if (!(status & HAVEBOND)) detail::fail("Setting elements to a tensor without bonds is not supported.");
  const Layout& L = *m_layout;
  const int nb = (int)bonds.size();
  std::vector<size_t> rAcc(nb, 1);
  for (int b = nb - 1; b > 0; --b) rAcc[b - 1] = rAcc[b] * bonds[b].dim();
  std::vector<double> work(m_elemNum, 0.0);
  std::vector<int> qi, sidx(nb), sdim(nb);
  for (auto& e : L.QidxEnc) {
    L.decode(e.first, qi);
    size_t roff = 0;
    for (int b = 0; b < nb; ++b) {
      roff += rAcc[b] * bonds[b].qoffsets()[qi[b]];
      sdim[b] = bonds[b].qdegs()[qi[b]];
      sidx[b] = 0;
    }
    size_t eoff, bcols, sr, sc;
    L.subblock(e.first, eoff, bcols, sr, sc);
    for (size_t r = 0; r < sr; ++r)
      for (size_t c = 0; c < sc; ++c) {
        work[eoff + r * bcols + c] = raw[roff];
        for (int bend = nb - 1; bend >= 0; --bend) {
          if (++sidx[bend] < sdim[bend]) {
            roff += rAcc[bend];
            break;
          }
          roff -= rAcc[bend] * (sidx[bend] - 1);
          sidx[bend] = 0;
        }
      }
  }
  detail::check(tn_memcpy_h2d(buf.get(), work.data(), m_elemNum * sizeof(double), detail::stream()),
                "setRawElem");
  status |= HAVEELEM;
}

void UniTensor::setRawElem(const std::vector<double>& raw) {
  size_t n = 1;
  for (auto& b : bonds) n *= b.dim();
  if (raw.size() < n) detail::fail("setRawElem: not enough elements");
  setRawElem(raw.data());
}

// (for a complex tensor the real forms read and write (re, im) pairs)
void UniTensor::setRawElem(const std::complex<double>* raw) {
  if (!m_ctype) detail::fail("setRawElem: complex elements for a real tensor");
  setRawElem(reinterpret_cast<const double*>(raw));
}

void UniTensor::setRawElem(const std::vector<std::complex<double>>& raw) {
  size_t n = 1;
  for (auto& b : bond()) n *= b.dim();
  if (raw.size() < n) detail::fail("setRawElem: not enough elements");
  setRawElem(raw.data());
}

void UniTensor::setElem(const std::complex<double>* elem, bool src_ongpu) {
  if (!m_ctype) detail::fail("setElem: complex elements for a real tensor");
  setElem(reinterpret_cast<const double*>(elem), src_ongpu);
}

std::complex<double> UniTensor::at(cflag, size_t idx) const {
  if (!m_ctype) detail::fail("at(CTYPE, idx): the tensor is real");
  if (idx >= m_elemNum / 2) detail::fail("Index exceeds the number of elements.");
  double v[2] = {0, 0};
  detail::check(tn_memcpy_d2h(v, buf.get() + 2 * idx, sizeof(v), detail::stream()), "at");
  return {v[0], v[1]};
}

std::vector<double> UniTensor::getHostElem() const {
  std::vector<double> h(m_elemNum);
  if (m_elemNum)
    detail::check(tn_memcpy_d2h(h.data(), buf.get(), m_elemNum * sizeof(double), detail::stream()),
                  "getHostElem");
  return h;
}

Matrix UniTensor::getRawElem() const {
  if (!(status & HAVEBOND)) {
    if (status & HAVEELEM) return Matrix(1, 1, buf.get(), false, true);
    return Matrix();
  }
  const Layout& L = *m_layout;
  const int nb = (int)bonds.size();
  size_t rows = 1, cols = 1;
  for (auto& b : bonds) (b.type() == BD_IN ? rows : cols) *= b.dim();
  std::vector<size_t> rAcc(nb, 1);
  for (int b = nb - 1; b > 0; --b) rAcc[b - 1] = rAcc[b] * bonds[b].dim();
  std::vector<double> h = getHostElem(), raw(rows * cols, 0.0);
  std::vector<int> qi, sidx(nb), sdim(nb);
  for (auto& e : L.QidxEnc) {
    L.decode(e.first, qi);
    size_t roff = 0;
    for (int b = 0; b < nb; ++b) {
      roff += rAcc[b] * bonds[b].qoffsets()[qi[b]];
      sdim[b] = bonds[b].qdegs()[qi[b]];
      sidx[b] = 0;
    }
    size_t eoff, bcols, sr, sc;
    L.subblock(e.first, eoff, bcols, sr, sc);
    for (size_t r = 0; r < sr; ++r)
      for (size_t c = 0; c < sc; ++c) {
        raw[roff] = h[eoff + r * bcols + c];
        for (int bend = nb - 1; bend >= 0; --bend) {
          if (++sidx[bend] < sdim[bend]) {
            roff += rAcc[bend];
            break;
          }
          roff -= rAcc[bend] * (sidx[bend] - 1);
          sidx[bend] = 0;
        }
      }
  }
  if (m_ctype) {
    // the pair bond is the last outgoing one: (re, im) adjacent in each row
    return Matrix(CTYPE, rows, cols / 2, reinterpret_cast<const std::complex<double>*>(raw.data()), false);
  }
  return Matrix(rows, cols, raw.data(), false, false);
}

void UniTensor::setElem(const double* elem, bool src_ongpu) {
  int rc = src_ongpu ? tn_memcpy_d2d(buf.get(), elem, m_elemNum * sizeof(double), detail::stream())
                     : tn_memcpy_h2d(buf.get(), elem, m_elemNum * sizeof(double), detail::stream());
  detail::check(rc, "setElem");
  status |= HAVEELEM;
}

void UniTensor::setElem(const std::vector<double>& elem, bool src_ongpu) {
  if (elem.size() < m_elemNum) detail::fail("setElem: not enough elements");
  setElem(elem.data(), src_ongpu);
}

double UniTensor::at(size_t idx) const {
  real_only("at (use at(CTYPE, idx))");
  if (idx >= m_elemNum) detail::fail("Index exceeds the number of elements.");
  double v = 0;
  detail::check(tn_memcpy_d2h(&v, buf.get() + idx, sizeof(double), detail::stream()), "at");
  return v;
}

double UniTensor::at(const std::vector<size_t>& idxs) const {
  real_only("at(idxs)");
  if (!(status & HAVEBOND)) detail::fail("The tensor is a scalar. Use UniTensor::operator[] instead.");
  if (idxs.size() != bonds.size())
    detail::fail("The size of input indices array does not match with the number of the bonds.");
  const Layout& L = *m_layout;
  const int nb = (int)bonds.size();
  std::vector<int> qi(nb);
  int64_t qoff = 0;
  for (int b = 0; b < nb; ++b) {
    if (!(idxs[b] < (size_t)bonds[b].dim())) detail::fail("The input indices are out of range.");
    const auto& of = bonds[b].qoffsets();
    int q = (int)of.size() - 1;
    while (q > 0 && (size_t)of[q] > idxs[b]) --q;
    qi[b] = q;
    qoff = qoff * (int64_t)bonds[b].qnums().size() + q;
  }
  auto it = std::lower_bound(L.QidxEnc.begin(), L.QidxEnc.end(), std::make_pair(qoff, (size_t)0),
                             [](const std::pair<int64_t, size_t>& a, const std::pair<int64_t, size_t>& b) {
                               return a.first < b.first;
                             });
  if (it == L.QidxEnc.end() || it->first != qoff) return 0.0;
  size_t eoff, bcols, sr, sc;
  L.subblock(qoff, eoff, bcols, sr, sc);
  size_t cnt = 0;
  for (int b = 0; b < nb; ++b) cnt = cnt * bonds[b].qdegs()[qi[b]] + (idxs[b] - bonds[b].qoffsets()[qi[b]]);
  return at(eoff + (cnt / sc) * bcols + cnt % sc);
}

void UniTensor::putBlock(const Matrix& mat) { putBlock(Qnum(0), mat); }

void UniTensor::putBlock(const Qnum& qnum, const Matrix& mat) {
  auto it = m_layout->blocks.find(qnum);
  if (it == m_layout->blocks.end()) {
    std::ostringstream err;
    err << "There is no block with the given quantum number " << qnum;
    detail::fail(err.str());
  }
  const BlockInfo& bi = it->second;
  if (mat.typeID() != typeID()) detail::fail("putBlock: the block and the tensor differ in type (real/complex)");
  if (m_ctype) {
    if (mat.isDiag()) detail::fail("putBlock: a complex diagonal block is not provided");
    if (!(mat.row() == bi.Rnum && 2 * mat.col() == bi.Cnum))
      detail::fail("The dimension of input matrix does not match for the dimension of the block");
    if (bi.Rnum * bi.Cnum)
      detail::check(tn_memcpy_d2d(buf.get() + bi.offset, mat.getElem(), bi.Rnum * bi.Cnum * sizeof(double),
                                  detail::stream()),
                    "putBlock");
    status |= HAVEELEM;
    return;
  }
  if (!(mat.row() == bi.Rnum && mat.col() == bi.Cnum)) {
    std::ostringstream err;
    err << "The dimension of input matrix does not match for the dimension of the block with quantum number "
        << qnum << "\n  Hint: Use Matrix::resize(int, int)";
    detail::fail(err.str());
  }
  double* dst = buf.get() + bi.offset;
  if (mat.isDiag()) {
    std::vector<double> d = mat.getHostElem(), full(bi.Rnum * bi.Cnum, 0.0);
    for (size_t i = 0; i < d.size(); ++i) full[i * bi.Cnum + i] = d[i];
    detail::check(tn_memcpy_h2d(dst, full.data(), full.size() * sizeof(double), detail::stream()), "putBlock");
  } else if (bi.Rnum * bi.Cnum) {
    detail::check(tn_memcpy_d2d(dst, mat.getElem(), bi.Rnum * bi.Cnum * sizeof(double), detail::stream()),
                  "putBlock");
  }
  status |= HAVEELEM;
}

Matrix UniTensor::getBlock(bool diag) const { return getBlock(Qnum(0), diag); }

Matrix UniTensor::getBlock(const Qnum& qnum, bool diag) const {
  auto it = m_layout->blocks.find(qnum);
  if (it == m_layout->blocks.end()) {
    std::ostringstream err;
    err << "There is no block with the given quantum number " << qnum;
    detail::fail(err.str());
  }
  Matrix m(it->second.Rnum, it->second.Cnum, buf.get() + it->second.offset, false, true);
  if (m_ctype) {
    if (diag) detail::fail("getBlock: a complex diagonal block is not provided");
    m.Cnum /= 2;     // the pair bond doubles the columns: the same doubles, read as complex
    m.ctype = true;
    return m;
  }
  return diag ? m.getDiag() : m;
}

std::map<Qnum, Matrix> UniTensor::getBlocks() const {
  std::map<Qnum, Matrix> out;
  for (auto& kv : m_layout->blocks) out.emplace(kv.first, getBlock(kv.first, false));
  return out;
}

// ---------------------------------------------------------------------------
// permute
namespace {
std::vector<std::pair<int, int>> rec_swap(const std::vector<int>& order, const std::vector<int>& ordF) {
  // uni10_tools.cpp:52-69: bubble sort recording each swapped bond pair
  std::vector<int> ord = order;
  std::vector<std::pair<int, int>> swaps;
  int n = (int)ord.size();
  for (int i = 0; i < n - 1; ++i)
    for (int j = 0; j < n - i - 1; ++j)
      if (ord[j] > ord[j + 1]) {
        swaps.emplace_back(ordF[ord[j + 1]], ordF[ord[j]]);
        std::swap(ord[j], ord[j + 1]);
      }
  return swaps;
}

struct PermPlanEntry {
  tn_blockperm_plan plan = nullptr;
};

std::mutex g_perm_mu;
std::unordered_map<std::string, std::shared_ptr<PermPlanEntry>> g_perm_cache;
}  // namespace

// Builds (or fetches) the device move list that maps the blocks of `in`
// (labels `inLabels`) to the blocks of `out` for the permutation rsp.
static tn_blockperm_plan block_perm_plan(const Layout& in, const std::vector<int>& inLabels, const Layout& out,
                                         const std::vector<int>& rsp, const std::vector<int>& newLabels) {
  const int nb = (int)in.bonds.size();
  const bool ferm = Qnum::isFermionic();
  std::string key = in.key + "#" + out.key + "#";
  for (int r : rsp) key += std::to_string(r) + ",";
  if (ferm) {
    key += "F";
    for (int l : inLabels) key += std::to_string(l) + ",";
    key += "/";
    for (int l : newLabels) key += std::to_string(l) + ",";
  }
  {
    std::lock_guard<std::mutex> g(g_perm_mu);
    auto it = g_perm_cache.find(key);
    if (it != g_perm_cache.end()) return it->second->plan;
  }
  // fermionic swap list (UniTensorReal.cpp:657-681)
  std::vector<std::pair<int, int>> swaps;
  if (ferm) {
    const int RB = in.RBondNum, oRB = out.RBondNum;
    std::vector<int> inF(nb), outF(nb), ordF(nb);
    for (int b = 0; b < RB; ++b) {
      inF[b] = inLabels[b];
      ordF[b] = b;
    }
    for (int b = 0; b < oRB; ++b) outF[b] = newLabels[b];
    for (int b = nb - 1; b >= RB; --b) {
      ordF[b] = nb - b + RB - 1;
      inF[ordF[b]] = inLabels[b];
    }
    for (int b = nb - 1; b >= oRB; --b) outF[nb - b + oRB - 1] = newLabels[b];
    std::vector<int> rspF(nb);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j)
        if (inF[i] == outF[j]) rspF[j] = i;
    swaps = rec_swap(rspF, ordF);
  }
  std::vector<int64_t> qot_acc(nb, 1);
  for (int b = nb - 1; b > 0; --b) qot_acc[b - 1] = qot_acc[b] * (int64_t)out.bonds[b].qnums().size();
  const size_t nm = in.QidxEnc.size();
  std::vector<int64_t> meta(6 * nm), sub(nb * nm), acc(nb * nm);
  std::vector<double> sign(nm, 1.0);
  std::vector<int> qin, qot(nb);
  std::vector<int64_t> sbot(nb);
  size_t m = 0;
  for (auto& e : in.QidxEnc) {
    in.decode(e.first, qin);
    int64_t qot_off = 0;
    for (int b = 0; b < nb; ++b) {
      qot[b] = qin[rsp[b]];
      qot_off += qot[b] * qot_acc[b];
    }
    std::fill(sbot.begin(), sbot.end(), 1);
    for (int b = nb - 1; b > 0; --b) sbot[rsp[b - 1]] = sbot[rsp[b]] * in.bonds[rsp[b]].qdegs()[qot[b]];
    size_t ein, bin_c, sr, sc, eot, bot_c, osr, osc;
    in.subblock(e.first, ein, bin_c, sr, sc);
    out.subblock(qot_off, eot, bot_c, osr, osc);
    int64_t* mm = &meta[6 * m];
    mm[0] = (int64_t)ein;
    mm[1] = (int64_t)bin_c;
    mm[2] = (int64_t)sc;
    mm[3] = (int64_t)eot;
    mm[4] = (int64_t)bot_c;
    mm[5] = (int64_t)osc;
    for (int b = 0; b < nb; ++b) {
      sub[nb * m + b] = in.bonds[b].qdegs()[qin[b]];
      acc[nb * m + b] = sbot[b];
    }
    if (!swaps.empty()) {
      int bit = 0;
      for (auto& s : swaps)
        bit ^= (int)in.bonds[s.first].qnums()[qin[s.first]].prtF() & (int)in.bonds[s.second].qnums()[qin[s.second]].prtF();
      sign[m] = bit ? -1.0 : 1.0;
    }
    ++m;
  }
  auto ent = std::make_shared<PermPlanEntry>();
  detail::check(tn_blockperm_plan_create(&ent->plan, nb, (int64_t)nm, meta.data(), sub.data(), acc.data(),
                                         sign.data(), (int64_t)out.elemNum),
                "tn_blockperm_plan_create");
  std::lock_guard<std::mutex> g(g_perm_mu);
  auto ins = g_perm_cache.emplace(key, ent);
  if (!ins.second) {
    // another thread built the same plan; keep theirs (ours is never used)
    tn_blockperm_plan_destroy(ent->plan);
  }
  return ins.first->second->plan;
}

UniTensor& UniTensor::permute(const std::vector<int>& newLabels, int rowBondNum) {
  if (!m_ctype) return permute_full(newLabels, rowBondNum);
  // complex (UniTensorComplex.cpp:592-825): the real permute with the pair
  // bond kept last (an outgoing bond), a bit-exact move of (re, im) pairs
  if (newLabels.size() != bondNum())
    detail::fail("The size of the input new labels does not match for the number of bonds.");
  if (rowBondNum < 0 || rowBondNum > (int)bondNum()) detail::fail("permute: inBondNum out of range");
  std::vector<int> full(newLabels);
  full.push_back(kPairLabel);
  m_ctype = false;
  try {
    permute_full(full, rowBondNum);
  } catch (...) {
    m_ctype = true;
    throw;
  }
  m_ctype = true;
  return *this;
}

UniTensor& UniTensor::permute_full(const std::vector<int>& newLabels, int rowBondNum) {
  if (!(status & HAVEBOND)) detail::fail("There is no bond in the tensor(scalar) to permute.");
  if (labels.size() != newLabels.size())
    detail::fail("The size of the input new labels does not match for the number of bonds.");
  const int nb = (int)bonds.size();
  std::vector<int> rsp(nb, 0);
  int cnt = 0;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j)
      if (labels[i] == newLabels[j]) {
        rsp[j] = i;
        cnt++;
      }
  if (cnt != nb) detail::fail("The input new labels do not 1-1 correspond to the labels of the tensor.");
  bool inorder = true;
  for (int i = 1; i < nb; ++i)
    if (rsp[i] != i) {
      inorder = false;
      break;
    }
  if (inorder && RBondNum == rowBondNum) return *this;
  std::vector<Bond> outBonds;
  bool withoutSym = true;
  for (int b = 0; b < nb; ++b) {
    outBonds.push_back(bonds[rsp[b]]);
    if (bonds[b].qnums().size() != 1) withoutSym = false;
  }
  for (int b = 0; b < nb; ++b) outBonds[b].change(b < rowBondNum ? BD_IN : BD_OUT);
  UniTensor out(outBonds, m_name, (status & HAVEELEM) ? NoZero : Zero);
  if (status & HAVEELEM) {
    if (withoutSym) {
      if (!inorder) {
        std::vector<int64_t> dims(nb);
        for (int b = 0; b < nb; ++b) dims[b] = bonds[b].dim();
        detail::check(tn_permute_dense_f64(buf.get(), out.buf.get(), nb, dims.data(), rsp.data(), detail::stream()),
                      "permute");
      } else {
        // the same element order, only the row / column split changes: the
        // permuted tensor takes this tensor's buffer (no copy)
        std::swap(out.buf, buf);
      }
    } else {
      tn_blockperm_plan p = block_perm_plan(*m_layout, labels, *out.m_layout, rsp, newLabels);
      detail::check(tn_blockperm_exec(p, buf.get(), out.buf.get(), detail::stream()), "permute");
    }
    out.status |= HAVEELEM;
  }
  *this = std::move(out);
  setLabel(newLabels);
  return *this;
}

UniTensor& UniTensor::permute(int* newLabels, int rowBondNum) {
  return permute(std::vector<int>(newLabels, newLabels + bondNum()), rowBondNum);
}

UniTensor& UniTensor::permute(int rowBondNum) { return permute(label(), rowBondNum); }

UniTensor& UniTensor::transpose() {
  real_only("transpose");
  if (!(status & HAVEBOND)) detail::fail("There is no bond in the tensor(scalar) to perform transposition.");
  const int nb = (int)bonds.size();
  int rb = 0;
  for (int b = 0; b < nb; ++b) {
    if (bonds[b].type() == BD_IN)
      rb++;
    else
      break;
  }
  const int cb = nb - rb;
  std::vector<int> rsp(nb), outLabels(nb);
  for (int b = 0; b < nb; ++b) rsp[b] = (b < cb) ? rb + b : b - cb;
  std::vector<Bond> outBonds;
  for (int b = 0; b < nb; ++b) {
    outBonds.push_back(bonds[rsp[b]]);
    outLabels[b] = labels[rsp[b]];
  }
  for (int b = 0; b < nb; ++b) outBonds[b].dummy_change(b < cb ? BD_IN : BD_OUT);
  UniTensor out(outBonds, m_name, (status & HAVEELEM) ? NoZero : Zero);
  out.setLabel(outLabels);
  if (status & HAVEELEM) {
    for (auto& kv : m_layout->blocks) {
      const BlockInfo& bi = kv.second;
      const BlockInfo& bo = out.m_layout->blocks.at(kv.first);
      if (bi.Rnum * bi.Cnum)
        detail::check(tn_transpose_f64(buf.get() + bi.offset, (int64_t)bi.Rnum, (int64_t)bi.Cnum,
                                       out.buf.get() + bo.offset, detail::stream()),
                      "transpose");
    }
    out.status |= HAVEELEM;
  }
  *this = std::move(out);
  return *this;
}

// exSwap (UniTensor.cpp:1237-1268): every pair (bond shared with Tb, bond
// not shared), in bond order.
std::vector<_Swap> UniTensor::exSwap(const UniTensor& Tb) const {
  real_only("exSwap");
  std::vector<_Swap> swaps;
  if (status & Tb.status & HAVEBOND) {
    std::vector<int> shared, left;
    for (int a = 0; a < (int)labels.size(); ++a) {
      bool found = false;
      for (int l : Tb.labels)
        if (labels[a] == l) found = true;
      (found ? shared : left).push_back(a);
    }
    for (int i : shared)
      for (int j : left) swaps.push_back(_Swap{i, j});
  }
  return swaps;
}

namespace {
struct GatePlanEntry {
  tn_rectscale_plan plan = nullptr;
};
std::mutex g_gate_mu;
std::unordered_map<std::string, std::shared_ptr<GatePlanEntry>> g_gate_cache;
}  // namespace

// addGate (UniTensorReal.cpp:868-921): the sub-block of Qidx q is multiplied
// by (-1)^(XOR over gates of prtF(b1, q) & prtF(b2, q)).  Only the negated
// sub-blocks are listed in the device plan, which scales them in place.
void UniTensor::addGate(const std::vector<_Swap>& swaps) {
  real_only("addGate");
  if (!(status & HAVEBOND)) detail::fail("Adding swap gates to a tensor without bonds(scalar).");
  if (!(status & HAVEELEM)) detail::fail("Cannot add swap gates to a tensor before setting its elements.");
  const int nb = (int)bonds.size();
  for (const _Swap& sw : swaps)
    if (sw.b1 < 0 || sw.b1 >= nb || sw.b2 < 0 || sw.b2 >= nb)
      detail::fail("addGate: swap gate bond index out of range.");
  if (swaps.empty()) return;
  std::string key = m_layout->key + "#G";
  for (const _Swap& sw : swaps) key += std::to_string(sw.b1) + "," + std::to_string(sw.b2) + ";";
  std::shared_ptr<GatePlanEntry> ent;
  {
    std::lock_guard<std::mutex> g(g_gate_mu);
    auto it = g_gate_cache.find(key);
    if (it != g_gate_cache.end()) ent = it->second;
  }
  if (!ent) {
    std::vector<int64_t> rects;
    std::vector<int> q;
    for (auto& e : m_layout->QidxEnc) {
      m_layout->decode(e.first, q);
      int bit = 0;
      for (const _Swap& sw : swaps)
        bit ^= (int)bonds[sw.b1].qnums()[q[sw.b1]].prtF() & (int)bonds[sw.b2].qnums()[q[sw.b2]].prtF();
      if (!bit) continue;
      size_t eoff, bcols, sr, sc;
      m_layout->subblock(e.first, eoff, bcols, sr, sc);
      rects.insert(rects.end(), {(int64_t)eoff, (int64_t)sr, (int64_t)sc, (int64_t)bcols});
    }
    ent = std::make_shared<GatePlanEntry>();
    detail::check(tn_rectscale_plan_create(&ent->plan, (int64_t)(rects.size() / 4), rects.data()),
                  "tn_rectscale_plan_create");
    std::lock_guard<std::mutex> g(g_gate_mu);
    auto ins = g_gate_cache.emplace(key, ent);
    if (!ins.second) {
      tn_rectscale_plan_destroy(ent->plan);
      ent = ins.first->second;
    }
  }
  detail::check(tn_rectscale_exec(ent->plan, buf.get(), -1.0, detail::stream()), "addGate");
}

UniTensor& UniTensor::operator*=(double a) {
  detail::check(tn_scale_f64(buf.get(), (int64_t)m_elemNum, a, detail::stream()), "operator*=");
  return *this;
}

double UniTensor::norm() const {
  DevBuf out(1);
  detail::check(tn_norm2_f64(buf.get(), (int64_t)m_elemNum, out.get(), detail::stream()), "norm");
  double v = 0;
  detail::check(tn_memcpy_d2h(&v, out.get(), sizeof(double), detail::stream()), "norm");
  return v;
}

UniTensor& UniTensor::normalize() {
  double n = norm();
  if (n > 0) *this *= 1.0 / n;
  return *this;
}

void UniTensor::randomize(uint64_t seed) {
  detail::check(tn_fill_uniform_f64(buf.get(), (int64_t)m_elemNum, seed, detail::stream()), "randomize");
  status |= HAVEELEM;
}

void UniTensor::set_zero() {
  detail::check(tn_zero(buf.get(), m_elemNum * sizeof(double), detail::stream()), "set_zero");
  status |= HAVEELEM;
}

// ---------------------------------------------------------------------------
// contract
namespace {
struct GemmPlanEntry {
  tn_gemm_plan plan = nullptr;
};
std::mutex g_gemm_mu;
std::unordered_map<std::string, std::shared_ptr<GemmPlanEntry>> g_gemm_cache;
}  // namespace

// One grouped GEMM over the charges present in both A and B (the per-Qnum
// loop of UniTensorTools.cpp:191-203); returns the plan and whether every
// block of C is written (else C must be zero-filled first).
// Tile autotuning: the first time a cached GEMM plan runs, every tile shape
// is timed on the real operands and the fastest kept (results are identical
// for every tile shape).  Thread-local switch; on by default.
namespace {
thread_local bool t_autotune = true;
std::unordered_map<tn_gemm_plan, bool> g_tuned;   // guarded by g_gemm_mu
}  // namespace

void setAutotune(bool on) { t_autotune = on; }

static void maybe_autotune(tn_gemm_plan plan, const double* A, const double* B, double* C) {
  if (!t_autotune) return;
  {
    std::lock_guard<std::mutex> g(g_gemm_mu);
    if (!g_tuned.emplace(plan, true).second) return;
  }
  detail::check(tn_gemm_plan_autotune(plan, A, B, C, 3, detail::stream(), nullptr), "tn_gemm_plan_autotune");
}

static tn_gemm_plan contract_plan(const Layout& A, const Layout& B, const Layout& C, int mode, bool* covers_c) {
  std::string key = A.key + "#" + B.key + "#" + C.key + "#" + std::to_string(mode);
  size_t covered = 0;
  std::vector<tn_gemm_desc> descs;
  for (auto& kv : A.blocks) {
    auto itb = B.blocks.find(kv.first);
    if (itb == B.blocks.end()) continue;
    auto itc = C.blocks.find(kv.first);
    const BlockInfo& a = kv.second;
    const BlockInfo& b = itb->second;
    if (itc == C.blocks.end() || !(a.Rnum == itc->second.Rnum && b.Cnum == itc->second.Cnum && a.Cnum == b.Rnum))
      detail::fail("The dimensions the bonds to be contracted out are different.");
    const BlockInfo& c = itc->second;
    covered += c.Rnum * c.Cnum;
    tn_gemm_desc d;
    d.M = (int64_t)a.Rnum;
    d.N = (int64_t)b.Cnum;
    d.K = (int64_t)a.Cnum;
    d.a_off = (int64_t)a.offset;
    d.b_off = (int64_t)b.offset;
    d.c_off = (int64_t)c.offset;
    d.lda = std::max<int64_t>(d.K, 1);
    d.ldb = std::max<int64_t>(d.N, 1);
    d.ldc = std::max<int64_t>(d.N, 1);
    descs.push_back(d);
  }
  *covers_c = (covered == C.elemNum);
  {
    std::lock_guard<std::mutex> g(g_gemm_mu);
    auto it = g_gemm_cache.find(key);
    if (it != g_gemm_cache.end()) return it->second->plan;
  }
  auto ent = std::make_shared<GemmPlanEntry>();
  detail::check(tn_gemm_plan_create(&ent->plan, mode, (int64_t)descs.size(), descs.data()), "tn_gemm_plan_create");
  std::lock_guard<std::mutex> g(g_gemm_mu);
  auto ins = g_gemm_cache.emplace(key, ent);
  if (!ins.second) tn_gemm_plan_destroy(ent->plan);
  return ins.first->second->plan;
}

static bool is_dense(const UniTensor& T) {
  for (const Bond& b : T.bond())
    if (b.qnums().size() != 1) return false;
  return true;
}

// Gather-addressed GEMM of dense tensors: A rows = A's free bonds (A order),
// K = shared bonds in A's order, B columns = B's free bonds (B order).
static tn_gemm_plan gather_plan(const UniTensor& Ta, const std::vector<int>& a_free, const std::vector<int>& a_sh,
                                const UniTensor& Tb, const std::vector<int>& b_sh, const std::vector<int>& b_free,
                                const Layout& C, int mode) {
  auto strides = [](const UniTensor& T) {
    std::vector<int64_t> s(T.bondNum(), 1);
    for (int b = (int)T.bondNum() - 1; b > 0; --b) s[b - 1] = s[b] * T.bond(b).dim();
    return s;
  };
  const std::vector<int64_t> sa = strides(Ta), sb = strides(Tb);
  std::string key = "G#" + Ta.layout()->key + "#" + Tb.layout()->key + "#" + std::to_string(mode) + "#";
  for (int v : a_free) key += std::to_string(v) + ",";
  key += "|";
  for (int v : a_sh) key += std::to_string(v) + ",";
  key += "|";
  for (int v : b_sh) key += std::to_string(v) + ",";
  key += "|";
  for (int v : b_free) key += std::to_string(v) + ",";
  {
    std::lock_guard<std::mutex> g(g_gemm_mu);
    auto it = g_gemm_cache.find(key);
    if (it != g_gemm_cache.end()) return it->second->plan;
  }
  // Index map of the multi-index over `axes` (last fastest): unit axes drop,
  // an axis merges into the previous level when they are adjacent in storage.
  auto make_map = [](const UniTensor& T, const std::vector<int>& axes, const std::vector<int64_t>& st,
                     tn_index_map& m, int64_t& range) {
    std::vector<std::pair<int64_t, int64_t>> lv;   // (extent, stride), outermost first
    range = 1;
    for (int a : axes) {
      const int64_t e = T.bond(a).dim();
      range *= e;
      if (e == 1) continue;
      if (!lv.empty() && lv.back().second == st[a] * e) {
        lv.back().first *= e;
        lv.back().second = st[a];
      } else {
        lv.emplace_back(e, st[a]);
      }
    }
    if (lv.empty()) lv.emplace_back(1, 0);
    if (lv.size() > TN_MAX_MAP_LEVELS) return false;
    std::memset(&m, 0, sizeof(m));
    m.n = (int32_t)lv.size();
    for (size_t l = 0; l < lv.size(); ++l) {
      m.ext[l] = lv[l].first;
      m.stride[l] = lv[l].second;
    }
    return true;
  };
  tn_gemm_desc_ex d;
  std::memset(&d, 0, sizeof(d));
  int64_t kb = 0;
  if (!make_map(Ta, a_free, sa, d.a_row, d.M) || !make_map(Ta, a_sh, sa, d.a_col, d.K) ||
      !make_map(Tb, b_sh, sb, d.b_row, kb) || !make_map(Tb, b_free, sb, d.b_col, d.N))
    return nullptr;   // more than TN_MAX_MAP_LEVELS strided levels: use the permute path
  if (kb != d.K) detail::fail("The dimensions the bonds to be contracted out are different.");
  if (d.M >= (int64_t(1) << 31) || d.N >= (int64_t(1) << 31) || d.K >= (int64_t(1) << 31)) return nullptr;
  const BlockInfo& cb = C.blocks.begin()->second;
  if (C.blocks.size() != 1 || (int64_t)cb.Rnum != d.M || (int64_t)cb.Cnum != d.N)
    detail::fail("contract: unexpected output layout for dense operands");
  d.lda = std::max<int64_t>(d.K, 1);
  d.ldb = std::max<int64_t>(d.N, 1);
  d.ldc = std::max<int64_t>(d.N, 1);
  auto ent = std::make_shared<GemmPlanEntry>();
  detail::check(tn_gemm_plan_create_ex(&ent->plan, mode, 1, &d), "tn_gemm_plan_create_ex");
  std::lock_guard<std::mutex> g(g_gemm_mu);
  auto ins = g_gemm_cache.emplace(key, ent);
  if (!ins.second) tn_gemm_plan_destroy(ent->plan);
  return ins.first->second->plan;
}

UniTensor contract_impl(UniTensor& Ta, UniTensor& Tb, bool fast, bool allow_gather) {
  if (!(Ta.status & Tb.status & UniTensor::HAVEELEM))
    detail::fail("Cannot perform contraction of two tensors before setting their elements.");
  if (Ta.m_ctype || Tb.m_ctype) return contract_complex(Ta, Tb, fast, allow_gather);
  if (&Ta == &Tb) {
    UniTensor tmp = Tb;
    return contract_impl(Ta, tmp, fast, allow_gather);
  }
  if ((Ta.status & UniTensor::HAVEBOND) && (Tb.status & UniTensor::HAVEBOND)) {
    const int na = (int)Ta.bonds.size(), nbb = (int)Tb.bonds.size();
    std::vector<int> oldA = Ta.labels, oldB = Tb.labels;
    const int oldRA = Ta.RBondNum, oldRB = Tb.RBondNum;
    std::vector<int> newA, inter, newB, newC, markB(nbb, 0);
    for (int a = 0; a < na; ++a) {
      bool match = false;
      for (int b = 0; b < nbb; ++b)
        if (Ta.labels[a] == Tb.labels[b]) {
          markB[b] = 1;
          inter.push_back(Ta.labels[a]);
          newB.push_back(Tb.labels[b]);
          if (Ta.bonds[a].dim() != Tb.bonds[b].dim())
            detail::fail("Cannot contract two bonds having different dimensions");
          match = true;
          break;
        }
      if (!match) {
        newA.push_back(Ta.labels[a]);
        newC.push_back(Ta.labels[a]);
      }
    }
    newA.insert(newA.end(), inter.begin(), inter.end());
    for (int b = 0; b < nbb; ++b)
      if (!markB[b]) {
        newB.push_back(Tb.labels[b]);
        newC.push_back(Tb.labels[b]);
      }
    const int con = (int)inter.size();
    bool gather = allow_gather && is_dense(Ta) && is_dense(Tb);
    UniTensor Tc(std::vector<Bond>(), "", UniTensor::NoZero);
    if (gather) {
      // Dense operands: read A and B in their stored bond order through
      // multi-stride index maps; no permute of the inputs (their observable
      // state is what the reference leaves after its permute + restore).
      std::vector<int> a_free, a_sh, b_sh, b_free;
      for (int a = 0; a < na; ++a)
        (std::find(inter.begin(), inter.end(), Ta.labels[a]) == inter.end() ? a_free : a_sh).push_back(a);
      for (int l : inter)
        for (int b = 0; b < nbb; ++b)
          if (Tb.labels[b] == l) b_sh.push_back(b);
      // shared bonds of A in the order of `inter` (A's own order)
      for (int b = 0; b < nbb; ++b)
        if (!markB[b]) b_free.push_back(b);
      std::vector<Bond> cBonds;
      for (int a : a_free) cBonds.push_back(Bond(Ta.bonds[a]).change(BD_IN));
      for (int b : b_free) cBonds.push_back(Bond(Tb.bonds[b]).change(BD_OUT));
      Tc = UniTensor(cBonds, "", UniTensor::NoZero);
      if (!cBonds.empty()) Tc.setLabel(newC);
      tn_gemm_plan plan = gather_plan(Ta, a_free, a_sh, Tb, b_sh, b_free, *Tc.m_layout, detail::gemm_mode());
      if (plan) {
        maybe_autotune(plan, Ta.buf.get(), Tb.buf.get(), Tc.buf.get());
        detail::check(tn_gemm_exec(plan, Ta.buf.get(), Tb.buf.get(), Tc.buf.get(), detail::stream()), "contract");
      } else {
        gather = false;
      }
    }
    if (!gather) {
      Ta.permute(newA, na - con);
      Tb.permute(newB, con);
      std::vector<Bond> cBonds;
      for (int i = 0; i < na - con; ++i) cBonds.push_back(Ta.bonds[i]);
      for (int i = con; i < nbb; ++i) cBonds.push_back(Tb.bonds[i]);
      Tc = UniTensor(cBonds, "", UniTensor::NoZero);
      if (!cBonds.empty()) Tc.setLabel(newC);
      bool covers = false;
      tn_gemm_plan plan = contract_plan(*Ta.m_layout, *Tb.m_layout, *Tc.m_layout, detail::gemm_mode(), &covers);
      maybe_autotune(plan, Ta.buf.get(), Tb.buf.get(), Tc.buf.get());
      if (!covers)
        detail::check(tn_zero(Tc.buf.get(), Tc.m_elemNum * sizeof(double), detail::stream()), "contract zero");
      detail::check(tn_gemm_exec(plan, Ta.buf.get(), Tb.buf.get(), Tc.buf.get(), detail::stream()), "contract");
    }
    Tc.status |= UniTensor::HAVEELEM;
    if (con == 0) {
      std::vector<int> lab;
      for (int i = 0; i < oldRA; ++i) lab.push_back(oldA[i]);
      for (int i = 0; i < oldRB; ++i) lab.push_back(oldB[i]);
      for (int i = oldRA; i < na; ++i) lab.push_back(oldA[i]);
      for (int i = oldRB; i < nbb; ++i) lab.push_back(oldB[i]);
      Tc.permute(lab, oldRA + oldRB);
    }
    if (!fast && !gather) {
      Ta.permute(oldA, oldRA);
      Tb.permute(oldB, oldRB);
    }
    return Tc;
  }
  if (Ta.status & UniTensor::HAVEBOND) return Ta * Tb.at((size_t)0);
  if (Tb.status & UniTensor::HAVEBOND) return Ta.at((size_t)0) * Tb;
  return UniTensor(Ta.at((size_t)0) * Tb.at((size_t)0));
}

// Complex contraction (UniTensorTools.cpp:58-82, contract(CTYPE, ...)): a
// real operand is promoted to a complex copy (RtoC) first.  With both
// complex, B's elements b = br + i bi are expanded to the real 2 x 2 factor
// X = [[br, bi], [-bi, br]] (one more outgoing pair bond); contracting A's
// pair bond with X's first pair axis is the complex product, and X's second
// pair axis is the result's pair bond, last.  All block GEMMs are the real
// ones of the augmented tensors.
UniTensor contract_complex(UniTensor& Ta, UniTensor& Tb, bool fast, bool allow_gather) {
  auto promote = [](const UniTensor& R) {
    UniTensor C(CTYPE, R.bond(), R.getName());
    if (R.hasBond()) {
      std::vector<int> l = R.label();
      C.setLabel(l);
    }
    if (R.m_elemNum)
      detail::check(tn_copy_strided_f64(C.buf.get(), 2, R.buf.get(), 1, (int64_t)R.m_elemNum, detail::stream()),
                    "contract: promote to complex");
    C.status |= UniTensor::HAVEELEM;
    return C;
  };
  if (!Ta.m_ctype) {
    UniTensor A = promote(Ta);
    return contract_complex(A, Tb, fast, allow_gather);
  }
  if (!Tb.m_ctype) {
    UniTensor B = promote(Tb);
    return contract_complex(Ta, B, fast, allow_gather);
  }
  const std::vector<int> visA = Ta.label(), visB = Tb.label();
  const int oldRA = Ta.RBondNum, oldRB = Tb.RBondNum;
  std::vector<Bond> xb = Tb.bonds;
  xb.push_back(Bond(BD_OUT, 2));
  UniTensor X(xb, "", UniTensor::NoZero);
  std::vector<int> xl = Tb.labels;   // ends with the pair label, matching A's
  xl.push_back(kPairLabel2);
  X.setLabel(xl);
  detail::check(tn_expand_c128(Tb.buf.get(), (int64_t)(Tb.m_elemNum / 2), X.buf.get(), detail::stream()),
                "contract: complex expansion");
  X.status |= UniTensor::HAVEELEM;
  Ta.m_ctype = false;
  UniTensor Tc;
  try {
    Tc = contract_impl(Ta, X, fast, allow_gather);
  } catch (...) {
    Ta.m_ctype = true;
    throw;
  }
  Ta.m_ctype = true;
  Tc.labels.back() = kPairLabel;
  Tc.m_ctype = true;
  // the visible bonds shared by A and B (B's side of the reference's permute)
  std::vector<int> inter, freeB;
  for (int l : visA)
    if (std::find(visB.begin(), visB.end(), l) != visB.end()) inter.push_back(l);
  for (int l : visB)
    if (std::find(inter.begin(), inter.end(), l) == inter.end()) freeB.push_back(l);
  if (inter.empty() && Tc.bondNum() > 0) {
    // outer product: the real path's order (rows of A, rows of B, columns of A, columns of B)
    std::vector<int> lab;
    for (int i = 0; i < oldRA; ++i) lab.push_back(visA[i]);
    for (int i = 0; i < oldRB; ++i) lab.push_back(visB[i]);
    for (int i = oldRA; i < (int)visA.size(); ++i) lab.push_back(visA[i]);
    for (int i = oldRB; i < (int)visB.size(); ++i) lab.push_back(visB[i]);
    Tc.permute(lab, oldRA + oldRB);
  }
  if (fast) {
    // fast leaves B permuted to (shared bonds, free bonds), as the real path
    std::vector<int> newB(inter);
    newB.insert(newB.end(), freeB.begin(), freeB.end());
    Tb.permute(newB, (int)inter.size());
  }
  return Tc;
}

UniTensor contract(UniTensor& Ta, UniTensor& Tb, bool fast) {
  // fast=true leaves the inputs permuted (observable, UniTensorTools.cpp:178-179),
  // so only fast=false may skip the input permutes.
  return contract_impl(Ta, Tb, fast, !fast);
}

UniTensor contract_nopermute(UniTensor& Ta, UniTensor& Tb) { return contract_impl(Ta, Tb, true, true); }

UniTensor operator*(const UniTensor& Ta, const UniTensor& Tb) {
  UniTensor a = Ta, b = Tb;
  return contract(a, b, true);
}

UniTensor operator*(const UniTensor& T, double a) {
  UniTensor c = T;
  c *= a;
  return c;
}

UniTensor operator*(double a, const UniTensor& T) { return T * a; }

UniTensor otimes(const UniTensor& Ta, const UniTensor& Tb) {
  if (Ta.typeID() != 1 || Tb.typeID() != 1) detail::fail("otimes: not provided for complex tensors");
  UniTensor T1 = Ta, T2 = Tb;
  std::vector<int> l1(T1.bondNum()), l2(T2.bondNum());
  for (size_t i = 0; i < T1.bondNum(); ++i)
    l1[i] = (i < T1.inBondNum()) ? (int)i : (int)(T2.inBondNum() + i);
  for (size_t i = 0; i < T2.bondNum(); ++i)
    l2[i] = (i < T2.inBondNum()) ? (int)(i + T1.inBondNum()) : (int)(i + T1.bondNum());
  T1.setLabel(l1);
  T2.setLabel(l2);
  return contract(T1, T2, true);
}

// ---------------------------------------------------------------------------
// binary file format (UniTensor.cpp:922-986 writes; :486-582 reads)
void UniTensor::save(const std::string& fname) const {
  real_only("save");
  if (!(status & HAVEBOND)) detail::fail("Saving a tensor without bonds(scalar) is not supported.");
  FILE* fp = std::fopen(fname.c_str(), "wb");
  if (!fp) detail::fail("Error in writing to file '" + fname + "'.");
  int rflag = 1, cflag = 0;   // RTYPE, CNULL
  std::fwrite(&rflag, sizeof(int), 1, fp);
  std::fwrite(&cflag, sizeof(int), 1, fp);
  int namelen = (int)m_name.size();
  std::fwrite(&namelen, sizeof(int), 1, fp);
  if (namelen) std::fwrite(m_name.data(), 1, namelen, fp);
  std::fwrite(&status, sizeof(int), 1, fp);
  int bn = (int)bonds.size();
  std::fwrite(&bn, sizeof(int), 1, fp);
  size_t qsz = 12;
  std::fwrite(&qsz, sizeof(size_t), 1, fp);
  for (const Bond& b : bonds) {
    int tp = (int)b.type(), nq = (int)b.qnums().size();
    std::fwrite(&tp, sizeof(int), 1, fp);
    std::fwrite(&nq, sizeof(int), 1, fp);
    for (const Qnum& q : b.qnums()) {
      int t[3] = {q.U1(), (int)q.prt(), (int)q.prtF()};
      std::fwrite(t, sizeof(int), 3, fp);
    }
    std::fwrite(b.qdegs().data(), sizeof(int), nq, fp);
  }
  int nl = (int)labels.size();
  std::fwrite(&nl, sizeof(int), 1, fp);
  std::fwrite(labels.data(), sizeof(int), nl, fp);
  if (status & HAVEELEM) {
    size_t ne = m_elemNum;
    std::fwrite(&ne, sizeof(size_t), 1, fp);
    std::vector<double> h = getHostElem();
    std::fwrite(h.data(), sizeof(double), ne, fp);
  }
  std::fclose(fp);
}

static bool parse_tensor_file(const std::vector<char>& b, bool header, std::string& name, int& st,
                              std::vector<Bond>& bonds, std::vector<int>& labels, std::vector<double>& elem) {
  size_t o = 0;
  auto rd = [&](void* dst, size_t n) {
    if (o + n > b.size()) return false;
    std::memcpy(dst, b.data() + o, n);
    o += n;
    return true;
  };
  bonds.clear();
  labels.clear();
  elem.clear();
  if (header) {
    int rf, cf, nl;
    if (!rd(&rf, 4) || !rd(&cf, 4) || !rd(&nl, 4) || nl < 0 || nl > 4096) return false;
    name.assign(b.data() + o, std::min<size_t>(nl, b.size() - o));
    o += nl;
  }
  int bn;
  size_t qs;
  if (!rd(&st, 4) || !rd(&bn, 4) || !rd(&qs, 8) || qs != 12 || bn < 0 || bn > 64) return false;
  for (int i = 0; i < bn; ++i) {
    int tp, nq;
    if (!rd(&tp, 4) || !rd(&nq, 4) || nq <= 0 || nq > (1 << 20) || (tp != 1 && tp != -1)) return false;
    std::vector<Qnum> qn;
    std::vector<int> dg(nq);
    for (int q = 0; q < nq; ++q) {
      int t[3];
      if (!rd(t, 12)) return false;
      qn.push_back(Qnum((parityFType)t[2], t[0], (parityType)t[1]));
    }
    if (!rd(dg.data(), 4 * (size_t)nq)) return false;
    std::vector<Qnum> full;
    for (int q = 0; q < nq; ++q) full.insert(full.end(), dg[q], qn[q]);
    bonds.emplace_back((bondType)tp, full);
  }
  int nl;
  if (!rd(&nl, 4) || nl != bn) return false;
  labels.resize(nl);
  if (nl && !rd(labels.data(), 4 * (size_t)nl)) return false;
  if (st & 2) {
    size_t ne;
    if (!rd(&ne, 8) || ne > (b.size() - o) / 8) return false;
    elem.resize(ne);
    if (ne && !rd(elem.data(), 8 * ne)) return false;
  }
  return o == b.size();
}

UniTensor::UniTensor(const std::string& fname) {
  FILE* fp = std::fopen(fname.c_str(), "rb");
  if (!fp) detail::fail("Error in opening file '" + fname + "'.");
  std::vector<char> b;
  char chunk[65536];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) b.insert(b.end(), chunk, chunk + n);
  std::fclose(fp);
  std::string name;
  int st = 0;
  std::vector<Bond> bd;
  std::vector<int> lab;
  std::vector<double> el;
  bool ok = false;
  try {
    ok = parse_tensor_file(b, true, name, st, bd, lab, el);
  } catch (const std::exception&) {
    ok = false;
  }
  if (!ok) {
    try {
      ok = parse_tensor_file(b, false, name, st, bd, lab, el);
    } catch (const std::exception&) {
      ok = false;
    }
  }
  if (!ok) detail::fail("Error in reading file '" + fname + "' in.");
  m_name = name;
  bonds = bd;
  initUniT();
  labels = lab;
  if (st & HAVEELEM) {
    if (el.size() != m_elemNum) detail::fail("Error in reading file '" + fname + "': element count mismatch.");
    setElem(el);
  }
}

std::string UniTensor::printDiagram() const {
  std::ostringstream os;
  os << "***************** " << m_name << " *****************\n";
  for (size_t b = 0; b < bonds.size(); ++b)
    os << (bonds[b].type() == BD_IN ? "  in " : "  out") << " label " << labels[b] << " dim " << bonds[b].dim()
       << "\n";
  return os.str();
}

std::ostream& operator<<(std::ostream& os, const UniTensor& T) {
  os << T.printDiagram();
  os << "===============BLOCKS===============\n";
  for (auto& kv : T.getBlocks()) os << "--- " << kv.first << ": " << kv.second;
  os << "Total elemNum: " << T.m_elemNum << "\n";
  return os;
}

std::string UniTensor::profile(bool print) {
  std::ostringstream os;
  os << "\n===== Tensor profile =====\n"
     << "Existing Tensors: " << COUNTER << "\n"
     << "Allocated Elements: " << ELEMNUM << "\n"
     << "Max Allocated Elements: " << MAXELEMNUM << "\n"
     << "Max Allocated Elements for a Tensor: " << MAXELEMTEN << "\n"
     << "============================\n\n";
  if (print) std::cout << os.str();
  return os.str();
}

}  // namespace uni10
