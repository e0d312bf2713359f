// Qnum and Bond: charge algebra and run-length bonds.
// Follows uni10/src/uni10/datatype/lib/Qnum.cpp:33-100 and
// uni10/src/uni10/data-structure/lib/Bond.cpp:90-252.
#include <iomanip>
#include <iostream>

#include "common.h"

namespace uni10 {

static bool g_fermionic = false;

static void check_u1(int u1) {
  if (!(u1 < Qnum::U1_UPB && u1 > Qnum::U1_LOB)) {
    std::ostringstream err;
    err << "U1 is out of range. " << Qnum::U1_LOB << " < U1 < " << Qnum::U1_UPB << ".";
    detail::fail(err.str());
  }
}

Qnum::Qnum(int U1, parityType prt) : m_U1(U1), m_prt(prt), m_prtF(PRTF_EVEN) { check_u1(U1); }

Qnum::Qnum(parityFType prtF, int U1, parityType prt) : m_U1(U1), m_prt(prt), m_prtF(prtF) {
  check_u1(U1);
  if (prtF == PRTF_ODD) g_fermionic = true;
}

void Qnum::assign(int U1, parityType prt) {
  m_U1 = U1;
  m_prt = prt;
  m_prtF = PRTF_EVEN;
}

void Qnum::assign(parityFType prtF, int U1, parityType prt) {
  m_U1 = U1;
  m_prt = prt;
  m_prtF = prtF;
  if (prtF == PRTF_ODD) g_fermionic = true;
}

bool Qnum::isFermionic() { return g_fermionic; }

Qnum operator-(const Qnum& q) { return Qnum(q.m_prtF, -q.m_U1, q.m_prt); }

Qnum operator*(const Qnum& a, const Qnum& b) {
  return Qnum((parityFType)(a.m_prtF ^ b.m_prtF), a.m_U1 + b.m_U1, (parityType)(a.m_prt ^ b.m_prt));
}

std::ostream& operator<<(std::ostream& os, const Qnum& q) {
  os << "(U1 = " << q.m_U1 << ", P = " << (int)q.m_prt << ", " << (int)q.m_prtF << ")";
  return os;
}

// ---------------------------------------------------------------------------
Bond::Bond(bondType tp, size_t dim) : m_type(tp), m_dim(0) {
  setting(std::vector<Qnum>(dim, Qnum(0)));
}

Bond::Bond(bondType tp, const std::vector<Qnum>& qnums) : m_type(tp), m_dim(0) { setting(qnums); }

void Bond::assign(bondType tp, size_t dim) {
  m_type = tp;
  setting(std::vector<Qnum>(dim, Qnum(0)));
}

void Bond::assign(bondType tp, const std::vector<Qnum>& qnums) {
  m_type = tp;
  setting(qnums);
}

void Bond::setting(const std::vector<Qnum>& qnums) {
  if (qnums.empty()) detail::fail("Cannot create a bond of dimension 0.");
  Qnums.clear();
  Qdegs.clear();
  offsets.clear();
  m_dim = 0;
  for (size_t i = 0; i < qnums.size(); ++i) {
    if (i == 0 || !(qnums[i] == qnums[i - 1])) {
      Qnums.push_back(qnums[i]);
      Qdegs.push_back(1);
      offsets.push_back(m_dim);
    } else {
      Qdegs.back()++;
    }
    m_dim++;
  }
}

std::map<Qnum, int> Bond::degeneracy() const {
  std::map<Qnum, int> h;
  for (size_t i = 0; i < Qnums.size(); ++i) h[Qnums[i]] += Qdegs[i];
  return h;
}

std::vector<Qnum> Bond::Qlist() const {
  std::vector<Qnum> out;
  out.reserve(m_dim);
  for (size_t q = 0; q < Qnums.size(); ++q)
    for (int d = 0; d < Qdegs[q]; ++d) out.push_back(Qnums[q]);
  return out;
}

Bond& Bond::change(bondType tp) {
  if (m_type != tp) {
    for (auto& q : Qnums) q = -q;
    m_type = tp;
  }
  return *this;
}

Bond& Bond::dummy_change(bondType tp) {
  m_type = tp;
  return *this;
}

Bond& Bond::combine(Bond bd) {
  bd.change(m_type);
  std::vector<Qnum> qn;
  std::vector<int> qd, of;
  int dim = 0;
  for (size_t q = 0; q < Qnums.size(); ++q)
    for (int d = 0; d < Qdegs[q]; ++d)
      for (size_t qq = 0; qq < bd.Qnums.size(); ++qq) {
        Qnum fused = Qnums[q] * bd.Qnums[qq];
        int deg = bd.Qdegs[qq];
        if (qn.empty() || !(fused == qn.back())) {
          qn.push_back(fused);
          qd.push_back(deg);
          of.push_back(dim);
        } else {
          qd.back() += deg;
        }
        dim += deg;
      }
  Qnums.swap(qn);
  Qdegs.swap(qd);
  offsets.swap(of);
  m_dim = dim;
  return *this;
}

bool operator==(const Bond& a, const Bond& b) {
  return a.m_type == b.m_type && a.Qnums == b.Qnums && a.Qdegs == b.Qdegs;
}

std::ostream& operator<<(std::ostream& os, const Bond& b) {
  os << (b.m_type == BD_IN ? "IN : " : "OUT: ");
  for (size_t i = 0; i < b.Qnums.size(); ++i) os << b.Qnums[i] << "|" << b.Qdegs[i] << ", ";
  os << "Dim = " << b.m_dim << std::endl;
  return os;
}

Bond combine(bondType tp, const std::vector<Bond>& bds) {
  if (bds.empty()) detail::fail("There should be at least one bond in the input vector to be combined.");
  if (bds.size() == 1) {
    Bond b = bds[0];
    b.change(tp);
    return b;
  }
  const int n = (int)bds.size();
  Bond o1 = bds[n - 1];
  Bond o2 = bds[n - 2];
  o2.change(tp);
  o2.combine(o1);
  int b = 0;
  for (b = 0; b < n - 2; ++b) {
    if (b % 2 == 0) {
      o1 = bds[n - 3 - b];
      o1.change(tp);
      o1.combine(o2);
    } else {
      o2 = bds[n - 3 - b];
      o2.change(tp);
      o2.combine(o1);
    }
  }
  return (b % 2 == 0) ? o2 : o1;
}

Bond combine(const std::vector<Bond>& bds) {
  if (bds.empty()) detail::fail("There should be at least one bond in the input vector to be combined.");
  return combine(bds[0].type(), bds);
}

}  // namespace uni10
