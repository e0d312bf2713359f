// Batched iTEBD over the kernel C ABI.
//
// Dense chain (bond dimension c, physical dimension d, D = c d), per update,
// with X the Gamma*lambda tensor being rebuilt (ALa in update(A,B)), Y the
// other one, LX / LY the lambdas between and outside them (iTEBD.cpp:148-173
// with itebd.net, updateA.net):
//   T1[a,s,b,t]   = sum_m X[a,m,s] Y[m,b,t]                     GEMM  D x c x D
//   C[a,s',b,t']  = sum_{s,t} T1[a,s,b,t] G[(s,t),(s',t')]      GEMM  c^2 x d^2 x d^2
//   Theta         = LY[a] C[(a,s'),(b,t')]                      same launch (row-scaled second output)
//   Theta         = U S VT                                      SVD (D x D)
//   LX            = S[:c] / ||S[:c]||,  Y = VT[:c, :]           truncation
//   X[a,m',s']    = sum_{b,t'} C[a,s',b,t'] VT[m',(b,t')] / ||S[:c]||
//
// U(1) chain: the same update on block-sparse tensors.  Theta is
// block-diagonal in the charge q of the middle bond; block q has rows
// (s, a with q_a = q - s) and columns (t, b with q_b = q + t).  T1 is built
// per middle sector (one GEMM per (s, t)) into arrays keyed by (q_a, s + t),
// where the gate, which conserves s + t, is one GEMM per output (s', t')
// writing straight into theta's blocks.  Every block is factorised on its
// own (Block::svd per charge, BlockReal.cpp:226-250) and keeps its sector's
// number of singular values, so shapes never change; LX is normalised over
// all kept values of all blocks.
//
// Every stage is one launch over all chains: the GEMMs are grouped plans
// whose index maps read each operand in its stored layout, the SVD is a
// batched plan warm-started from the bond's previous VT, and the vector
// steps are vector-op plans.  All chains live in one arena.
#include "batch_tebd.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "common.h"
#include "tebd.h"

namespace uni10 {

namespace {
constexpr int kQmax = 12;       // U(1) charges (2 Sz) of a bond stay within [-12, 12]
constexpr int kMaxChi = 2048;

tn_index_map imap(std::initializer_list<std::pair<int64_t, int64_t>> lv) {
  tn_index_map m;
  std::memset(&m, 0, sizeof(m));
  m.n = (int32_t)lv.size();
  int l = 0;
  for (const auto& e : lv) {
    m.ext[l] = e.first;
    m.stride[l] = e.second;
    ++l;
  }
  return m;
}

tn_gemm_desc_ex desc(int64_t M, int64_t N, int64_t K, int64_t a_off, int64_t b_off, int64_t c_off) {
  tn_gemm_desc_ex d;
  std::memset(&d, 0, sizeof(d));
  d.M = M;
  d.N = N;
  d.K = K;
  d.a_off = a_off;
  d.b_off = b_off;
  d.c_off = c_off;
  d.lda = K;
  d.ldb = N;
  d.ldc = N;
  return d;
}

// 2 Sz of the physical states in basis order
std::vector<int> phys_charges(int d) { return d == 2 ? std::vector<int>{1, -1} : std::vector<int>{2, 0, -2}; }

// The sums s + t of two physical charges, descending, and for each the
// (s, t) index pairs with that sum, s ascending.
void sigma_combos(const std::vector<int>& phys, std::vector<int>& sig, std::vector<std::vector<int>>& cs) {
  const int d = (int)phys.size();
  sig.clear();
  for (int s = 0; s < d; ++s)
    for (int t = 0; t < d; ++t) sig.push_back(phys[s] + phys[t]);
  std::sort(sig.rbegin(), sig.rend());
  sig.erase(std::unique(sig.begin(), sig.end()), sig.end());
  cs.assign(sig.size(), {});
  for (size_t k = 0; k < sig.size(); ++k)
    for (int s = 0; s < d; ++s)
      for (int t = 0; t < d; ++t)
        if (phys[s] + phys[t] == sig[k]) cs[k].push_back(s);   // t follows from s
}

std::vector<int> largest_remainder(int total, const std::vector<double>& w) {
  const int n = (int)w.size();
  double W = 0.0;
  for (double v : w) W += v;
  std::vector<double> rem(n);
  std::vector<int64_t> base(n);
  int64_t sum = 0;
  for (int i = 0; i < n; ++i) {
    const double x = w[i] / W * (double)(total - n);
    base[i] = (int64_t)std::floor(x);
    rem[i] = x - (double)base[i];
    sum += base[i];
  }
  const int64_t shortfall = (int64_t)(total - n) - sum;
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rem[a] > rem[b]; });
  for (int64_t k = 0; k < shortfall && k < n; ++k) base[order[k]] += 1;
  std::vector<int> out(n);
  for (int i = 0; i < n; ++i) out[i] = (int)base[i] + 1;
  return out;
}
}  // namespace

int U1Sectors::find(int charge) const {
  const auto it = std::lower_bound(q.begin(), q.end(), charge);
  return (it != q.end() && *it == charge) ? (int)(it - q.begin()) : -1;
}

void u1_bond_sectors(int chi, int d, int qmax, U1Sectors& bond0, U1Sectors& bond1) {
  const std::vector<int> phys = phys_charges(d);
  const double sigma = 2.0 * (1.0 + std::log2((double)chi) / 4.0);
  auto pick = [&](std::vector<int> cands, U1Sectors& out) {
    std::sort(cands.begin(), cands.end());
    cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
    std::vector<double> w(cands.size());
    double W = 0.0;
    for (size_t i = 0; i < cands.size(); ++i) {
      const double c = (double)cands[i];
      w[i] = std::exp(-(c * c) / (2.0 * sigma * sigma));
      W += w[i];
    }
    const size_t imax = std::max_element(w.begin(), w.end()) - w.begin();
    std::vector<int> kq;
    std::vector<double> kw;
    for (size_t i = 0; i < cands.size(); ++i)
      if ((double)chi * w[i] / W >= 0.5 || i == imax) {
        kq.push_back(cands[i]);
        kw.push_back(w[i]);
      }
    if ((int)kq.size() > chi) {
      std::vector<int> order(kq.size());
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return kw[a] > kw[b]; });
      order.resize(chi);
      std::sort(order.begin(), order.end());
      std::vector<int> q2;
      std::vector<double> w2;
      for (int o : order) {
        q2.push_back(kq[o]);
        w2.push_back(kw[o]);
      }
      kq = q2;
      kw = w2;
    }
    out.q = kq;
    out.n = largest_remainder(chi, kw);
  };
  std::vector<int> c1, c0;
  for (int q = -qmax; q <= qmax; ++q)
    if (q % 2 == 0) c1.push_back(q);
  pick(c1, bond1);
  for (int q : bond1.q)
    for (int s : phys)
      if (std::abs(q + s) <= qmax) c0.push_back(q + s);
  pick(c0, bond0);
  U1Sectors* b[2] = {&bond0, &bond1};
  for (int it = 0; it < 64; ++it) {
    bool changed = false;
    for (int p = 0; p < 2; ++p) {
      U1Sectors& R = *b[p];
      const U1Sectors& L = *b[1 - p];
      std::vector<int> room(R.q.size(), 0);
      for (size_t i = 0; i < R.q.size(); ++i)
        for (int s : phys) {
          const int j = L.find(R.q[i] - s);
          if (j >= 0) room[i] += L.n[j];
        }
      for (size_t i = 0; i < R.q.size(); ++i)
        while (R.n[i] > room[i]) {
          R.n[i] -= 1;
          size_t j = 0;
          for (size_t k = 1; k < R.q.size(); ++k) {
            const long sk = room[k] - R.n[k], sj = room[j] - R.n[j];
            if (sk > sj || (sk == sj && (std::abs(R.q[k]) < std::abs(R.q[j]) ||
                                         (std::abs(R.q[k]) == std::abs(R.q[j]) && R.q[k] < R.q[j]))))
              j = k;
          }
          if (room[j] - R.n[j] <= 0) detail::fail("u1_bond_sectors: chi does not fit the charge range");
          R.n[j] += 1;
          changed = true;
        }
    }
    if (!changed) break;
  }
  for (U1Sectors* s : b) {
    s->off.assign(s->q.size(), 0);
    for (size_t i = 1; i < s->q.size(); ++i) s->off[i] = s->off[i - 1] + s->n[i - 1];
  }
}

void BatchTebd::layout_u1(int i) {
  U1Chain& u = u1_[i];
  const int d = d_[i];
  u.phys = phys_charges(d);
  u1_bond_sectors(chi_[i], d, kQmax, u.bond[0], u.bond[1]);
  std::vector<int> sig;
  std::vector<std::vector<int>> cs;
  sigma_combos(u.phys, sig, cs);
  for (int p = 0; p < 2; ++p) {
    const U1Sectors& L = u.bond[1 - p];
    const U1Sectors& R = u.bond[p];
    // tensor p: blocks by left-bond sector
    int64_t o = 0;
    const size_t nL = L.q.size(), nR = R.q.size();
    u.xblk[p].assign(nL, 0);
    u.xw[p].assign(nL, 0);
    u.xpart[p].assign(nL, std::vector<int64_t>(d, -1));
    for (size_t a = 0; a < nL; ++a) {
      int64_t w = 0;
      for (int s = 0; s < d; ++s) {
        const int j = R.find(L.q[a] + u.phys[s]);
        if (j >= 0) {
          u.xpart[p][a][s] = w;
          w += R.n[j];
        }
      }
      u.xblk[p][a] = o;
      u.xw[p][a] = w;
      o += (int64_t)L.n[a] * w;
    }
    u.xsize[p] = o;
    // update p: theta blocks by sector of bond p
    u.D[p].assign(nR, 0);
    u.sq[p].assign(nR, 0);
    u.sv[p].assign(nR, 0);
    u.disc[p].assign(nR, 0);
    u.part[p].assign(nR, std::vector<int64_t>(2 * d, -1));
    int64_t sqo = 0, svo = 0, dio = 0;
    for (size_t j = 0; j < nR; ++j) {
      const int q = R.q[j];
      int64_t D = 0, Dc = 0;
      for (int s = 0; s < d; ++s) {
        const int a = L.find(q - u.phys[s]);
        if (a >= 0) {
          u.part[p][j][s] = D;
          D += L.n[a];
        }
        const int b = L.find(q + u.phys[s]);
        if (b >= 0) {
          u.part[p][j][d + s] = Dc;
          Dc += L.n[b];
        }
      }
      if (D != Dc) detail::fail("BatchTebd: theta block is not square");
      if (D < R.n[j]) detail::fail("BatchTebd: a sector is larger than its theta block");
      u.D[p][j] = D;
      u.sq[p][j] = sqo;
      u.sv[p][j] = svo;
      u.disc[p][j] = dio;
      sqo += D * D;
      svo += D;
      dio += D - R.n[j];
    }
    u.sqsize[p] = sqo;
    u.svsize[p] = svo;
    u.tmpsize[p] = R.total() + dio;
    // T1 arrays of (outer sector a, sigma): rows (a, b in sector q_a + sigma), combos fastest
    const size_t ns = sig.size();
    u.pair[p].assign(nL * ns, -1);
    int64_t t1o = 0;
    for (size_t a = 0; a < nL; ++a)
      for (size_t k = 0; k < ns; ++k) {
        const int b = L.find(L.q[a] + sig[k]);
        if (b < 0) continue;
        u.pair[p][a * ns + k] = t1o;
        t1o += (int64_t)L.n[a] * L.n[b] * (int64_t)cs[k].size();
      }
    u.t1size[p] = t1o;
  }
  // the block rows of VT are the blocks of the tensor they become
  for (int p = 0; p < 2; ++p)
    for (size_t j = 0; j < u.bond[p].q.size(); ++j)
      if (u.xw[1 - p][j] != u.D[p][j]) detail::fail("BatchTebd: VT block and tensor block differ");
}

BatchTebd::BatchTebd(const std::vector<int>& chi, const std::vector<double>& delta,
                     const std::vector<uint64_t>& seed, double tau, int gemm_mode, const std::string& netdir,
                     const std::vector<int>& d, const std::vector<int>& sym)
    : chi_(chi), delta_(delta), netdir_(detail::netdir_or_default(netdir)) {
  const size_t nc = chi.size();
  if (delta.size() != nc || seed.size() != nc) detail::fail("BatchTebd: chi, delta and seed differ in length");
  if (!d.empty() && d.size() != nc) detail::fail("BatchTebd: d and chi differ in length");
  if (!sym.empty() && sym.size() != nc) detail::fail("BatchTebd: sym and chi differ in length");
  if (nc == 0) detail::fail("BatchTebd: no chains");
  d_ = d.empty() ? std::vector<int>(nc, 2) : d;
  sym_ = sym.empty() ? std::vector<int>(nc, 0) : sym;
  for (size_t i = 0; i < nc; ++i) {
    if (chi[i] < 1 || chi[i] > kMaxChi) detail::fail("BatchTebd: chi must be in [1, 2048]");
    if (d_[i] != 2 && d_[i] != 3) detail::fail("BatchTebd: d must be 2 (spin 1/2) or 3 (spin 1)");
    if (sym_[i] != 0 && sym_[i] != 1) detail::fail("BatchTebd: sym must be 0 (dense) or 1 (U(1))");
  }
  // a failure part-way (the device memory budget, say) returns everything
  // taken so far before the error reaches the caller
  try {
    build(seed, tau, gemm_mode);
  } catch (...) {
    release();
    throw;
  }
}

void BatchTebd::build(const std::vector<uint64_t>& seed, double tau, int gemm_mode) {
  const size_t nc = chi_.size();
  const std::vector<int>& chi = chi_;
  const std::vector<double>& delta = delta_;
  // the engine's own stream (default priority): independent engines overlap
  detail::check(tn_stream_create(&own_stream_), "BatchTebd stream");
  stream_ = own_stream_;
  u1_.resize(nc);
  for (size_t i = 0; i < nc; ++i)
    if (sym_[i]) layout_u1((int)i);
  // arena layout
  int64_t o = 0;
  auto take = [&](int64_t n) {
    const int64_t r = o;
    o += (n + 1) & ~int64_t(1);   // keep every region 16-byte aligned
    return r;
  };
  off_.resize(nc);
  for (size_t i = 0; i < nc; ++i) {
    const int64_t c = chi[i], dd = d_[i], D = c * dd, t = c * c * dd * dd;
    Off& f = off_[i];
    if (!sym_[i]) {
      f.x[0] = take(c * c * dd);
      f.x[1] = take(c * c * dd);
      f.l[0] = take(c);
      f.l[1] = take(c);
      f.g = take(dd * dd * dd * dd);
      f.t1[0] = f.t1[1] = take(t);
      f.cb[0] = f.cb[1] = take(t);
      f.th[0] = f.th[1] = take(t);
      f.s[0] = f.s[1] = take(D);
      f.vt[0] = take(D * D);
      f.vt[1] = take(D * D);
      f.sc[0] = take(2);
      f.sc[1] = take(2);
      f.tmp[0] = f.tmp[1] = -1;
    } else {
      const U1Chain& u = u1_[i];
      f.x[0] = take(u.xsize[0]);
      f.x[1] = take(u.xsize[1]);
      f.l[0] = take(c);
      f.l[1] = take(c);
      f.g = take(dd * dd * dd * dd);
      for (int p = 0; p < 2; ++p) {
        f.t1[p] = take(u.t1size[p]);
        f.cb[p] = take(u.sqsize[p]);
        f.th[p] = take(u.sqsize[p]);
        f.s[p] = take(u.svsize[p]);
        f.vt[p] = take(u.sqsize[p]);
        f.tmp[p] = take(u.tmpsize[p]);
        f.sc[p] = take(2);
      }
    }
  }
  arena_.reset((size_t)o, true);
  detail::sync();   // the arena is cleared on the caller's stream

  // initial state of every chain and its gate
  tau_ = tau;
  std::vector<double> host((size_t)o, 0.0);
  std::vector<int> all(nc);
  for (size_t i = 0; i < nc; ++i) all[i] = (int)i;
  const std::vector<std::vector<double>> g = gates(all, delta);
  for (size_t i = 0; i < nc; ++i) {
    const int64_t end = (i + 1 < nc) ? off_[i + 1].x[0] : o;
    region_.push_back(end - off_[i].x[0]);
    init_chain((int)i, delta[i], seed[i], host.data() + off_[i].x[0], &g[i]);
  }

  // every chain starts active (device flags read by every launch)
  active_host_.assign(nc, 1);
  detail::check(tn_malloc(reinterpret_cast<void**>(&active_), sizeof(int) * nc), "BatchTebd active");
  detail::check(tn_memcpy_h2d(active_, active_host_.data(), sizeof(int) * nc, stream_), "BatchTebd active");

  // plans, one set per update parity p (p = 0: X = ALa, Y = BLb; p = 1: swapped)
  for (int p = 0; p < 2; ++p) {
    std::vector<tn_gemm_desc_ex> d1, d2, d3;
    std::vector<int64_t> svd[7];   // m, n, a, u, s, vt, v0
    std::vector<tn_vecop> gather, trunc, scale;
    // the chain of every GEMM group, SVD problem and vector item (for the mask)
    std::vector<int32_t> c1, c2, c3, cs, cgather, ctrunc, cscale;
    for (size_t i = 0; i < nc; ++i) {
      if (sym_[i])
        plans_u1((int)i, p, d1, d2, d3, gather, trunc, scale, svd);
      else
        plans_dense((int)i, p, d1, d2, d3, gather, trunc, scale, svd);
      c1.resize(d1.size(), (int32_t)i);
      c2.resize(d2.size(), (int32_t)i);
      c3.resize(d3.size(), (int32_t)i);
      cs.resize(svd[0].size(), (int32_t)i);
      cgather.resize(gather.size(), (int32_t)i);
      ctrunc.resize(trunc.size(), (int32_t)i);
      cscale.resize(scale.size(), (int32_t)i);
    }
    auto gemm = [&](tn_gemm_plan* g, const std::vector<tn_gemm_desc_ex>& ds, const char* what) {
      detail::check(tn_gemm_plan_create_ex(g, gemm_mode, (int64_t)ds.size(), ds.data()), what);
    };
    gemm(&g1_[p], d1, "BatchTebd gemm1");
    gemm(&g2_[p], d2, "BatchTebd gemm2");
    gemm(&g3_[p], d3, "BatchTebd gemm3");
    // tile shapes picked by timing each plan on the (still zero) arena; every
    // shape gives the same bits, so the choice never changes a result
    for (tn_gemm_plan g : {g1_[p], g2_[p], g3_[p]}) {
      float ms = 0;
      double* A = arena_.get();
      detail::check(tn_gemm_plan_autotune(g, A, A, A, 2, stream_, &ms), "BatchTebd autotune");
    }
    const int64_t ns = (int64_t)svd[0].size();
    // every SVD starts from the VT of the previous update on its bond, in
    // place; a chain's start (and a reloaded slot) holds VT = I, so its first
    // update is a cold start through the same launches
    detail::check(tn_svd_plan_create(&svd_[p], ns, svd[0].data(), svd[1].data(), svd[2].data(), svd[6].data(),
                                     svd[3].data(), svd[4].data(), svd[5].data()),
                  "BatchTebd svd");
    auto vec = [&](tn_vecops_plan* v, const std::vector<tn_vecop>& items, const char* what) {
      detail::check(tn_vecops_plan_create(v, (int64_t)items.size(), items.data()), what);
    };
    vec(&gather_[p], gather, "BatchTebd gather");
    vec(&trunc_[p], trunc, "BatchTebd trunc");
    vec(&scale_[p], scale, "BatchTebd scale");
    // an inactive chain's work is skipped on the device, its state untouched
    detail::check(tn_gemm_plan_set_mask(g1_[p], active_, c1.data()), "BatchTebd mask");
    detail::check(tn_gemm_plan_set_mask(g2_[p], active_, c2.data()), "BatchTebd mask");
    detail::check(tn_gemm_plan_set_mask(g3_[p], active_, c3.data()), "BatchTebd mask");
    detail::check(tn_svd_plan_set_mask(svd_[p], active_, cs.data()), "BatchTebd mask");
    detail::check(tn_vecops_plan_set_mask(gather_[p], active_, cgather.data()), "BatchTebd mask");
    detail::check(tn_vecops_plan_set_mask(trunc_[p], active_, ctrunc.data()), "BatchTebd mask");
    detail::check(tn_vecops_plan_set_mask(scale_[p], active_, cscale.data()), "BatchTebd mask");
  }
  detail::check(tn_memcpy_h2d(arena_.get(), host.data(), sizeof(double) * (size_t)o, stream_), "BatchTebd upload");
}

void BatchTebd::plans_dense(int i, int p, std::vector<tn_gemm_desc_ex>& d1, std::vector<tn_gemm_desc_ex>& d2,
                            std::vector<tn_gemm_desc_ex>& d3, std::vector<tn_vecop>&,
                            std::vector<tn_vecop>& trunc, std::vector<tn_vecop>& scale, std::vector<int64_t>* svd) {
  const int64_t c = chi_[i], kd = d_[i], D = c * kd, dd = kd * kd;
  const Off& f = off_[i];
  const int64_t X = f.x[p], Y = f.x[1 - p], LX = f.l[p], LY = f.l[1 - p];
  // T1[(a,s),(b,t)] = sum_m X[a,m,s] Y[m,(b,t)]
  tn_gemm_desc_ex g1 = desc(D, D, c, X, Y, f.t1[p]);
  g1.a_row = imap({{c, c * kd}, {kd, 1}});
  g1.a_col = imap({{c, kd}});
  d1.push_back(g1);
  // C[a,s',b,t'] = sum_{s,t} T1[a,s,b,t] G[(s,t),(s',t')]; rows (a,b)
  tn_gemm_desc_ex g2 = desc(c * c, dd, dd, f.t1[p], f.g, f.cb[p]);
  g2.a_row = imap({{c, kd * c * kd}, {c, kd}});
  g2.a_col = imap({{kd, c * kd}, {kd, 1}});
  g2.c_row = imap({{c, kd * c * kd}, {c, kd}});
  g2.c_col = imap({{kd, c * kd}, {kd, 1}});
  // and, in the same launch, Theta = LY[a] C[(a,s'),(b,t')] (row m = (a, b))
  g2.d_off = f.th[p];
  g2.rs_off = LY;
  g2.rs_div = c;
  d2.push_back(g2);
  // SVD of Theta (D x D) into U, S, VT[p]
  const int64_t sv[7] = {D, D, f.th[p], -1, f.s[p], f.vt[p], f.vt[p]};   // U is not formed
  for (int k = 0; k < 7; ++k) svd[k].push_back(sv[k]);
  // LX = S[:c]/||S[:c]||, sc = {norm, discarded}; Y = VT[:c, :]
  trunc.push_back(tn_vecop{TN_VECOP_TRUNCATE, 0, D, f.s[p], LX, f.sc[p], c});
  trunc.push_back(tn_vecop{TN_VECOP_COPY, 0, c * D, f.vt[p], Y, 0, 0});
  // X[a,m',s'] = sum_{b,t'} C[a,s',b,t'] VT[m',(b,t')]
  tn_gemm_desc_ex g3 = desc(D, c, D, f.cb[p], f.vt[p], X);
  g3.b_row = imap({{D, 1}});
  g3.b_col = imap({{c, D}});
  g3.c_row = imap({{c, c * kd}, {kd, 1}});
  g3.c_col = imap({{c, kd}});
  d3.push_back(g3);
  // X /= ||S[:c]||
  scale.push_back(tn_vecop{TN_VECOP_SCALE_INV, 0, c * c * kd, X, X, f.sc[p], 0});
  if (p == 0) flops_ += 2.0 * (2.0 * D * D * c + 2.0 * c * c * dd * dd + 2.0 * D * c * D);
}

void BatchTebd::plans_u1(int i, int p, std::vector<tn_gemm_desc_ex>& d1, std::vector<tn_gemm_desc_ex>& d2,
                         std::vector<tn_gemm_desc_ex>& d3, std::vector<tn_vecop>& gather,
                         std::vector<tn_vecop>& trunc, std::vector<tn_vecop>& scale, std::vector<int64_t>* svd) {
  const U1Chain& u = u1_[i];
  const Off& f = off_[i];
  const int d = d_[i];
  const U1Sectors& L = u.bond[1 - p];
  const U1Sectors& R = u.bond[p];
  std::vector<int> sig;
  std::vector<std::vector<int>> cs;
  sigma_combos(u.phys, sig, cs);
  const size_t ns = sig.size();
  auto sig_index = [&](int sum) { return (size_t)(std::find(sig.begin(), sig.end(), sum) - sig.begin()); };
  double fl = 0.0;
  // T1 of middle sector j: rows (s, a), columns (t, b), into the (a, sigma) arrays
  for (size_t j = 0; j < R.q.size(); ++j) {
    const int q = R.q[j];
    for (int s = 0; s < d; ++s) {
      const int a = L.find(q - u.phys[s]);
      if (a < 0) continue;
      for (int t = 0; t < d; ++t) {
        const int b = L.find(q + u.phys[t]);
        if (b < 0) continue;
        const size_t k = sig_index(u.phys[s] + u.phys[t]);
        const int64_t nc = (int64_t)cs[k].size();
        const int64_t combo = std::find(cs[k].begin(), cs[k].end(), s) - cs[k].begin();
        const int64_t M = L.n[a], N = L.n[b], K = R.n[j];
        tn_gemm_desc_ex g = desc(M, N, K, f.x[p] + u.xblk[p][a] + u.xpart[p][a][s],
                                 f.x[1 - p] + u.xblk[1 - p][j] + u.xpart[1 - p][j][t],
                                 f.t1[p] + u.pair[p][a * ns + k] + combo);
        g.lda = u.xw[p][a];
        g.ldb = u.xw[1 - p][j];
        g.c_row = imap({{M, N * nc}});
        g.c_col = imap({{N, nc}});
        d1.push_back(g);
        fl += 2.0 * M * N * K;
      }
    }
  }
  // gate: C block (s', t') of theta block q_a + s' from the (a, sigma) array
  for (size_t a = 0; a < L.q.size(); ++a)
    for (size_t k = 0; k < ns; ++k) {
      const int64_t po = u.pair[p][a * ns + k];
      if (po < 0) continue;
      const int b = L.find(L.q[a] + sig[k]);
      const int64_t nc = (int64_t)cs[k].size();
      // input combo c is (s0 + c, t0 - c): G rows in0 + c (d - 1)
      const int s0 = cs[k][0];
      int t0 = 0;
      while (u.phys[s0] + u.phys[t0] != sig[k]) ++t0;
      const int64_t in0 = (int64_t)s0 * d + t0;
      for (int so : cs[k]) {
        int to = 0;
        while (u.phys[so] + u.phys[to] != sig[k]) ++to;
        const int j = R.find(L.q[a] + u.phys[so]);
        if (j < 0) continue;   // the charge lies outside the middle bond's sectors
        const int64_t D = u.D[p][j];
        const int64_t M = (int64_t)L.n[a] * L.n[b];
        tn_gemm_desc_ex g = desc(M, 1, nc, f.t1[p] + po, f.g + in0 * d * d + (int64_t)so * d + to,
                                 f.cb[p] + u.sq[p][j] + u.part[p][j][so] * D + u.part[p][j][d + to]);
        g.lda = nc;
        g.b_row = imap({{nc, (int64_t)(d - 1) * d * d}});
        g.b_col = imap({{1, 1}});
        g.c_row = imap({{L.n[a], D}, {L.n[b], 1}});
        g.c_col = imap({{1, 1}});
        // and theta = LY[a] C in the same launch: theta row part(so) + a_idx,
        // a_idx = m / n_b
        g.d_off = g.c_off - f.cb[p] + f.th[p];
        g.rs_off = f.l[1 - p] + L.off[a];
        g.rs_div = L.n[b];
        d2.push_back(g);
        fl += 2.0 * M * nc;
      }
    }
  int64_t xn = 0;
  for (size_t j = 0; j < R.q.size(); ++j) {
    const int q = R.q[j];
    const int64_t D = u.D[p][j], k = R.n[j], sq = u.sq[p][j];
    const int64_t sv[7] = {D, D, f.th[p] + sq, -1, f.s[p] + u.sv[p][j], f.vt[p] + sq, f.vt[p] + sq};
    for (int e = 0; e < 7; ++e) svd[e].push_back(sv[e]);
    // kept singular values to the sector's place, the rest after all kept
    // ones; kept VT rows are the new block j of the other tensor
    gather.push_back(tn_vecop{TN_VECOP_COPY, 0, k, f.s[p] + u.sv[p][j], f.tmp[p] + R.off[j], 0, 0});
    if (D > k)
      gather.push_back(
          tn_vecop{TN_VECOP_COPY, 0, D - k, f.s[p] + u.sv[p][j] + k, f.tmp[p] + R.total() + u.disc[p][j], 0, 0});
    gather.push_back(tn_vecop{TN_VECOP_COPY, 0, k * D, f.vt[p] + sq, f.x[1 - p] + u.xblk[1 - p][j], 0, 0});
    // X block of (s, a): C rows of part s times VT^T
    for (int s = 0; s < d; ++s) {
      const int a = L.find(q - u.phys[s]);
      if (a < 0) continue;
      const int64_t M = L.n[a];
      tn_gemm_desc_ex g = desc(M, k, D, f.cb[p] + sq + u.part[p][j][s] * D, f.vt[p] + sq,
                               f.x[p] + u.xblk[p][a] + u.xpart[p][a][s]);
      g.lda = D;
      g.b_row = imap({{D, 1}});
      g.b_col = imap({{k, D}});
      g.ldc = u.xw[p][a];
      d3.push_back(g);
      fl += 2.0 * M * k * D;
      xn += M * k;
    }
  }
  trunc.push_back(tn_vecop{TN_VECOP_TRUNCATE, 0, u.tmpsize[p], f.tmp[p], f.l[p], f.sc[p], R.total()});
  scale.push_back(tn_vecop{TN_VECOP_SCALE_INV, 0, u.xsize[p], f.x[p], f.x[p], f.sc[p], 0});
  if (xn != u.xsize[p]) detail::fail("BatchTebd: U(1) update does not rebuild its whole tensor");
  flops_ += fl;
}

// exp(-tau H(delta_j)) of chains[j], all in one batched eigendecomposition
// (H = U^T diag(lambda) U with U the eigenvector rows, as exph,
// MatrixTools.cpp:109-121); the products exp(-tau H) = U^T diag(e) U are
// formed on the host in a fixed order, so a chain's gate is the same bits
// whether it is built with the others or alone (a reloaded slot).
std::vector<std::vector<double>> BatchTebd::gates(const std::vector<int>& chains,
                                                  const std::vector<double>& delta) const {
  const size_t nb = chains.size();
  std::vector<int64_t> n(nb), ao(nb), eo(nb), vo(nb);
  std::vector<double> hs;
  int64_t ne = 0, nv = 0;
  for (size_t j = 0; j < nb; ++j) {
    const int kd = d_[chains[j]];
    const std::vector<double> h = (kd == 2) ? xxz_two_site_host(delta[j], 1.0) : xxz_two_site_spin1_host(delta[j], 1.0);
    n[j] = kd * kd;
    ao[j] = (int64_t)hs.size();
    eo[j] = ne;
    vo[j] = nv;
    hs.insert(hs.end(), h.begin(), h.end());
    ne += n[j];
    nv += n[j] * n[j];
  }
  std::vector<std::vector<double>> out(nb);
  if (nb == 0) return out;
  DevBuf dh((size_t)hs.size()), de((size_t)ne), dv((size_t)nv);
  detail::check(tn_memcpy_h2d(dh.get(), hs.data(), sizeof(double) * hs.size(), own_stream_), "BatchTebd gates");
  detail::check(tn_eigh_f64_batched((int64_t)nb, n.data(), dh.get(), ao.data(), de.get(), eo.data(), dv.get(),
                                    vo.data(), own_stream_),
                "BatchTebd gates");
  std::vector<double> he((size_t)ne), hv((size_t)nv);
  detail::check(tn_memcpy_d2h(he.data(), de.get(), sizeof(double) * he.size(), own_stream_), "BatchTebd gates");
  detail::check(tn_memcpy_d2h(hv.data(), dv.get(), sizeof(double) * hv.size(), own_stream_), "BatchTebd gates");
  for (size_t j = 0; j < nb; ++j) {
    const int64_t m = n[j];
    const double* lam = he.data() + eo[j];
    const double* U = hv.data() + vo[j];   // row k: eigenvector k
    std::vector<double> ex((size_t)m), g((size_t)(m * m), 0.0);
    for (int64_t k = 0; k < m; ++k) ex[k] = std::exp(-tau_ * lam[k]);
    for (int64_t r = 0; r < m; ++r)
      for (int64_t c = 0; c < m; ++c) {
        double acc = 0.0;
        for (int64_t k = 0; k < m; ++k) acc += U[k * m + r] * (ex[k] * U[k * m + c]);
        g[r * m + c] = acc;
      }
    out[j] = std::move(g);
  }
  return out;
}

// Host image of chain i's arena region (base = its x[0]): Gamma lambda
// tensors and lambdas from the seeded start, the gate exp(-tau H(delta)),
// and identity warm-start bases (so a chain's SVD sequence never depends on
// an earlier occupant).  A dense chain starts as ITebd (splitmix seeds
// seed..seed+3); a U(1) chain from the same dense Gammas restricted to its
// charge structure, with each lambda sorted within its sectors.
void BatchTebd::init_chain(int i, double delta, uint64_t seed, double* base, const std::vector<double>* gate) {
  const int c = chi_[i], kd = d_[i];
  const Off& f = off_[i];
  auto at = [&](int64_t off) { return base + (off - f.x[0]); };
  std::fill(base, base + region_[i], 0.0);
  std::vector<double> la(c), lb(c);
  for (int e = 0; e < c; ++e) {
    la[e] = std::fabs(detail::splitmix_uniform(seed + 2, e + 1));
    lb[e] = std::fabs(detail::splitmix_uniform(seed + 3, e + 1));
  }
  if (!sym_[i]) {
    std::sort(la.rbegin(), la.rend());
    std::sort(lb.rbegin(), lb.rend());
  } else {
    const U1Chain& u = u1_[i];
    for (size_t k = 0; k < u.bond[0].q.size(); ++k)
      std::sort(la.begin() + u.bond[0].off[k], la.begin() + u.bond[0].off[k] + u.bond[0].n[k], std::greater<double>());
    for (size_t k = 0; k < u.bond[1].q.size(); ++k)
      std::sort(lb.begin() + u.bond[1].off[k], lb.begin() + u.bond[1].off[k] + u.bond[1].n[k], std::greater<double>());
  }
  double na = 0, nb = 0;
  for (int e = 0; e < c; ++e) {
    na += la[e] * la[e];
    nb += lb[e] * lb[e];
  }
  na = std::sqrt(na);
  nb = std::sqrt(nb);
  double* l0 = at(f.l[0]);
  double* l1 = at(f.l[1]);
  for (int e = 0; e < c; ++e) {
    l0[e] = la[e] / na;
    l1[e] = lb[e] / nb;
  }
  const double* lam[2] = {l0, l1};
  if (!sym_[i]) {
    double* x0 = at(f.x[0]);
    double* x1 = at(f.x[1]);
    for (int a = 0; a < c; ++a)
      for (int s = 0; s < kd; ++s)
        for (int b = 0; b < c; ++b) {
          const uint64_t e = ((uint64_t)a * kd + s) * c + b + 1;
          x0[((size_t)a * c + b) * kd + s] = detail::splitmix_uniform(seed, e) * l0[b];
          x1[((size_t)a * c + b) * kd + s] = detail::splitmix_uniform(seed + 1, e) * l1[b];
        }
    const int64_t D = (int64_t)c * kd;
    for (int p = 0; p < 2; ++p)
      for (int64_t e = 0; e < D; ++e) at(f.vt[p])[e * D + e] = 1.0;
  } else {
    const U1Chain& u = u1_[i];
    for (int p = 0; p < 2; ++p) {
      const U1Sectors& L = u.bond[1 - p];
      const U1Sectors& R = u.bond[p];
      double* x = at(f.x[p]);
      for (size_t a = 0; a < L.q.size(); ++a)
        for (int s = 0; s < kd; ++s) {
          if (u.xpart[p][a][s] < 0) continue;
          const int j = R.find(L.q[a] + u.phys[s]);
          for (int ar = 0; ar < L.n[a]; ++ar)
            for (int mr = 0; mr < R.n[j]; ++mr) {
              const int64_t ga = L.off[a] + ar, gm = R.off[j] + mr;
              const uint64_t e = ((uint64_t)ga * kd + s) * c + gm + 1;
              x[u.xblk[p][a] + ar * u.xw[p][a] + u.xpart[p][a][s] + mr] =
                  detail::splitmix_uniform(seed + p, e) * lam[p][gm];
            }
        }
      for (size_t j = 0; j < R.q.size(); ++j) {
        const int64_t D = u.D[p][j];
        double* v = at(f.vt[p]) + u.sq[p][j];
        for (int64_t e = 0; e < D; ++e) v[e * D + e] = 1.0;
      }
    }
  }
  const std::vector<double> g = gate ? *gate : gates({i}, {delta})[0];
  std::copy(g.begin(), g.end(), at(f.g));
  delta_[i] = delta;
}

void BatchTebd::load(int chain, double delta, uint64_t seed) {
  if (chain < 0 || chain >= chains()) detail::fail("BatchTebd::load: bad index");
  std::vector<double> host((size_t)region_[chain]);
  init_chain(chain, delta, seed, host.data());
  // on the engine's own stream: the copy is complete when this returns, and
  // a stream shared with other engines' steps is not waited for
  detail::check(tn_memcpy_h2d(arena_.get() + off_[chain].x[0], host.data(), sizeof(double) * host.size(),
                              own_stream_),
                "BatchTebd load");
}

BatchTebd::~BatchTebd() { release(); }

void BatchTebd::release() noexcept {
  if (stream_) tn_stream_sync(stream_);
  if (own_stream_ && own_stream_ != stream_) tn_stream_sync(own_stream_);
  if (graph_) tn_graph_destroy(graph_);
  graph_ = nullptr;
  for (int p = 0; p < 2; ++p) {
    tn_gemm_plan_destroy(g1_[p]);
    tn_gemm_plan_destroy(g2_[p]);
    tn_gemm_plan_destroy(g3_[p]);
    tn_svd_plan_destroy(svd_[p]);
    tn_vecops_plan_destroy(gather_[p]);
    tn_vecops_plan_destroy(trunc_[p]);
    tn_vecops_plan_destroy(scale_[p]);
    g1_[p] = g2_[p] = g3_[p] = nullptr;
    svd_[p] = nullptr;
    gather_[p] = trunc_[p] = scale_[p] = nullptr;
  }
  tn_free(active_);
  active_ = nullptr;
  tn_stream_destroy(own_stream_);
  own_stream_ = stream_ = nullptr;
}

void BatchTebd::set_stream(void* stream) {
  detail::check(tn_stream_sync(stream_), "BatchTebd set_stream");
  stream_ = stream ? stream : own_stream_;
  active_dirty_ = true;     // the flags are uploaded again, in the new stream's order
}

void BatchTebd::update(int p) {
  double* A = arena_.get();
  detail::check(tn_gemm_exec(g1_[p], A, A, A, stream_), "BatchTebd gemm1");
  detail::check(tn_gemm_exec(g2_[p], A, A, A, stream_), "BatchTebd gemm2");
  last_svd_ = svd_[p];
  detail::check(tn_svd_plan_exec(last_svd_, A, A, A, A, A, stream_), "BatchTebd svd");
  detail::check(tn_vecops_plan_exec(gather_[p], A, stream_), "BatchTebd gather");
  detail::check(tn_vecops_plan_exec(trunc_[p], A, stream_), "BatchTebd trunc");
  detail::check(tn_gemm_exec(g3_[p], A, A, A, stream_), "BatchTebd gemm3");
  detail::check(tn_vecops_plan_exec(scale_[p], A, stream_), "BatchTebd scale");
}

void BatchTebd::run_step() {
  update(0);
  update(1);
  warm_[0] = warm_[1] = true;
}

void BatchTebd::set_graph(bool on) {
  use_graph_ = on;
  if (!on && graph_) {
    detail::check(tn_graph_destroy(graph_), "BatchTebd graph");
    graph_ = nullptr;
  }
}

void BatchTebd::step(int n) {
  launch(n);
  wait();
}

void BatchTebd::set_active(int chain, bool on) {
  if (chain < 0 || chain >= chains()) detail::fail("BatchTebd::set_active: bad index");
  if (active_host_[chain] != (on ? 1 : 0)) {
    active_host_[chain] = on ? 1 : 0;
    active_dirty_ = true;
  }
}

void BatchTebd::launch(int n) {
  if (active_dirty_) {
    // in stream order before the next step (graph replays read the flags)
    detail::check(tn_memcpy_h2d(active_, active_host_.data(), sizeof(int) * active_host_.size(), stream_),
                  "BatchTebd active");
    active_dirty_ = false;
  }
  for (int k = 0; k < n; ++k) {
    if (use_graph_ && warm_[0] && warm_[1]) {
      // every later step issues the same launches: capture once, replay
      if (!graph_) {
        detail::check(tn_graph_begin(stream_), "BatchTebd capture");
        run_step();
        detail::check(tn_graph_end(stream_, &graph_), "BatchTebd capture");
      }
      detail::check(tn_graph_launch(graph_, stream_), "BatchTebd replay");
    } else {
      run_step();
    }
  }
}

void BatchTebd::wait() {
  if (!last_svd_) return;
  // Both updates' SVD plans: their counters cover every execution since the
  // last check (all steps of launch(n), replayed or issued directly).
  int bad = 0, most = 0;
  for (int p = 0; p < 2; ++p) {
    if (!svd_[p]) continue;
    int nonconv = 0, sw = 0;
    detail::check(tn_svd_plan_check(svd_[p], &nonconv, &sw, stream_), "BatchTebd svd");
    bad += nonconv;
    most = std::max(most, sw);
  }
  sweeps_ = most;
  if (bad) detail::fail("BatchTebd: " + std::to_string(bad) + " SVD problem(s) did not converge");
}

bool BatchTebd::ready() const {
  int done = 0;
  detail::check(tn_stream_query(stream_, &done), "BatchTebd ready");
  return done != 0;
}

std::vector<double> BatchTebd::lambda(int chain, int which) const {
  if (chain < 0 || chain >= chains() || which < 0 || which > 1) detail::fail("BatchTebd::lambda: bad index");
  std::vector<double> h(chi_[chain]);
  detail::check(tn_memcpy_d2h(h.data(), arena_.get() + off_[chain].l[which], sizeof(double) * h.size(), stream_),
                "BatchTebd lambda");
  return h;
}

double BatchTebd::trunc_err(int chain) const {
  if (chain < 0 || chain >= chains()) detail::fail("BatchTebd::trunc_err: bad index");
  double h[2];
  detail::check(tn_memcpy_d2h(h, arena_.get() + off_[chain].sc[1], sizeof(h), stream_), "BatchTebd trunc_err");
  return h[1];
}

const U1Sectors* BatchTebd::sectors(int chain, int bond) const {
  if (chain < 0 || chain >= chains() || bond < 0 || bond > 1) detail::fail("BatchTebd::sectors: bad index");
  return sym_[chain] ? &u1_[chain].bond[bond] : nullptr;
}

double BatchTebd::energy(int chain) {
  if (chain < 0 || chain >= chains()) detail::fail("BatchTebd::energy: bad index");
  detail::check(tn_stream_sync(stream_), "BatchTebd energy");   // the networks run on the caller's stream
  const int c = chi_[chain], kd = d_[chain];
  const Off& f = off_[chain];
  std::vector<Bond> b3 = {Bond(BD_IN, c), Bond(BD_IN, c), Bond(BD_OUT, kd)};
  UniTensor ALa(b3, "ALa"), BLb(b3, "BLb");
  if (!sym_[chain]) {
    ALa.setElem(arena_.get() + f.x[0], true);
    BLb.setElem(arena_.get() + f.x[1], true);
  } else {
    // the dense (chi, chi, d) layout of the block tensors
    const U1Chain& u = u1_[chain];
    for (int p = 0; p < 2; ++p) {
      std::vector<double> blk((size_t)u.xsize[p]), dense((size_t)c * c * kd, 0.0);
      detail::check(tn_memcpy_d2h(blk.data(), arena_.get() + f.x[p], sizeof(double) * blk.size(), stream_),
                    "BatchTebd energy");
      const U1Sectors& L = u.bond[1 - p];
      const U1Sectors& R = u.bond[p];
      for (size_t a = 0; a < L.q.size(); ++a)
        for (int s = 0; s < kd; ++s) {
          if (u.xpart[p][a][s] < 0) continue;
          const int j = R.find(L.q[a] + u.phys[s]);
          for (int ar = 0; ar < L.n[a]; ++ar)
            for (int mr = 0; mr < R.n[j]; ++mr)
              dense[((size_t)(L.off[a] + ar) * c + R.off[j] + mr) * kd + s] =
                  blk[u.xblk[p][a] + ar * u.xw[p][a] + u.xpart[p][a][s] + mr];
        }
      (p == 0 ? ALa : BLb).setRawElem(dense);
    }
  }
  Matrix La(c, c, arena_.get() + f.l[0], true, true), Lb(c, c, arena_.get() + f.l[1], true, true);
  std::vector<Bond> bH = {Bond(BD_IN, kd), Bond(BD_IN, kd), Bond(BD_OUT, kd), Bond(BD_OUT, kd)};
  UniTensor H(bH, "H");
  H.putBlock(kd == 2 ? xxz_two_site(delta_[chain], 1.0) : xxz_two_site_spin1(delta_[chain], 1.0));
  Network MPS(netdir_ + "/MPS.net"), meas(netdir_ + "/measure.net");
  return itebd_energy(ALa, BLb, La, Lb, H, MPS, meas);
}

}  // namespace uni10
