// Batched iTEBD over the kernel C ABI.
//
// Per chain (bond dimension c, d = 2, D = c d) and per update, with X the
// Gamma*lambda tensor being rebuilt (ALa in update(A,B)), Y the other one,
// LX / LY the lambdas between and outside them (iTEBD.cpp:148-173 with
// itebd.net, updateA.net):
//   T1[a,s,b,t]   = sum_m X[a,m,s] Y[m,b,t]                     GEMM  D x c x D
//   C[a,s',b,t']  = sum_{s,t} T1[a,s,b,t] G[(s,t),(s',t')]      GEMM  c^2 x d^2 x d^2
//   Theta         = LY[a] C[(a,s'),(b,t')]                      row scaling
//   Theta         = U S VT                                      SVD (D x D)
//   LX            = S[:c] / ||S[:c]||,  Y = VT[:c, :]           truncation
//   X[a,m',s']    = sum_{b,t'} C[a,s',b,t'] VT[m',(b,t')] / ||S[:c]||
// Every stage is one launch over all chains: the GEMMs are grouped plans
// whose index maps read each operand in its stored layout, the SVD is a
// batched plan warm-started from the bond's previous VT, and the vector
// steps are vector-op plans.  All chains live in one arena.
#include "batch_tebd.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"
#include "tebd.h"

namespace uni10 {

namespace {
constexpr int kD = 2;

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
}  // namespace

BatchTebd::BatchTebd(const std::vector<int>& chi, const std::vector<double>& delta,
                     const std::vector<uint64_t>& seed, double tau, int gemm_mode, const std::string& netdir)
    : chi_(chi), delta_(delta), netdir_(netdir) {
  const size_t nc = chi.size();
  if (delta.size() != nc || seed.size() != nc) detail::fail("BatchTebd: chi, delta and seed differ in length");
  if (nc == 0) detail::fail("BatchTebd: no chains");
  for (int c : chi)
    if (c < 1 || c > 256) detail::fail("BatchTebd: chi must be in [1, 256]");
  stream_ = detail::stream();
  // arena layout
  int64_t o = 0;
  auto take = [&](int64_t n) {
    const int64_t r = o;
    o += (n + 1) & ~int64_t(1);   // keep every region 16-byte aligned
    return r;
  };
  off_.resize(nc);
  for (size_t i = 0; i < nc; ++i) {
    const int64_t c = chi[i], D = c * kD, t = c * c * kD * kD;
    Off& f = off_[i];
    f.x[0] = take(c * c * kD);
    f.x[1] = take(c * c * kD);
    f.l[0] = take(c);
    f.l[1] = take(c);
    f.g = take(kD * kD * kD * kD);
    f.t1 = take(t);
    f.cb = take(t);
    f.th = take(t);
    f.u = take(D * D);
    f.s = take(D);
    f.vt[0] = take(D * D);
    f.vt[1] = take(D * D);
    f.sc[0] = take(2);
    f.sc[1] = take(2);
  }
  arena_.reset((size_t)o, true);

  // initial state of every chain (as ITebd) and its gate
  tau_ = tau;
  std::vector<double> host((size_t)o, 0.0);
  for (size_t i = 0; i < nc; ++i) {
    const int64_t end = (i + 1 < nc) ? off_[i + 1].x[0] : o;
    region_.push_back(end - off_[i].x[0]);
    init_chain((int)i, delta[i], seed[i], host.data() + off_[i].x[0]);
  }
  detail::check(tn_memcpy_h2d(arena_.get(), host.data(), sizeof(double) * (size_t)o, stream_), "BatchTebd upload");

  // plans, one set per update parity p (p = 0: X = ALa, Y = BLb; p = 1: swapped)
  for (int p = 0; p < 2; ++p) {
    std::vector<tn_gemm_desc_ex> d1, d2, d3;
    std::vector<int64_t> sm, sn, sa, su, ss, svt, sv0;
    std::vector<tn_vecop> rows, trunc, scale;
    for (size_t i = 0; i < nc; ++i) {
      const int64_t c = chi[i], D = c * kD, dd = kD * kD;
      const Off& f = off_[i];
      const int64_t X = f.x[p], Y = f.x[1 - p], LX = f.l[p], LY = f.l[1 - p];
      // T1[(a,s),(b,t)] = sum_m X[a,m,s] Y[m,(b,t)]
      tn_gemm_desc_ex g1 = desc(D, D, c, X, Y, f.t1);
      g1.a_row = imap({{c, c * kD}, {kD, 1}});
      g1.a_col = imap({{c, kD}});
      d1.push_back(g1);
      // C[a,s',b,t'] = sum_{s,t} T1[a,s,b,t] G[(s,t),(s',t')]; rows (a,b)
      tn_gemm_desc_ex g2 = desc(c * c, dd, dd, f.t1, f.g, f.cb);
      g2.a_row = imap({{c, kD * c * kD}, {c, kD}});
      g2.a_col = imap({{kD, c * kD}, {kD, 1}});
      g2.c_row = imap({{c, kD * c * kD}, {c, kD}});
      g2.c_col = imap({{kD, c * kD}, {kD, 1}});
      d2.push_back(g2);
      // Theta = LY[a] C[(a,s'),(b,t')]
      rows.push_back(tn_vecop{TN_VECOP_ROWSCALE, 0, D * D, f.cb, f.th, LY, c * kD * kD});
      // SVD of Theta (D x D) into U, S, VT[p]
      sm.push_back(D);
      sn.push_back(D);
      sa.push_back(f.th);
      su.push_back(f.u);
      ss.push_back(f.s);
      svt.push_back(f.vt[p]);
      sv0.push_back(f.vt[p]);
      // LX = S[:c]/||S[:c]||, sc = {norm, discarded}; Y = VT[:c, :]
      trunc.push_back(tn_vecop{TN_VECOP_TRUNCATE, 0, D, f.s, LX, f.sc[p], c});
      trunc.push_back(tn_vecop{TN_VECOP_COPY, 0, c * D, f.vt[p], Y, 0, 0});
      // X[a,m',s'] = sum_{(b,t')} C[(a,s'),(b,t')] VT[m',(b,t')]
      tn_gemm_desc_ex g3 = desc(D, c, D, f.cb, f.vt[p], X);
      g3.b_row = imap({{D, 1}});
      g3.b_col = imap({{c, D}});
      g3.c_row = imap({{c, c * kD}, {kD, 1}});
      g3.c_col = imap({{c, kD}});
      d3.push_back(g3);
      // X /= ||S[:c]||
      scale.push_back(tn_vecop{TN_VECOP_SCALE_INV, 0, c * c * kD, X, X, f.sc[p], 0});
      if (p == 0)
        flops_ += 2.0 * (2.0 * D * D * c + 2.0 * c * c * dd * dd + 2.0 * D * c * D);
    }
    const int64_t n = (int64_t)nc;
    detail::check(tn_gemm_plan_create_ex(&g1_[p], gemm_mode, n, d1.data()), "BatchTebd gemm1");
    detail::check(tn_gemm_plan_create_ex(&g2_[p], gemm_mode, n, d2.data()), "BatchTebd gemm2");
    detail::check(tn_gemm_plan_create_ex(&g3_[p], gemm_mode, n, d3.data()), "BatchTebd gemm3");
    detail::check(tn_svd_plan_create(&svd_cold_[p], n, sm.data(), sn.data(), sa.data(), nullptr, su.data(),
                                     ss.data(), svt.data()),
                  "BatchTebd svd");
    detail::check(tn_svd_plan_create(&svd_warm_[p], n, sm.data(), sn.data(), sa.data(), sv0.data(), su.data(),
                                     ss.data(), svt.data()),
                  "BatchTebd svd");
    detail::check(tn_vecops_plan_create(&rows_[p], (int64_t)rows.size(), rows.data()), "BatchTebd rows");
    detail::check(tn_vecops_plan_create(&trunc_[p], (int64_t)trunc.size(), trunc.data()), "BatchTebd trunc");
    detail::check(tn_vecops_plan_create(&scale_[p], (int64_t)scale.size(), scale.data()), "BatchTebd scale");
  }
}

// Host image of chain i's arena region (base = its x[0]): Gamma lambda
// tensors and lambdas from the seeded start of ITebd (splitmix seeds
// seed..seed+3), the gate exp(-tau H(delta)), and identity warm-start bases
// (so a chain's SVD sequence never depends on an earlier occupant).
void BatchTebd::init_chain(int i, double delta, uint64_t seed, double* base) {
  const int c = chi_[i];
  const Off& f = off_[i];
  auto at = [&](int64_t off) { return base + (off - f.x[0]); };
  std::fill(base, base + region_[i], 0.0);
  const size_t n = (size_t)c * kD * c;
  std::vector<double> ga(n), gb(n), la(c), lb(c);
  for (size_t e = 0; e < n; ++e) {
    ga[e] = detail::splitmix_uniform(seed, e + 1);
    gb[e] = detail::splitmix_uniform(seed + 1, e + 1);
  }
  for (int e = 0; e < c; ++e) {
    la[e] = std::fabs(detail::splitmix_uniform(seed + 2, e + 1));
    lb[e] = std::fabs(detail::splitmix_uniform(seed + 3, e + 1));
  }
  std::sort(la.rbegin(), la.rend());
  std::sort(lb.rbegin(), lb.rend());
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
  double* x0 = at(f.x[0]);
  double* x1 = at(f.x[1]);
  for (int a = 0; a < c; ++a)
    for (int s = 0; s < kD; ++s)
      for (int b = 0; b < c; ++b) {
        x0[((size_t)a * c + b) * kD + s] = ga[((size_t)a * kD + s) * c + b] * l0[b];
        x1[((size_t)a * c + b) * kD + s] = gb[((size_t)a * kD + s) * c + b] * l1[b];
      }
  const std::vector<double> g = takeExp(-tau_, xxz_two_site(delta, 1.0)).getHostElem();
  std::copy(g.begin(), g.end(), at(f.g));
  const int64_t D = (int64_t)c * kD;
  for (int p = 0; p < 2; ++p)
    for (int64_t e = 0; e < D; ++e) at(f.vt[p])[e * D + e] = 1.0;
  delta_[i] = delta;
}

void BatchTebd::load(int chain, double delta, uint64_t seed) {
  if (chain < 0 || chain >= chains()) detail::fail("BatchTebd::load: bad index");
  std::vector<double> host((size_t)region_[chain]);
  init_chain(chain, delta, seed, host.data());
  detail::check(tn_memcpy_h2d(arena_.get() + off_[chain].x[0], host.data(), sizeof(double) * host.size(), stream_),
                "BatchTebd load");
}

BatchTebd::~BatchTebd() {
  if (graph_) tn_graph_destroy(graph_);
  for (int p = 0; p < 2; ++p) {
    tn_gemm_plan_destroy(g1_[p]);
    tn_gemm_plan_destroy(g2_[p]);
    tn_gemm_plan_destroy(g3_[p]);
    tn_svd_plan_destroy(svd_cold_[p]);
    tn_svd_plan_destroy(svd_warm_[p]);
    tn_vecops_plan_destroy(rows_[p]);
    tn_vecops_plan_destroy(trunc_[p]);
    tn_vecops_plan_destroy(scale_[p]);
  }
}

void BatchTebd::update(int p, bool warm) {
  double* A = arena_.get();
  detail::check(tn_gemm_exec(g1_[p], A, A, A, stream_), "BatchTebd gemm1");
  detail::check(tn_gemm_exec(g2_[p], A, A, A, stream_), "BatchTebd gemm2");
  detail::check(tn_vecops_plan_exec(rows_[p], A, stream_), "BatchTebd rows");
  last_svd_ = warm ? svd_warm_[p] : svd_cold_[p];
  detail::check(tn_svd_plan_exec(last_svd_, A, A, A, A, A, stream_), "BatchTebd svd");
  detail::check(tn_vecops_plan_exec(trunc_[p], A, stream_), "BatchTebd trunc");
  detail::check(tn_gemm_exec(g3_[p], A, A, A, stream_), "BatchTebd gemm3");
  detail::check(tn_vecops_plan_exec(scale_[p], A, stream_), "BatchTebd scale");
}

void BatchTebd::run_step() {
  update(0, warm_[0]);
  update(1, warm_[1]);
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
  int nonconv = 0;
  // status of the last SVD launch of the step (the second update's)
  detail::check(tn_svd_plan_check(last_svd_, &nonconv, &sweeps_, stream_), "BatchTebd svd");
  if (nonconv) detail::fail("BatchTebd: an SVD did not converge");
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

double BatchTebd::energy(int chain) {
  if (chain < 0 || chain >= chains()) detail::fail("BatchTebd::energy: bad index");
  const int c = chi_[chain];
  const Off& f = off_[chain];
  std::vector<Bond> b3 = {Bond(BD_IN, c), Bond(BD_IN, c), Bond(BD_OUT, kD)};
  UniTensor ALa(b3, "ALa"), BLb(b3, "BLb");
  ALa.setElem(arena_.get() + f.x[0], true);
  BLb.setElem(arena_.get() + f.x[1], true);
  Matrix La(c, c, arena_.get() + f.l[0], true, true), Lb(c, c, arena_.get() + f.l[1], true, true);
  std::vector<Bond> bH = {Bond(BD_IN, kD), Bond(BD_IN, kD), Bond(BD_OUT, kD), Bond(BD_OUT, kD)};
  UniTensor H(bH, "H");
  H.putBlock(xxz_two_site(delta_[chain], 1.0));
  Network MPS(netdir_ + "/MPS.net"), meas(netdir_ + "/measure.net");
  return itebd_energy(ALa, BLb, La, Lb, H, MPS, meas);
}

}  // namespace uni10
