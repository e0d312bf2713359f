// iTEBD driver written against the Uni10 API, as uni10/examples/iTEBD.cpp
// (update :148-173, measure :189-203, expectation :205-216) with the
// .net files shipped beside it; spin-1/2 XXZ two-site gate exp(-tau H).
#include <dlfcn.h>

#include <algorithm>
#include <cmath>

#include "common.h"
#include "tebd.h"

namespace uni10 {

std::string detail::netdir_or_default(const std::string& netdir) {
  if (!netdir.empty()) return netdir;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&detail::netdir_or_default), &info) == 0 || !info.dli_fname)
    detail::fail("cannot locate libuni10_amd.so to find its network files; pass a directory");
  std::string lib(info.dli_fname);
  const size_t slash = lib.rfind('/');
  return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/nets";
}

std::vector<double> xxz_two_site_host(double delta, double J) {
  // J (Sx Sx + Sy Sy + delta Sz Sz), basis |uu>,|ud>,|du>,|dd>
  std::vector<double> h(16, 0.0);
  h[0] = h[15] = 0.25 * delta * J;
  h[5] = h[10] = -0.25 * delta * J;
  h[6] = h[9] = 0.5 * J;
  return h;
}

Matrix xxz_two_site(double delta, double J) {
  const std::vector<double> h = xxz_two_site_host(delta, J);
  return Matrix(4, 4, h.data(), false, false);
}

Matrix xxz_two_site_spin1(double delta, double J) {
  const std::vector<double> h = xxz_two_site_spin1_host(delta, J);
  return Matrix(9, 9, h.data(), false, false);
}

std::vector<double> xxz_two_site_spin1_host(double delta, double J) {
  // S+ = sqrt(2) (|+1><0| + |0><-1|); (S+S- + S-S+)/2 moves one unit of Sz
  // between the sites with amplitude 1 (sqrt(2) sqrt(2) / 2)
  std::vector<double> h(81, 0.0);
  auto idx = [](int s, int t) { return 3 * s + t; };   // state index 0: +1, 1: 0, 2: -1
  const double sz[3] = {1.0, 0.0, -1.0};
  for (int s = 0; s < 3; ++s)
    for (int t = 0; t < 3; ++t) h[idx(s, t) * 9 + idx(s, t)] = J * delta * sz[s] * sz[t];
  for (int s = 1; s < 3; ++s)      // site 1 raised (s -> s - 1), site 2 lowered (t -> t + 1)
    for (int t = 0; t < 2; ++t) {
      const int a = idx(s, t), b = idx(s - 1, t + 1);
      h[b * 9 + a] = J;
      h[a * 9 + b] = J;
    }
  return h;
}

ITebd::ITebd(int chi, int d, double delta, double tau, uint64_t seed, const std::string& netdir)
    : chi_(chi), d_(d), tau_(tau),
      iTEBD_(detail::netdir_or_default(netdir) + "/itebd.net"),
      updateA_(detail::netdir_or_default(netdir) + "/updateA.net"),
      MPS_(detail::netdir_or_default(netdir) + "/MPS.net"), meas_(detail::netdir_or_default(netdir) + "/measure.net") {
  if (d != 2) detail::fail("ITebd: only the spin-1/2 chain (d = 2) is provided");
  Matrix Hm = xxz_two_site(delta, 1.0);
  std::vector<Bond> bH = {Bond(BD_IN, d), Bond(BD_IN, d), Bond(BD_OUT, d), Bond(BD_OUT, d)};
  H_ = UniTensor(bH, "H");
  H_.putBlock(Hm);
  U_ = UniTensor(bH, "U");
  U_.putBlock(takeExp(-tau, Hm));
  // Gamma_A, Gamma_B (chi, d, chi) ~ U[-1,1) with seeds seed, seed+1; lambdas
  // sorted |U[-1,1)| (seeds seed+2, seed+3) normalised; ALa = Gamma_A lambda_A
  // in the (chi_left, chi_right | d) layout the update networks use.
  const size_t n = (size_t)chi * d * chi;
  std::vector<double> ga(n), gb(n), la(chi), lb(chi);
  for (size_t i = 0; i < n; ++i) {
    ga[i] = detail::splitmix_uniform(seed, i + 1);
    gb[i] = detail::splitmix_uniform(seed + 1, i + 1);
  }
  for (int i = 0; i < chi; ++i) {
    la[i] = std::fabs(detail::splitmix_uniform(seed + 2, i + 1));
    lb[i] = std::fabs(detail::splitmix_uniform(seed + 3, i + 1));
  }
  std::sort(la.rbegin(), la.rend());
  std::sort(lb.rbegin(), lb.rend());
  double na = 0, nbb = 0;
  for (int i = 0; i < chi; ++i) {
    na += la[i] * la[i];
    nbb += lb[i] * lb[i];
  }
  na = std::sqrt(na);
  nbb = std::sqrt(nbb);
  for (int i = 0; i < chi; ++i) {
    la[i] /= na;
    lb[i] /= nbb;
  }
  std::vector<double> ala(n), blb(n);
  for (int a = 0; a < chi; ++a)
    for (int s = 0; s < d; ++s)
      for (int b = 0; b < chi; ++b) {
        ala[((size_t)a * chi + b) * d + s] = ga[((size_t)a * d + s) * chi + b] * la[b];
        blb[((size_t)a * chi + b) * d + s] = gb[((size_t)a * d + s) * chi + b] * lb[b];
      }
  std::vector<Bond> b3 = {Bond(BD_IN, chi), Bond(BD_IN, chi), Bond(BD_OUT, d)};
  ALa_ = UniTensor(b3, "ALa");
  BLb_ = UniTensor(b3, "BLb");
  ALa_.setElem(ala);
  BLb_.setElem(blb);
  La_ = Matrix(chi, chi, la.data(), true, false);
  Lb_ = Matrix(chi, chi, lb.data(), true, false);
}

// SVD of theta started from the right singular vectors of the previous
// theta on the same bond (the two updates of a step alternate between two
// bonds): successive thetas differ by one small imaginary-time step, so the
// Jacobi sweeps start close to convergence.  The factors of Matrix::svd that
// the update uses: S (k x k diagonal) and VT (k x C); U is not formed.  The
// warm-started SVD of each bond runs through a plan built once for its shape:
// no allocation, copy or wait per update; its convergence is checked by
// check().
std::vector<Matrix> ITebd::svd_warm(const Matrix& T) {
  Matrix& vt_prev = vt_prev_[parity_];
  const size_t R = T.row(), C = T.col(), k = std::min(R, C);
  // S and VT are written in full by the SVD: no zero fill
  std::vector<Matrix> outs;
  outs.push_back(Matrix(Matrix::uninit_t{}, k, k, true));
  outs.push_back(Matrix(Matrix::uninit_t{}, k, C, false));
  if (k == 0) return outs;
  if (R >= C && vt_prev.row() == C && vt_prev.col() == C) {
    tn_svd_plan& P = plan_[parity_];
    if (!P || plan_shape_[parity_][0] != (int64_t)R || plan_shape_[parity_][1] != (int64_t)C) {
      if (P) {
        check();
        tn_svd_plan_destroy(P);
        P = nullptr;
      }
      const int64_t m = (int64_t)R, n = (int64_t)C, z = 0, none = -1;
      detail::check(tn_svd_plan_create(&P, 1, &m, &n, &z, &z, &none, &z, &z), "ITebd svd plan");
      plan_shape_[parity_][0] = m;
      plan_shape_[parity_][1] = n;
    }
    detail::check(tn_svd_plan_exec(P, T.getElem(), vt_prev.getElem(), nullptr, outs[0].getElem(),
                                   outs[1].getElem(), detail::stream()),
                  "ITebd svd");
  } else {
    detail::check(tn_svd_f64(T.getElem(), (int64_t)R, (int64_t)C, nullptr, outs[0].getElem(),
                             outs[1].getElem(), detail::stream()),
                  "ITebd svd");
  }
  // the caller keeps outs[1] as this bond's next start (update())
  return outs;
}

ITebd::~ITebd() {
  for (tn_svd_plan& P : plan_) tn_svd_plan_destroy(P);
}

void ITebd::check() {
  int bad = 0;
  for (tn_svd_plan P : plan_) {
    if (!P) continue;
    int nonconv = 0, sw = 0;
    detail::check(tn_svd_plan_check(P, &nonconv, &sw, detail::stream()), "ITebd svd");
    bad += nonconv;
  }
  if (bad) detail::fail("ITebd: " + std::to_string(bad) + " SVD(s) did not converge");
}

void ITebd::update(UniTensor& ALa, UniTensor& BLb, Matrix& La, Matrix& Lb) {
  Qnum q0(0);
  iTEBD_.putTensor("ALa", &ALa);
  iTEBD_.putTensor("BLb", &BLb);
  iTEBD_.putTensor("expH", &U_);
  UniTensor C = iTEBD_.launch();
  UniTensor Theta(C.bond(), "Theta");
  Theta.putBlock(q0, Lb * C.getBlock(q0));
  Theta.permute(2);
  std::vector<Matrix> rets = svd_warm(Theta.getBlock(q0));
  parity_ ^= 1;
  const int k = (int)rets[0].row();
  const int dim = chi_ < k ? chi_ : k;
  // lambda = S[:dim] / ||S[:dim]|| (iTEBD.cpp:157-161), the norm and the
  // discarded weight stay on the device: the update never waits for the host
  Matrix lam(Matrix::uninit_t{}, dim, dim, true);  // every entry written below
  detail::check(tn_truncate_f64(rets[0].getElem(), k, chi_, lam.getElem(), scal_.get(), scal_.get() + 1,
                                detail::stream()),
                "ITebd truncate");
  La = std::move(lam);
  Bond bdi(BD_IN, dim);
  std::vector<Bond> bond3 = ALa.bond();
  bond3[0] = bdi;
  BLb.assign(bond3);
  // one block (no symmetry), its elements the first dim rows of VT
  BLb.setElem(rets[1].getElem(), true);
  // VT starts this bond's next SVD (svd_warm flipped parity_)
  vt_prev_[parity_ ^ 1] = std::move(rets[1]);
  updateA_.putTensor("BLb", &BLb);
  updateA_.putTensor("C", &C);
  ALa = updateA_.launch();
  detail::check(tn_scale_inv_dev_f64(ALa.getElem(), (int64_t)ALa.elemNum(), scal_.get(), detail::stream()),
                "ITebd scale");
}

double ITebd::truncation_error() const {
  double e = 0;
  detail::check(tn_memcpy_d2h(&e, scal_.get() + 1, sizeof(double), detail::stream()), "ITebd trunc_err");
  return e;
}

void ITebd::step() {
  update(ALa_, BLb_, La_, Lb_);
  update(BLb_, ALa_, Lb_, La_);
}

// <psi|H|psi> / <psi|psi> of the two-site state L R (iTEBD.cpp:205-216).
static double expectation(UniTensor& L, UniTensor& R, UniTensor& H, Network& MPS, Network& meas) {
  Qnum q0(0);
  MPS.putTensor("L", &L);
  MPS.putTensor("R", &R);
  UniTensor psi = MPS.launch();
  double norm = psi.getBlock(q0).norm();
  norm *= norm;
  meas.putTensor("bra", &psi);
  meas.putTensorT("ket", &psi);
  meas.putTensor("Op", &H);
  return meas.launch()[0] / norm;
}

double itebd_energy(UniTensor& ALa, UniTensor& BLb, const Matrix& La, const Matrix& Lb, UniTensor& H,
                    Network& MPS, Network& meas) {
  // iTEBD.cpp:189-203: mean of the two bond energies
  Qnum q0(0);
  UniTensor A1(ALa);
  A1.permute(1);
  A1.putBlock(q0, Lb * A1.getBlock(q0));
  A1.permute(2);
  UniTensor B1(BLb);
  B1.permute(1);
  B1.putBlock(q0, La * B1.getBlock(q0));
  B1.permute(2);
  double val = expectation(A1, BLb, H, MPS, meas);
  val += expectation(B1, ALa, H, MPS, meas);
  return val / 2;
}

double ITebd::energy() {
  check();
  return itebd_energy(ALa_, BLb_, La_, Lb_, H_, MPS_, meas_);
}

}  // namespace uni10
