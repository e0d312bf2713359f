// C facade of the Uni10 C++ API (include/uni10_capi.h).
#include "uni10_capi.h"

#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "batch_tebd.h"
#include "common.h"
#include "tebd.h"

using namespace uni10;

struct u10_tensor_s {
  UniTensor T;
};
struct u10_network_s {
  explicit u10_network_s(const std::string& f) : N(f) {}
  Network N;
};
struct u10_itebd_s {
  template <typename... A>
  explicit u10_itebd_s(A&&... a) : E(std::forward<A>(a)...) {}
  ITebd E;
};
struct u10_batch_tebd_s {
  template <typename... A>
  explicit u10_batch_tebd_s(A&&... a) : B(std::forward<A>(a)...) {}
  BatchTebd B;
};

namespace {
thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
  try {
    f();
    g_err.clear();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  } catch (...) {
    g_err = "unknown error";
    return -1;
  }
}

Qnum qn(const int* t) { return Qnum((parityFType)t[2], t[0], (parityType)t[1]); }

void need(const void* p, const char* what) {
  if (!p) detail::fail(std::string(what) + ": NULL argument");
}
}  // namespace

extern "C" {

const char* u10_last_error(void) { return g_err.c_str(); }

int u10_set_gemm_mode(int mode) {
  return guard([&] {
    if (mode < GEMM_F64 || mode > GEMM_BF16_EXACT) detail::fail("unknown GEMM mode");
    setGemmMode((GemmMode)mode);
  });
}

int u10_set_stream(void* s) {
  return guard([&] { setStream(s); });
}

int u10_set_autotune(int on) {
  return guard([&] { setAutotune(on != 0); });
}

int u10_tensor_create(int nbonds, const int* types, const int* nruns, const int* run_qnums, const int* run_degs,
                      const char* name, u10_tensor* out) {
  return guard([&] {
    need(out, "u10_tensor_create");
    std::vector<Bond> bonds;
    int r = 0;
    for (int b = 0; b < nbonds; ++b) {
      std::vector<Qnum> full;
      for (int k = 0; k < nruns[b]; ++k, ++r) full.insert(full.end(), run_degs[r], qn(run_qnums + 3 * r));
      bonds.emplace_back(types[b] == 1 ? BD_IN : BD_OUT, full);
    }
    *out = new u10_tensor_s{UniTensor(bonds, name ? name : "")};
  });
}

int u10_tensor_create_c(int nbonds, const int* types, const int* nruns, const int* run_qnums, const int* run_degs,
                        const char* name, u10_tensor* out) {
  return guard([&] {
    need(out, "u10_tensor_create_c");
    std::vector<Bond> bonds;
    int r = 0;
    for (int b = 0; b < nbonds; ++b) {
      std::vector<Qnum> full;
      for (int k = 0; k < nruns[b]; ++k, ++r) full.insert(full.end(), run_degs[r], qn(run_qnums + 3 * r));
      bonds.emplace_back(types[b] == 1 ? BD_IN : BD_OUT, full);
    }
    *out = new u10_tensor_s{UniTensor(CTYPE, bonds, name ? name : "")};
  });
}

int u10_tensor_type(u10_tensor t, int* type) {
  return guard([&] {
    need(type, "u10_tensor_type");
    *type = t->T.typeID();
  });
}

int u10_bond_combine(int nbonds, const int* types, const int* nruns, const int* run_qnums, const int* run_degs,
                     int use_free_fn, int tp, int* out_type, int* out_dim, int* out_nruns, int* out_qnums,
                     int* out_degs, int* out_ndeg, int* deg_qnums, int* deg_counts, int cap) {
  return guard([&] {
    need(types, "u10_bond_combine");
    if (nbonds < 1) detail::fail("There should be at least one bond in the input vector to be combined.");
    std::vector<Bond> bonds;
    int r = 0;
    for (int b = 0; b < nbonds; ++b) {
      std::vector<Qnum> full;
      for (int k = 0; k < nruns[b]; ++k, ++r) full.insert(full.end(), run_degs[r], qn(run_qnums + 3 * r));
      bonds.emplace_back(types[b] == 1 ? BD_IN : BD_OUT, full);
    }
    Bond out;
    if (use_free_fn) {
      out = combine(tp == 1 ? BD_IN : BD_OUT, bonds);
    } else {
      out = bonds[0];
      for (int b = 1; b < nbonds; ++b) out.combine(bonds[b]);
    }
    *out_type = (int)out.type();
    *out_dim = out.dim();
    *out_nruns = (int)out.qnums().size();
    for (int i = 0; i < *out_nruns && i < cap; ++i) {
      const Qnum& q = out.qnums()[i];
      out_qnums[3 * i] = q.U1();
      out_qnums[3 * i + 1] = (int)q.prt();
      out_qnums[3 * i + 2] = (int)q.prtF();
      out_degs[i] = out.qdegs()[i];
    }
    std::map<Qnum, int> dg = out.degeneracy();
    *out_ndeg = (int)dg.size();
    int i = 0;
    for (auto& kv : dg) {
      if (i < cap) {
        deg_qnums[3 * i] = kv.first.U1();
        deg_qnums[3 * i + 1] = (int)kv.first.prt();
        deg_qnums[3 * i + 2] = (int)kv.first.prtF();
        deg_counts[i] = kv.second;
      }
      ++i;
    }
  });
}

int u10_tensor_scalar(double v, u10_tensor* out) {
  return guard([&] { *out = new u10_tensor_s{UniTensor(v)}; });
}

int u10_tensor_load(const char* fname, u10_tensor* out) {
  return guard([&] { *out = new u10_tensor_s{UniTensor(std::string(fname))}; });
}

int u10_tensor_save(u10_tensor t, const char* fname) {
  return guard([&] { t->T.save(fname); });
}

int u10_tensor_copy(u10_tensor t, u10_tensor* out) {
  return guard([&] { *out = new u10_tensor_s{UniTensor(t->T)}; });
}

int u10_tensor_free(u10_tensor t) {
  return guard([&] { delete t; });
}

int u10_tensor_info(u10_tensor t, int* nb, int* ib, int64_t* ne, int* nblk) {
  return guard([&] {
    if (nb) *nb = (int)t->T.bondNum();
    if (ib) *ib = (int)t->T.inBondNum();
    if (ne) *ne = (int64_t)t->T.elemNum();
    if (nblk) *nblk = (int)t->T.blockNum();
  });
}

int u10_tensor_labels(u10_tensor t, int* labels) {
  return guard([&] {
    auto l = t->T.label();
    std::copy(l.begin(), l.end(), labels);
  });
}

int u10_tensor_set_label(u10_tensor t, int n, const int* labels) {
  return guard([&] { t->T.setLabel(std::vector<int>(labels, labels + n)); });
}

int u10_tensor_bond(u10_tensor t, int b, int* type, int* dim, int* nruns) {
  return guard([&] {
    Bond bd = t->T.bond((size_t)b);
    if (type) *type = (int)bd.type();
    if (dim) *dim = bd.dim();
    if (nruns) *nruns = (int)bd.qnums().size();
  });
}

int u10_tensor_bond_runs(u10_tensor t, int b, int* q, int* d) {
  return guard([&] {
    Bond bd = t->T.bond((size_t)b);
    for (size_t i = 0; i < bd.qnums().size(); ++i) {
      q[3 * i] = bd.qnums()[i].U1();
      q[3 * i + 1] = (int)bd.qnums()[i].prt();
      q[3 * i + 2] = (int)bd.qnums()[i].prtF();
      d[i] = bd.qdegs()[i];
    }
  });
}

int u10_tensor_block(u10_tensor t, int i, int* q, int64_t* rows, int64_t* cols, int64_t* offset) {
  return guard([&] {
    const auto& bl = t->T.blockInfo();
    if (i < 0 || i >= (int)bl.size()) detail::fail("block index out of range");
    auto it = bl.begin();
    std::advance(it, i);
    q[0] = it->first.U1();
    q[1] = (int)it->first.prt();
    q[2] = (int)it->first.prtF();
    *rows = (int64_t)it->second.Rnum;
    *cols = (int64_t)it->second.Cnum;
    *offset = (int64_t)it->second.offset;
  });
}

int u10_tensor_set_elem(u10_tensor t, const double* host) {
  return guard([&] { t->T.setElem(host, false); });
}

int u10_tensor_get_elem(u10_tensor t, double* host) {
  return guard([&] {
    auto h = t->T.getHostElem();
    std::memcpy(host, h.data(), h.size() * sizeof(double));
  });
}

int u10_tensor_set_raw(u10_tensor t, const double* raw) {
  return guard([&] { t->T.setRawElem(raw); });
}

int u10_tensor_raw_shape(u10_tensor t, int64_t* rows, int64_t* cols) {
  return guard([&] {
    int64_t r = 1, c = 1;
    for (auto& b : t->T.bond()) (b.type() == BD_IN ? r : c) *= b.dim();
    *rows = r;
    *cols = c;
  });
}

int u10_tensor_get_raw(u10_tensor t, double* raw) {
  return guard([&] {
    Matrix m = t->T.getRawElem();
    auto h = m.getHostElem();
    std::memcpy(raw, h.data(), h.size() * sizeof(double));
  });
}

int u10_tensor_at(u10_tensor t, const int64_t* idxs, double* out) {
  return guard([&] {
    std::vector<size_t> ix(idxs, idxs + t->T.bondNum());
    *out = t->T.at(ix);
  });
}

int u10_tensor_put_block(u10_tensor t, const int* q, int64_t rows, int64_t cols, const double* host) {
  return guard([&] {
    if (t->T.typeID() == 2)
      t->T.putBlock(qn(q), Matrix(CTYPE, (size_t)rows, (size_t)cols,
                                  reinterpret_cast<const std::complex<double>*>(host), false));
    else
      t->T.putBlock(qn(q), Matrix((size_t)rows, (size_t)cols, host, false, false));
  });
}

int u10_tensor_get_block(u10_tensor t, const int* q, double* host) {
  return guard([&] {
    auto h = t->T.getBlock(qn(q)).getHostElem();
    std::memcpy(host, h.data(), h.size() * sizeof(double));
  });
}

int u10_tensor_permute(u10_tensor t, int n, const int* labels, int in_bond_num) {
  return guard([&] { t->T.permute(std::vector<int>(labels, labels + n), in_bond_num); });
}

int u10_tensor_transpose(u10_tensor t) {
  return guard([&] { t->T.transpose(); });
}

int u10_tensor_randomize(u10_tensor t, uint64_t seed) {
  return guard([&] { t->T.randomize(seed); });
}

int u10_tensor_scale(u10_tensor t, double a) {
  return guard([&] { t->T *= a; });
}

int u10_tensor_norm(u10_tensor t, double* out) {
  return guard([&] { *out = t->T.norm(); });
}

int u10_tensor_device_ptr(u10_tensor t, void** dptr) {
  return guard([&] { *dptr = t->T.getElem(); });
}

int u10_contract(u10_tensor a, u10_tensor b, int fast, u10_tensor* out) {
  return guard([&] { *out = new u10_tensor_s{contract(a->T, b->T, fast != 0)}; });
}

int u10_otimes(u10_tensor a, u10_tensor b, u10_tensor* out) {
  return guard([&] { *out = new u10_tensor_s{otimes(a->T, b->T)}; });
}

int u10_tensor_mul(u10_tensor a, u10_tensor b, u10_tensor* out) {
  return guard([&] { *out = new u10_tensor_s{a->T * b->T}; });
}

int u10_tensor_ex_swap(u10_tensor a, u10_tensor b, int* pairs, int cap, int* n) {
  return guard([&] {
    std::vector<_Swap> sw = a->T.exSwap(b->T);
    *n = (int)sw.size();
    for (int i = 0; i < (int)sw.size() && i < cap; ++i) {
      pairs[2 * i] = sw[i].b1;
      pairs[2 * i + 1] = sw[i].b2;
    }
  });
}

int u10_tensor_add_gate(u10_tensor t, int n, const int* pairs) {
  return guard([&] {
    if (n > 0) need(pairs, "u10_tensor_add_gate");
    std::vector<_Swap> sw(n);
    for (int i = 0; i < n; ++i) sw[i] = _Swap{pairs[2 * i], pairs[2 * i + 1]};
    t->T.addGate(sw);
  });
}

int u10_network_create(const char* fname, u10_network* out) {
  return guard([&] { *out = new u10_network_s(fname); });
}

int u10_network_put(u10_network net, const char* name, u10_tensor t, int transposed) {
  return guard([&] {
    if (transposed)
      net->N.putTensorT(name, &t->T);
    else
      net->N.putTensor(name, &t->T);
  });
}

int u10_network_launch(u10_network net, const char* name, u10_tensor* out) {
  return guard([&] { *out = new u10_tensor_s{net->N.launch(name ? name : "")}; });
}

int u10_network_tree(u10_network net, char* buf, int len) {
  return guard([&] {
    std::string s = net->N.tree();
    std::snprintf(buf, (size_t)len, "%s", s.c_str());
  });
}

int u10_network_set_graph(u10_network net, int on) {
  return guard([&] { net->N.setGraph(on != 0); });
}

int u10_network_print(u10_network net, char* buf, int len) {
  return guard([&] {
    std::ostringstream os;
    os << net->N;
    std::snprintf(buf, (size_t)len, "%s", os.str().c_str());
  });
}

int u10_network_free(u10_network net) {
  return guard([&] { delete net; });
}

int u10_rdotr(int64_t ar, int64_t ac, int adiag, const double* a, int64_t br, int64_t bc, int bdiag,
              const double* b, int64_t* cr, int64_t* cc, int* cdiag, double* c, int64_t cap) {
  return guard([&] {
    Matrix A((size_t)ar, (size_t)ac, a, adiag != 0, false), B((size_t)br, (size_t)bc, b, bdiag != 0, false);
    Matrix C = RDotR(A, B);
    *cr = (int64_t)C.row();
    *cc = (int64_t)C.col();
    *cdiag = C.isDiag() ? 1 : 0;
    std::vector<double> h = C.getHostElem();
    if ((int64_t)h.size() > cap) detail::fail("u10_rdotr: output larger than the given buffer");
    std::copy(h.begin(), h.end(), c);
  });
}

int u10_svd(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT) {
  return guard([&] {
    Matrix a((size_t)M, (size_t)N, A, false, false);
    auto r = a.svd();
    auto u = r[0].getHostElem(), s = r[1].getHostElem(), vt = r[2].getHostElem();
    std::memcpy(U, u.data(), u.size() * 8);
    std::memcpy(S, s.data(), s.size() * 8);
    std::memcpy(VT, vt.data(), vt.size() * 8);
  });
}

int u10_svd_c128(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT) {
  return guard([&] {
    Matrix a(CTYPE, (size_t)M, (size_t)N, reinterpret_cast<const std::complex<double>*>(A), false);
    auto r = a.svd();
    auto u = r[0].getHostElem(), s = r[1].getHostElem(), vt = r[2].getHostElem();
    std::memcpy(U, u.data(), u.size() * 8);
    std::memcpy(S, s.data(), s.size() * 8);
    std::memcpy(VT, vt.data(), vt.size() * 8);
  });
}

int u10_eigh_c128(const double* A, int64_t N, double* eig, double* vec) {
  return guard([&] {
    Matrix a(CTYPE, (size_t)N, (size_t)N, reinterpret_cast<const std::complex<double>*>(A), false);
    auto r = a.eigh();
    auto e = r[0].getHostElem(), v = r[1].getHostElem();
    std::memcpy(eig, e.data(), e.size() * 8);
    std::memcpy(vec, v.data(), v.size() * 8);
  });
}

int u10_eigh(const double* A, int64_t N, double* eig, double* vec) {
  return guard([&] {
    Matrix a((size_t)N, (size_t)N, A, false, false);
    auto r = a.eigh();
    auto e = r[0].getHostElem(), v = r[1].getHostElem();
    std::memcpy(eig, e.data(), e.size() * 8);
    std::memcpy(vec, v.data(), v.size() * 8);
  });
}

int u10_lanczos_eigh(const double* A, int64_t N, double* psi, uint64_t max_iter, double err_tol, double* E0,
                     uint64_t* iters) {
  return guard([&] {
    if (!A || !psi || !E0 || !iters) detail::fail("u10_lanczos_eigh: NULL argument");
    Matrix a((size_t)N, (size_t)N, A, false, false);
    Matrix v((size_t)N, 1, psi, false, false);
    double e0 = 0.0;
    const size_t it = lanczosEigh(a, e0, v, (size_t)max_iter, err_tol);
    auto h = v.getHostElem();
    std::memcpy(psi, h.data(), h.size() * 8);
    *E0 = e0;
    *iters = it;
  });
}

int u10_exph(double a, const double* H, int64_t N, double* out) {
  return guard([&] {
    Matrix h((size_t)N, (size_t)N, H, false, false);
    auto g = exph(a, h).getHostElem();
    std::memcpy(out, g.data(), g.size() * 8);
  });
}

int u10_itebd_create(int chi, int d, double delta, double tau, uint64_t seed, const char* netdir, u10_itebd* out) {
  return guard([&] { *out = new u10_itebd_s(chi, d, delta, tau, seed, std::string(netdir ? netdir : "")); });
}

int u10_itebd_set_graph(u10_itebd t, int on) {
  return guard([&] { t->E.set_graph(on != 0); });
}

int u10_itebd_step(u10_itebd t, int n) {
  return guard([&] {
    for (int i = 0; i < n; ++i) t->E.step();
    t->E.check();
  });
}

int u10_itebd_energy(u10_itebd t, double* out) {
  return guard([&] { *out = t->E.energy(); });
}

int u10_itebd_lambda(u10_itebd t, int which, double* out, int* n) {
  return guard([&] {
    const Matrix& L = which == 0 ? t->E.La() : t->E.Lb();
    auto h = L.getHostElem();
    *n = (int)h.size();
    if (out) std::memcpy(out, h.data(), h.size() * 8);
  });
}

int u10_itebd_trunc_err(u10_itebd t, double* out) {
  return guard([&] { *out = t->E.truncation_error(); });
}

int u10_itebd_free(u10_itebd t) {
  return guard([&] { delete t; });
}

int u10_batch_tebd_create(int n, const int* chi, const double* delta, const uint64_t* seed, double tau,
                          int gemm_mode, const char* netdir, u10_batch_tebd* out) {
  return guard([&] {
    if (n < 1 || !chi || !delta || !seed || !out) detail::fail("u10_batch_tebd_create: bad argument");
    *out = new u10_batch_tebd_s(std::vector<int>(chi, chi + n), std::vector<double>(delta, delta + n),
                                std::vector<uint64_t>(seed, seed + n), tau, gemm_mode, std::string(netdir ? netdir : ""));
  });
}

int u10_batch_tebd_create_ex(int n, const int* chi, const double* delta, const uint64_t* seed, const int* d,
                             const int* sym, double tau, int gemm_mode, const char* netdir, u10_batch_tebd* out) {
  return guard([&] {
    if (n < 1 || !chi || !delta || !seed || !out) detail::fail("u10_batch_tebd_create_ex: bad argument");
    *out = new u10_batch_tebd_s(std::vector<int>(chi, chi + n), std::vector<double>(delta, delta + n),
                                std::vector<uint64_t>(seed, seed + n), tau, gemm_mode, std::string(netdir ? netdir : ""),
                                d ? std::vector<int>(d, d + n) : std::vector<int>(),
                                sym ? std::vector<int>(sym, sym + n) : std::vector<int>());
  });
}

int u10_batch_tebd_sectors(u10_batch_tebd b, int chain, int bond, int* charges, int* dims, int* count) {
  return guard([&] {
    if (!count) detail::fail("u10_batch_tebd_sectors: NULL count");
    const U1Sectors* s = b->B.sectors(chain, bond);
    const int n = s ? (int)s->q.size() : 0;
    if ((charges || dims) && *count < n) detail::fail("u10_batch_tebd_sectors: output arrays too short");
    for (int i = 0; i < n; ++i) {
      if (charges) charges[i] = s->q[i];
      if (dims) dims[i] = s->n[i];
    }
    *count = n;
  });
}

int u10_u1_bond_sectors(int chi, int d, int bond, int* charges, int* dims, int* count) {
  return guard([&] {
    if (!count || chi < 1 || (d != 2 && d != 3) || bond < 0 || bond > 1)
      detail::fail("u10_u1_bond_sectors: bad argument");
    U1Sectors b[2];
    u1_bond_sectors(chi, d, 12, b[0], b[1]);
    const U1Sectors& s = b[bond];
    const int n = (int)s.q.size();
    if ((charges || dims) && *count < n) detail::fail("u10_u1_bond_sectors: output arrays too short");
    for (int i = 0; i < n; ++i) {
      if (charges) charges[i] = s.q[i];
      if (dims) dims[i] = s.n[i];
    }
    *count = n;
  });
}

int u10_batch_tebd_step(u10_batch_tebd b, int n) {
  return guard([&] { b->B.step(n); });
}

int u10_batch_tebd_launch(u10_batch_tebd b, int n) {
  return guard([&] { b->B.launch(n); });
}

int u10_batch_tebd_wait(u10_batch_tebd b) {
  return guard([&] { b->B.wait(); });
}

int u10_batch_tebd_ready(u10_batch_tebd b, int* ready) {
  return guard([&] {
    if (!ready) detail::fail("u10_batch_tebd_ready: NULL out pointer");
    *ready = b->B.ready() ? 1 : 0;
  });
}

int u10_batch_tebd_set_active(u10_batch_tebd b, int chain, int on) {
  return guard([&] { b->B.set_active(chain, on != 0); });
}

int u10_batch_tebd_set_stream(u10_batch_tebd b, void* stream) {
  return guard([&] { b->B.set_stream(stream); });
}

int u10_batch_tebd_set_graph(u10_batch_tebd b, int on) {
  return guard([&] { b->B.set_graph(on != 0); });
}

int u10_batch_tebd_load(u10_batch_tebd b, int chain, double delta, uint64_t seed) {
  return guard([&] { b->B.load(chain, delta, seed); });
}

int u10_batch_tebd_energy(u10_batch_tebd b, int chain, double* out) {
  return guard([&] { *out = b->B.energy(chain); });
}

int u10_batch_tebd_lambda(u10_batch_tebd b, int chain, int which, double* out, int* n) {
  return guard([&] {
    auto h = b->B.lambda(chain, which);
    *n = (int)h.size();
    if (out) std::memcpy(out, h.data(), h.size() * 8);
  });
}

int u10_batch_tebd_trunc_err(u10_batch_tebd b, int chain, double* out) {
  return guard([&] { *out = b->B.trunc_err(chain); });
}

int u10_batch_tebd_info(u10_batch_tebd b, double* flops, int* sweeps) {
  return guard([&] {
    if (flops) *flops = b->B.gemm_flops_per_step();
    if (sweeps) *sweeps = b->B.last_sweeps();
  });
}

int u10_batch_tebd_free(u10_batch_tebd b) {
  return guard([&] { delete b; });
}

}  // extern "C"
