// Exercises the backend seam (uni10_seam_amd.h) the way the reference's
// classes call it, on seeded inputs, with operands in device memory and in
// host memory; writes every result as raw doubles to OUTDIR/<name>.bin for
// a checker to compare with a textbook computation on the same inputs.
//   seam_check OUTDIR
// Inputs: SplitMix64 uniform[-1, 1), value_i = u(seed, i), i = 1, 2, ...
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "uni10_seam_amd.h"

using namespace uni10;

static double u(uint64_t seed, uint64_t i) {
  uint64_t z = seed + i * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

static std::vector<double> seeded(uint64_t seed, size_t n) {
  std::vector<double> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = u(seed, i + 1);
  return v;
}

static std::string OUT;

static void put(const std::string& name, const std::vector<double>& v) {
  FILE* f = std::fopen((OUT + "/" + name + ".bin").c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + name);
  std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
}

// a device copy of a host vector (elemAlloc + elemCopy, as the classes do)
static double* to_dev(const std::vector<double>& h) {
  bool ongpu = false;
  double* d = static_cast<double*>(elemAlloc(h.size() * sizeof(double), ongpu));
  if (!ongpu) throw std::runtime_error("elemAlloc did not return device memory");
  elemCopy(d, h.data(), h.size() * sizeof(double), true, false);
  return d;
}

static std::vector<double> to_host(const double* d, size_t n) {
  std::vector<double> h(n);
  elemCopy(h.data(), d, n * sizeof(double), false, true);
  return h;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUTDIR\n", argv[0]);
    return 2;
  }
  OUT = argv[1];
  try {
    const int M = 37, K = 29, N = 23;
    // matrixMul, device and host operands
    std::vector<double> a = seeded(1, M * K), b = seeded(2, K * N), c(M * N);
    double *da = to_dev(a), *db = to_dev(b), *dc = to_dev(c);
    matrixMul(da, db, M, N, K, dc, true, true, true);
    put("mm_dev", to_host(dc, M * N));
    matrixMul(a.data(), b.data(), M, N, K, c.data(), false, false, false);
    put("mm_host", c);
    // matrixSVD of a 40 x 31 matrix: all factors on the device; S only from host memory
    std::vector<double> s = seeded(3, 40 * 31), hs(31);
    double *ds = to_dev(s), *dU = to_dev(std::vector<double>(40 * 31)), *dS = to_dev(std::vector<double>(31)),
           *dVT = to_dev(std::vector<double>(31 * 31));
    matrixSVD(ds, 40, 31, dU, dS, dVT, true);
    put("svd_U", to_host(dU, 40 * 31));
    put("svd_S", to_host(dS, 31));
    put("svd_VT", to_host(dVT, 31 * 31));
    matrixSVD(s.data(), 40, 31, nullptr, hs.data(), nullptr, false);
    put("svd_S_host", hs);
    // eigSyDecompose of H = X + X^T (12 x 12)
    std::vector<double> x = seeded(4, 144), h(144), w(12), v(144);
    for (int i = 0; i < 12; ++i)
      for (int j = 0; j < 12; ++j) h[i * 12 + j] = x[i * 12 + j] + x[j * 12 + i];
    eigSyDecompose(h.data(), 12, w.data(), v.data(), false);
    put("eig_w", w);
    put("eig_v", v);
    // vector operations on 100 entries
    std::vector<double> y0 = seeded(5, 100), x0 = seeded(6, 100);
    double *dy = to_dev(y0), *dx = to_dev(x0);
    put("norm", {vectorNorm(dx, 100, 1, true), vectorNorm(dx, 34, 3, true), vectorNorm(x0.data(), 50, 2, false)});
    put("sum", {vectorSum(dx, 100, 1, true), vectorSum(dx, 50, 2, true), vectorSum(x0.data(), 100, 1, false)});
    vectorAdd(dy, dx, 100, true, true);
    put("add", to_host(dy, 100));
    vectorScal(-2.5, dy, 100, true);
    put("scal", to_host(dy, 100));
    std::vector<double> yh = y0;
    vectorMul(yh.data(), dx, 100, false, true);
    put("mul", yh);
    // diagonal scalings of a 37 x 29 matrix
    std::vector<double> dr = seeded(7, M), dcv = seeded(8, K);
    double* dm = to_dev(a);
    diagRowMul(dm, to_dev(dr), M, K, true, true);
    put("diagrow", to_host(dm, M * K));
    std::vector<double> am = a;
    diagColMul(am.data(), dcv.data(), M, K, false, false);
    put("diagcol", am);
    // transposes
    double* dT = to_dev(std::vector<double>(M * K));
    setTranspose(da, M, K, dT, true, true);
    put("transpose", to_host(dT, M * K));
    std::vector<double> ai = a;
    setTranspose(ai.data(), M, K, false);
    put("transpose_inplace", ai);
    // identity and diagonals of a 5 x 7 matrix
    std::vector<double> id(35);
    setIdentity(id.data(), 5, 7, false);
    put("identity", id);
    std::vector<double> g = seeded(9, 35), dg = seeded(10, 5), got(4);
    double* dG = to_dev(g);
    setDiag(dG, to_dev(dg), 5, 7, 5, true, true);
    put("setdiag", to_host(dG, 35));
    getDiag(g.data(), got.data(), 5, 7, 4, false, false);
    put("getdiag", got);
    // reshapeElem: (5,3,4,6) -> axes (2,0,3,1), the reference's offset table
    const int rank = 4;
    const size_t dims[4] = {5, 3, 4, 6}, perm[4] = {2, 0, 3, 1};
    std::vector<double> t = seeded(11, 360);
    size_t offset[8], so[4], nd[4], no[4];
    so[3] = 1;
    for (int i = 2; i >= 0; --i) so[i] = so[i + 1] * dims[i + 1];
    for (int j = 0; j < 4; ++j) nd[j] = dims[perm[j]];
    no[3] = 1;
    for (int j = 2; j >= 0; --j) no[j] = no[j + 1] * nd[j + 1];
    for (int i = 0; i < 4; ++i) offset[i] = so[i];
    for (int j = 0; j < 4; ++j) offset[rank + perm[j]] = no[j];
    double *dt = to_dev(t), *dn = to_dev(std::vector<double>(360));
    reshapeElem(dt, rank, 360, offset, dn);
    put("reshape", to_host(dn, 360));
    // element access
    setElemAt(7, 42.0, dn, true);
    put("elemat", {getElemAt(7, dn, true), getElemAt(8, dn, true)});
    // lanczosEV on a 60 x 60 symmetric matrix
    std::vector<double> xl = seeded(12, 3600), al(3600), psi = seeded(13, 60), vec(60);
    for (int i = 0; i < 60; ++i)
      for (int j = 0; j < 60; ++j) al[i * 60 + j] = xl[i * 60 + j] + xl[j * 60 + i];
    size_t iters = 200;
    double e0 = 0.0;
    const bool conv = lanczosEV(al.data(), psi.data(), 60, iters, 1e-14, e0, vec.data(), false);
    put("lanczos", {e0, (double)iters, conv ? 1.0 : 0.0});
    put("lanczos_vec", vec);
    for (double* p : {da, db, dc, ds, dU, dS, dVT, dy, dx, dm, dT, dG, dt, dn}) elemFree(p, 0, true);
    std::printf("seam_check ok\n");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "seam_check: %s\n", e.what());
    return 1;
  }
  return 0;
}
