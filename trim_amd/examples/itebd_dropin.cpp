// iTEBD of the spin-1/2 XXZ chain written the way a Uni10 caller writes it:
// only the public C++ API of uni10_amd/uni10.hpp, in the call pattern of the
// reference example (uni10/examples/iTEBD.cpp: update :148-173, measure
// :189-203, expectation :205-216) -- Networks, getBlock/putBlock, diagonal
// Matrix products, Matrix::svd, resize, norm, assign, permute.  It shows that
// caller code moves over by changing its include line; every operation runs
// on the device through the library.
//
//   itebd_dropin CHI STEPS SEED TAU DELTA [NETDIR]
//
// Start: Gamma_A, Gamma_B (CHI, 2, CHI) ~ SplitMix64 uniform[-1, 1) (seeds
// SEED, SEED+1), lambdas sorted |uniform| (seeds SEED+2, SEED+3), unit norm.
// Prints the energy per site before and after STEPS steps (a step is
// update(A, B) followed by update(B, A)).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "uni10_amd/uni10.hpp"

using namespace uni10;

static double uniform_pm1(uint64_t seed, uint64_t i) {
  uint64_t z = seed + i * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

static int CHI = 0;

static void update(UniTensor& ALa, UniTensor& BLb, std::map<Qnum, Matrix>& La, std::map<Qnum, Matrix>& Lb,
                   UniTensor& expH, Network& iTEBD, Network& updateA) {
  Qnum q0(0);
  iTEBD.putTensor("ALa", &ALa);
  iTEBD.putTensor("BLb", &BLb);
  iTEBD.putTensor("expH", &expH);
  UniTensor C = iTEBD.launch();
  UniTensor Theta(C.bond(), "Theta");
  Theta.putBlock(q0, Lb[q0] * C.getBlock(q0));
  Theta.permute(2);
  std::vector<Matrix> usv = Theta.getBlock(q0).svd();
  const int dim = CHI < (int)usv[1].row() ? CHI : (int)usv[1].row();
  const double norm = usv[1].resize(dim, dim).norm();
  usv[1] *= (1.0 / norm);
  La[q0] = usv[1];
  std::vector<Bond> bond3 = ALa.bond();
  bond3[0] = Bond(BD_IN, dim);
  BLb.assign(bond3);
  Matrix blk = BLb.getBlock(q0);
  blk.setElem(usv[2].getElem(), usv[2].isOngpu());
  BLb.putBlock(q0, blk);
  updateA.putTensor("BLb", &BLb);
  updateA.putTensor("C", &C);
  ALa = updateA.launch();
  ALa *= (1.0 / norm);
}

static double expectation(UniTensor& L, UniTensor& R, UniTensor& Op, Network& MPS, Network& meas) {
  Qnum q0(0);
  MPS.putTensor("L", &L);
  MPS.putTensor("R", &R);
  UniTensor psi = MPS.launch();
  double norm = psi.getBlock(q0).norm();
  norm *= norm;
  meas.putTensor("bra", &psi);
  meas.putTensorT("ket", &psi);
  meas.putTensor("Op", &Op);
  return meas.launch()[0] / norm;
}

static double measure(UniTensor& ALa, UniTensor& BLb, std::map<Qnum, Matrix>& La, std::map<Qnum, Matrix>& Lb,
                      UniTensor& Op, Network& MPS, Network& meas) {
  Qnum q0(0);
  UniTensor A1(ALa);
  A1.permute(1);
  A1.putBlock(q0, Lb[q0] * A1.getBlock(q0));
  A1.permute(2);
  UniTensor B1(BLb);
  B1.permute(1);
  B1.putBlock(q0, La[q0] * B1.getBlock(q0));
  B1.permute(2);
  double val = expectation(A1, BLb, Op, MPS, meas);
  val += expectation(B1, ALa, Op, MPS, meas);
  return val / 2;
}

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s CHI STEPS SEED TAU DELTA [NETDIR]\n", argv[0]);
    return 2;
  }
  CHI = std::atoi(argv[1]);
  const int steps = std::atoi(argv[2]);
  const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
  const double tau = std::atof(argv[4]), delta = std::atof(argv[5]);
  const std::string dir = argc > 6 ? std::string(argv[6]) + "/" : std::string();
  const int d = 2;
  try {
    // two-site XXZ Hamiltonian, basis |uu>, |ud>, |du>, |dd>
    std::vector<double> h(16, 0.0);
    h[0] = h[15] = 0.25 * delta;
    h[5] = h[10] = -0.25 * delta;
    h[6] = h[9] = 0.5;
    std::vector<Bond> bondH = {Bond(BD_IN, d), Bond(BD_IN, d), Bond(BD_OUT, d), Bond(BD_OUT, d)};
    UniTensor H(bondH, "H");
    H.setRawElem(h);
    UniTensor U(bondH, "U");
    U.putBlock(Qnum(0), takeExp(-tau, H.getBlock(Qnum(0))));

    const size_t n = (size_t)CHI * d * CHI;
    std::vector<double> la(CHI), lb(CHI), ala(n), blb(n);
    for (int i = 0; i < CHI; ++i) {
      la[i] = std::fabs(uniform_pm1(seed + 2, i + 1));
      lb[i] = std::fabs(uniform_pm1(seed + 3, i + 1));
    }
    std::sort(la.rbegin(), la.rend());
    std::sort(lb.rbegin(), lb.rend());
    double na = 0.0, nb = 0.0;
    for (int i = 0; i < CHI; ++i) {
      na += la[i] * la[i];
      nb += lb[i] * lb[i];
    }
    for (int i = 0; i < CHI; ++i) {
      la[i] /= std::sqrt(na);
      lb[i] /= std::sqrt(nb);
    }
    // ALa(a, b | s) = Gamma_A(a, s, b) lambda_A(b), likewise BLb
    for (int a = 0; a < CHI; ++a)
      for (int s = 0; s < d; ++s)
        for (int b = 0; b < CHI; ++b) {
          const size_t src = ((size_t)a * d + s) * CHI + b, dst = ((size_t)a * CHI + b) * d + s;
          ala[dst] = uniform_pm1(seed, src + 1) * la[b];
          blb[dst] = uniform_pm1(seed + 1, src + 1) * lb[b];
        }
    std::vector<Bond> bond3 = {Bond(BD_IN, CHI), Bond(BD_IN, CHI), Bond(BD_OUT, d)};
    UniTensor ALa(bond3, "ALa"), BLb(bond3, "BLb");
    ALa.setRawElem(ala);
    BLb.setRawElem(blb);
    std::map<Qnum, Matrix> La, Lb;
    La[Qnum(0)] = Matrix(CHI, CHI, la.data(), true);
    Lb[Qnum(0)] = Matrix(CHI, CHI, lb.data(), true);

    Network iTEBD(dir + "itebd.net"), updateA(dir + "updateA.net"), MPS(dir + "MPS.net"), meas(dir + "measure.net");
    std::printf("energy0 %.17g\n", measure(ALa, BLb, La, Lb, H, MPS, meas));
    for (int step = 0; step < steps; ++step) {
      update(ALa, BLb, La, Lb, U, iTEBD, updateA);
      update(BLb, ALa, Lb, La, U, iTEBD, updateA);
    }
    std::printf("energy %.17g\n", measure(ALa, BLb, La, Lb, H, MPS, meas));
    std::vector<double> a = La[Qnum(0)].getHostElem(), b = Lb[Qnum(0)].getHostElem();
    std::printf("lambdaA");
    for (double v : a) std::printf(" %.17g", v);
    std::printf("\nlambdaB");
    for (double v : b) std::printf(" %.17g", v);
    std::printf("\n");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "itebd_dropin: %s\n", e.what());
    return 1;
  }
  return 0;
}
