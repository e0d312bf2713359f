// iTEBD driver over the Uni10 API (the reference example's update/measure).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "tn_hip.h"
#include "uni10_amd/uni10.hpp"

namespace uni10 {

Matrix xxz_two_site(double delta, double J);
std::vector<double> xxz_two_site_host(double delta, double J);         // row-major 4 x 4
std::vector<double> xxz_two_site_spin1_host(double delta, double J);   // row-major 9 x 9
// Two-site spin-1 XXZ J ((S+S- + S-S+)/2 + delta Sz Sz), basis |s t> with
// s, t in (+1, 0, -1), s the slower index.
Matrix xxz_two_site_spin1(double delta, double J);
// Energy per site of the iTEBD state (ALa, BLb, La, Lb) for the two-site
// Hamiltonian H, through the MPS.net / measure.net networks.
double itebd_energy(UniTensor& ALa, UniTensor& BLb, const Matrix& La, const Matrix& Lb, UniTensor& H,
                    Network& MPS, Network& meas);

class ITebd {
 public:
  ITebd(int chi, int d, double delta, double tau, uint64_t seed, const std::string& netdir);
  ~ITebd();
  ITebd(const ITebd&) = delete;
  ITebd& operator=(const ITebd&) = delete;
  void step();                 // update(A,B) then update(B,A)
  // Waits for the device and throws if an SVD since the last check did not
  // converge (the updates themselves never wait for the host); energy()
  // checks first.
  void check();
  double energy();             // mean of the two bond energies
  double truncation_error() const;   // discarded weight of the last update
  // replay the update networks as captured graphs (Network::setGraph)
  void set_graph(bool on) {
    iTEBD_.setGraph(on);
    updateA_.setGraph(on);
  }
  const Matrix& La() const { return La_; }
  const Matrix& Lb() const { return Lb_; }
  UniTensor& ALa() { return ALa_; }
  UniTensor& BLb() { return BLb_; }

 private:
  void update(UniTensor& ALa, UniTensor& BLb, Matrix& La, Matrix& Lb);
  std::vector<Matrix> svd_warm(const Matrix& T);
  // warm-started SVD plan of each bond's theta (kept while the shape holds)
  tn_svd_plan plan_[2] = {nullptr, nullptr};
  int64_t plan_shape_[2][2] = {{-1, -1}, {-1, -1}};
  int chi_, d_;
  double tau_;
  DevBuf scal_{2, true};        // ||S[:dim]|| and the discarded weight of the last update
  Matrix vt_prev_[2];          // last VT of each bond's theta (warm starts)
  int parity_ = 0;
  UniTensor H_, U_, ALa_, BLb_;
  Matrix La_, Lb_;
  Network iTEBD_, updateA_, MPS_, meas_;
};

}  // namespace uni10
