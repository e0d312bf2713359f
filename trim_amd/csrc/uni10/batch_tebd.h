// Batched iTEBD: many independent spin-1/2 XXZ chains of uneven bond
// dimension advanced together, each stage of the update one launch for all
// chains (grouped GEMMs, batched SVD, batched vector operations), and a whole
// step replayable as one HIP graph.  Per chain, the update is the one of
// uni10/examples/iTEBD.cpp:148-173 (itebd.net / updateA.net), written as the
// GEMMs those networks perform.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "tn_hip.h"
#include "uni10_amd/uni10.hpp"

namespace uni10 {

class BatchTebd {
 public:
  // chain i: bond dimension chi[i], anisotropy delta[i], seed[i] (the same
  // seeded start as ITebd); time step tau; .net files for the energy.
  BatchTebd(const std::vector<int>& chi, const std::vector<double>& delta, const std::vector<uint64_t>& seed,
            double tau, int gemm_mode, const std::string& netdir);
  ~BatchTebd();
  BatchTebd(const BatchTebd&) = delete;
  BatchTebd& operator=(const BatchTebd&) = delete;

  void step(int n = 1);            // n x (update(A,B), update(B,A)) on every chain
  // restart chain `chain` from the seeded start of (delta, seed): the slot
  // reuse of continuous batching; other chains are untouched
  void load(int chain, double delta, uint64_t seed);
  void set_graph(bool on);         // replay steps through a captured HIP graph
  double energy(int chain);        // energy per site (networks, one chain)
  std::vector<double> lambda(int chain, int which) const;   // 0: lambda_A, 1: lambda_B
  double trunc_err(int chain) const;   // discarded weight of the chain's last update
  int chains() const { return (int)chi_.size(); }
  double gemm_flops_per_step() const { return flops_; }
  int last_sweeps() const { return sweeps_; }

 private:
  struct Off {
    int64_t x[2], l[2], g, t1, cb, th, u, s, vt[2], sc[2];
  };
  void update(int p, bool warm);
  void init_chain(int i, double delta, uint64_t seed, double* base);
  std::vector<int64_t> region_;   // arena doubles of each chain (from its x[0])
  double tau_ = 0;
  void run_step();
  std::vector<int> chi_;
  std::vector<double> delta_;
  std::vector<Off> off_;
  DevBuf arena_;
  void* stream_ = nullptr;
  tn_gemm_plan g1_[2] = {nullptr, nullptr}, g2_[2] = {nullptr, nullptr}, g3_[2] = {nullptr, nullptr};
  tn_svd_plan svd_cold_[2] = {nullptr, nullptr}, svd_warm_[2] = {nullptr, nullptr};
  tn_vecops_plan rows_[2] = {nullptr, nullptr}, trunc_[2] = {nullptr, nullptr}, scale_[2] = {nullptr, nullptr};
  tn_svd_plan last_svd_ = nullptr;
  bool warm_[2] = {false, false};
  bool use_graph_ = false;
  void* graph_ = nullptr;
  double flops_ = 0;
  int sweeps_ = 0;
  std::string netdir_;
};

}  // namespace uni10
