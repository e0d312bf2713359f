// Batched iTEBD: many independent XXZ chains (spin-1/2 or spin-1, dense or
// U(1)-symmetric, of uneven bond dimension) advanced together, each stage of
// the update one launch for all chains (grouped GEMMs, batched SVD, batched
// vector operations), and a whole step replayable as one HIP graph.  Per
// chain, the update is the one of uni10/examples/iTEBD.cpp:148-173
// (itebd.net / updateA.net), written as the GEMMs those networks perform;
// a U(1) chain runs it per charge block of its symmetric tensors.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "tn_hip.h"
#include "uni10_amd/uni10.hpp"

namespace uni10 {

// Charge sectors of one bond: charges ascending, dimensions, offsets.
struct U1Sectors {
  std::vector<int> q, n;
  std::vector<int64_t> off;
  int find(int charge) const;   // sector index of a charge, -1 if absent
  int64_t total() const { return off.empty() ? 0 : off.back() + n.back(); }
};

// The two bonds of a U(1) chain of bond dimension chi (bond 1, left of site
// A, holds even charges; bond 0 the charges one site away), charges within
// [-qmax, qmax]: Gaussian sector dimensions of width 2 (1 + log2(chi) / 4)
// rounded by largest remainder (at least 1 each) on the charges whose share
// is at least 1/2, no sector larger than the theta block producing it.
void u1_bond_sectors(int chi, int d, int qmax, U1Sectors& bond0, U1Sectors& bond1);

class BatchTebd {
 public:
  // chain i: bond dimension chi[i], anisotropy delta[i], seed[i] (the same
  // seeded start as ITebd for a dense chain), physical dimension d[i] (2:
  // spin 1/2, 3: spin 1; empty = all 2), sym[i] (0 dense, 1 U(1); empty =
  // all dense); time step tau; .net files for the energy.
  BatchTebd(const std::vector<int>& chi, const std::vector<double>& delta, const std::vector<uint64_t>& seed,
            double tau, int gemm_mode, const std::string& netdir, const std::vector<int>& d = {},
            const std::vector<int>& sym = {});
  ~BatchTebd();
  BatchTebd(const BatchTebd&) = delete;
  BatchTebd& operator=(const BatchTebd&) = delete;

  void step(int n = 1);            // n x (update(A,B), update(B,A)) on every chain
  // step without waiting: the engine's work goes to its own stream, so the
  // steps of several engines overlap on the device; wait() finishes them
  // and checks the SVDs' convergence
  void launch(int n = 1);
  void wait();
  // true once every launched step has finished (never blocks)
  bool ready() const;
  // restart chain `chain` from the seeded start of (delta, seed): the slot
  // reuse of continuous batching; other chains are untouched
  void load(int chain, double delta, uint64_t seed);
  void set_graph(bool on);         // replay steps through a captured HIP graph
  // An inactive chain is skipped by every launch (its state is kept); takes
  // effect from the next launch.  All chains start active.
  void set_active(int chain, bool on);
  // Run the engine's work on the caller's stream from now on (NULL: back to
  // its own); the engine must be idle.  Engines that share a stream run one
  // after another, in launch order.
  void set_stream(void* stream);
  double energy(int chain);        // energy per site (networks, one chain)
  // 0: lambda_A, 1: lambda_B (a U(1) chain: sector by sector, charges ascending)
  std::vector<double> lambda(int chain, int which) const;
  double trunc_err(int chain) const;   // discarded weight of the chain's last update
  const U1Sectors* sectors(int chain, int bond) const;   // nullptr for a dense chain
  int chains() const { return (int)chi_.size(); }
  double gemm_flops_per_step() const { return flops_; }
  int last_sweeps() const { return sweeps_; }

 private:
  struct Off {
    int64_t x[2], l[2], g, t1[2], cb[2], th[2], s[2], vt[2], sc[2], tmp[2];
  };
  // Block layout of a U(1) chain.  Tensor p (Gamma_p lambda_p) has left bond
  // 1 - p and right bond p and is stored by left-bond sector: block i holds
  // rows a of that sector and, for each physical state s in turn, the
  // columns m of the right-bond sector q_i + s.  Update p factorises theta
  // block j (charge q_j of bond p): rows (s, a in outer sector q_j - s),
  // columns (t, b in outer sector q_j + t), D_j each way, s and t in turn.
  struct U1Chain {
    std::vector<int> phys;
    U1Sectors bond[2];
    std::vector<int64_t> xblk[2], xw[2];             // tensor p: block offsets and row widths
    std::vector<std::vector<int64_t>> xpart[2];      // tensor p, block i: column offset of state s (-1: none)
    std::vector<int64_t> D[2], sq[2], sv[2], disc[2];   // update p, block j: size, D x D and vector offsets
    std::vector<std::vector<int64_t>> part[2];       // update p, block j: row/column offset of state s (-1: none)
    std::vector<int64_t> pair[2];                    // update p: T1 array of (outer sector i, sigma) (-1: none)
    int64_t xsize[2] = {0, 0}, sqsize[2] = {0, 0}, svsize[2] = {0, 0}, t1size[2] = {0, 0}, tmpsize[2] = {0, 0};
  };
  void layout_u1(int i);
  void plans_dense(int i, int p, std::vector<tn_gemm_desc_ex>& d1, std::vector<tn_gemm_desc_ex>& d2,
                   std::vector<tn_gemm_desc_ex>& d3, std::vector<tn_vecop>& gather,
                   std::vector<tn_vecop>& trunc, std::vector<tn_vecop>& scale, std::vector<int64_t>* svd);
  void plans_u1(int i, int p, std::vector<tn_gemm_desc_ex>& d1, std::vector<tn_gemm_desc_ex>& d2,
                std::vector<tn_gemm_desc_ex>& d3, std::vector<tn_vecop>& gather,
                std::vector<tn_vecop>& trunc, std::vector<tn_vecop>& scale, std::vector<int64_t>* svd);
  void update(int p);
  void release() noexcept;   // streams, plans, graph and flags (the arena frees itself)
  void build(const std::vector<uint64_t>& seed, double tau, int gemm_mode);
  void init_chain(int i, double delta, uint64_t seed, double* base, const std::vector<double>* gate = nullptr);
  std::vector<std::vector<double>> gates(const std::vector<int>& chains, const std::vector<double>& delta) const;
  std::vector<int64_t> region_;   // arena doubles of each chain (from its x[0])
  double tau_ = 0;
  void run_step();
  std::vector<int> chi_, d_, sym_;
  std::vector<double> delta_;
  std::vector<Off> off_;
  std::vector<U1Chain> u1_;       // per chain (empty layout for dense chains)
  DevBuf arena_;
  void* stream_ = nullptr;              // where the engine's work runs
  void* own_stream_ = nullptr;          // the engine's own stream
  tn_gemm_plan g1_[2] = {nullptr, nullptr}, g2_[2] = {nullptr, nullptr}, g3_[2] = {nullptr, nullptr};
  tn_svd_plan svd_[2] = {nullptr, nullptr};
  int* active_ = nullptr;               // device activity flags, one per chain
  std::vector<int> active_host_;
  bool active_dirty_ = false;
  tn_vecops_plan gather_[2] = {nullptr, nullptr}, trunc_[2] = {nullptr, nullptr}, scale_[2] = {nullptr, nullptr};
  tn_svd_plan last_svd_ = nullptr;
  bool warm_[2] = {false, false};
  bool use_graph_ = false;
  void* graph_ = nullptr;
  double flops_ = 0;
  int sweeps_ = 0;
  std::string netdir_;
};

}  // namespace uni10
