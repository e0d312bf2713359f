// Block one-sided Jacobi SVD: the path for matrices too large for one
// workgroup's LDS.  Internal to libtn_hip.so (used by the SVD plans).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace tn {
namespace bj {

// One problem as the caller describes it (A is M x N row-major).
struct Spec {
  int64_t M, N;
  int64_t a_off, u_off, s_off, vt_off;
  int64_t v0_off;   // warm start from an N x N V0^T at this offset (M >= N), -1 = cold
};

struct Plan;

// Builds the device descriptors, scratch and (for warm starts) the W = A V0
// product plan.  max_sweeps > 0 lowers the sweep cap (40) past which a
// problem is reported as not converged.  Returns a TN_* code.
int build(Plan** out, const std::vector<Spec>& specs, int max_sweeps = 0);
// Launches the whole solve on `st`; status[0] += problems not converged,
// status[1] = max(status[1], sweeps).  Under stream capture the sweep count
// is fixed (max_sweeps_captured); otherwise the host stops once a sweep
// rotates nothing (one 4-byte read per sweep).
int launch(Plan* P, const double* A, const double* V0T, double* U, double* S, double* VT, int* status,
           hipStream_t st);
void destroy(Plan* P);
// Problem k (in the order of the specs) is skipped by every kernel while
// mask[idx[k]] == 0 (idx[k] < 0 or mask == NULL: never skipped); a skipped
// problem's outputs and scratch keep their contents.  Returns a TN_* code.
int set_mask(Plan* P, const int* mask, const std::vector<int64_t>& idx);
bool empty(const Plan* P);

}  // namespace bj
}  // namespace tn
