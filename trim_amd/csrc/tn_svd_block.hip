// Block one-sided Jacobi SVD for matrices that do not fit one workgroup's
// LDS (the iTEBD theta of bond dimension 48 and up).
//
// The working matrix W (A, A^T or A V0; column-major, rows padded to a
// multiple of 16, columns to a multiple of 32 with zero columns) is split
// into column blocks of 16.  A sweep is a round-robin tournament over the
// blocks; in each round every block pair (P, Q) of every problem is handled
// by one workgroup, all in one launch:
//   G = [W_P W_Q]^T [W_P W_Q]                 32 x 32, FP64 MFMA over the rows
//   if a column pair of the panel is not yet orthogonal:
//     Q^T G Q = diag                            cyclic Jacobi on G in LDS
//     [W_P W_Q] <- [W_P W_Q] Q, [V_P V_Q] <- [V_P V_Q] Q    FP64 MFMA
// Each sweep first orders the columns by decreasing norm (de Rijk), so that
// a block holds columns of similar norm; this halves the sweeps a graded
// spectrum needs.  Every column pair meets in some block pair of a sweep,
// so a sweep that rotates nothing leaves W with orthogonal columns: the
// singular values are the column norms of W, U its normalised columns and V
// the accumulated rotations (A = W V^T).  The reference computes the same factorisation with
// dgesvd (uni10_lapack_cpu.cpp:289-313) or cuSOLVER gesvd
// (uni10_lapack_gpu.cu:418-500).
//
// Problems that need only S and VT (no U; M >= N, M <= 4096) are first
// factorised A V0 = Q R by a blocked Householder QR (32-column panels, 16
// beyond 1024 rows, 8 beyond 2048: one
// workgroup factors a panel and builds its compact-WY factor T, many
// workgroups apply I - Y T^T Y^T to the trailing columns), and the sweeps
// run on X = R^T, whose column Gram matrix is close to diagonal (Drmac and
// Veselic, SIAM J. Matrix Anal. Appl. 29 (2008)): TEBD thetas need about
// half the sweeps.  No rotation is accumulated: from X = U_X S V_X^T,
// A = (Q V_X) S (V0 U_X)^T, so VT = U_X^T V0^T (one GEMM at the end), with
// U_X the normalised columns of X (an exactly zero column is completed to
// an orthonormal basis by Gram-Schmidt).  These problems are orthogonalised
// relative to their own column norms only (no absolute floor), which keeps
// U_X orthonormal down to the rounding-level columns of a rank-deficient A.
//
// Convergence: a pair (p, q) is left alone when |w_p.w_q| is below
// 10 m eps |w_p||w_q| (the rounding noise of a dot product of orthogonal
// columns) or below m eps^2 ||A||_F^2 (columns at rounding level), the
// criterion of the one-workgroup kernels.  The Jacobi rotations of the
// panel's Gram matrix keep the relative structure of a graded panel (a
// rotation between a large and a small column has a small angle, computed
// to relative accuracy), so small columns are not swamped by the panel's
// large ones and the same criterion is met.
#include "tn_svd_block.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "tn_internal.h"

namespace tn {
namespace bj {
namespace {

constexpr int kB = 16;            // columns per block
constexpr int kPW = 2 * kB;       // panel width (two blocks)
constexpr int kThreads = 256;
constexpr int kMaxSweeps = 40;
constexpr int kMaxSweepsCaptured = kMaxSweeps;
constexpr int kInnerSweeps = 1;
constexpr int kSetupBlocks = 64;
constexpr int kNormBlocks = 16;
constexpr int kMaxN = 8192;       // the sort keeps n norms in LDS (64 KiB)
// QR panels are held in registers, 128 doubles per thread: 32 columns x 4
// rows (M <= 1024), 16 columns x 8 rows (M <= 2048) or 8 columns x 16 rows
// (M <= 4096)
constexpr int kQrMaxM = 4096;
constexpr int kQrWideMaxM = 1024;
constexpr int kQrMidMaxM = 2048;

typedef double d4 __attribute__((ext_vector_type(4)));

struct Prob {
  int64_t M, N;                           // caller's A: M x N row-major
  int64_t a_off, u_off, s_off, vt_off, v0_off;
  int64_t w_off, v_off, sig_off;          // scratch (doubles): W, V, column norms + ||A||_F^2
  int64_t y_off, t_off, ut_off;           // QR: panel reflectors (mp x 32), T (32 x 32), U_X^T (n x n)
  int32_t m, n;                           // working W: m x n, m >= n
  int32_t mp, nvp, np, nb;                // padded rows of W / V, padded columns, blocks
  int32_t tr;                             // W = A^T
  int32_t flag_off, rank_off, pos_off;    // int scratch: rotated-in-sweep flags, ranks, sweep order
  int32_t qr;                             // QR-preconditioned: sweeps on R^T, V holds V0
  int32_t qpw;                            // its QR panel width (32, 16 beyond 1024 rows, 8 beyond 2048)
  int32_t ms;                             // rows the sweeps touch (mp; n rounded up to 16 for QR)
  const int* mask;                        // with mask_idx >= 0: skipped while mask[mask_idx] == 0
  int64_t mask_idx;
};

__device__ inline bool prob_off(const Prob& p) {
  return p.mask != nullptr && p.mask_idx >= 0 && p.mask[p.mask_idx] == 0;
}

__device__ inline void rr_pair(int r, int k, int np, int& p, int& q) {
  const int L = np - 1;
  if (k == 0) {
    p = r;
    q = L;
  } else {
    p = (r + k) % L;
    q = (r - k + L) % L;
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

__device__ inline bool needs_rot(double gpq, double gpp, double gqq, double tol, double abs_tol) {
  // |g_pq| > tol sqrt(g_pp g_qq), compared in squares
  return gpq * gpq > (tol * tol) * (fmax(gpp, 0.0) * fmax(gqq, 0.0)) && fabs(gpq) > abs_tol;
}

__device__ inline double fast_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

__device__ inline double fast_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  y = y * fma(-0.5 * x * y, y, 1.5);
  return y;
}

// Pair k of step r of the round-robin over the 32 panel columns (p < q).
__device__ inline void panel_pair(int r, int k, int& p, int& q) {
  constexpr int L = kPW - 1;
  if (k == 0) {
    p = r;
    q = L;
  } else {
    p = r + k;
    p -= (p >= L) ? L : 0;
    q = r - k;
    q += (q < 0) ? L : 0;
    if (p > q) {
      const int x = p;
      p = q;
      q = x;
    }
  }
}

// Jacobi rotation of the pair from its G entries (c = 1, s = t = 0 when the
// pair is orthogonal enough): J^T G J zeroes G[p][q] with J = [c s; -s c] on
// rows/columns p, q.
__device__ inline void rotation(double app, double aqq, double apq, double tol, double abs_tol, double& c,
                                double& s, double& t) {
  c = 1.0;
  s = 0.0;
  t = 0.0;
  if (needs_rot(apq, app, aqq, tol, abs_tol)) {
    // The angle need not be exact (the sweeps converge on what is left), so
    // zeta and t use the hardware reciprocal estimates as they are; c is
    // refined to full precision so that c^2 + s^2 = 1 to rounding and every
    // rotation is orthogonal.
    const double zeta = (aqq - app) * (0.5 * __builtin_amdgcn_rcp(apq));
    const double z2 = fma(zeta, zeta, 1.0);
    t = copysign(1.0, zeta) * __builtin_amdgcn_rcp(fabs(zeta) + z2 * __builtin_amdgcn_rsq(z2));
    c = fast_rsqrt(fma(t, t, 1.0));
    s = c * t;
  }
}

__device__ inline double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
#pragma unroll
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// W (and V) of every problem, padding included: W = A, A^T or zero (a warm
// start's W = A V0 is written next by the GEMM), V = V0 or I.  Clears the
// problem's sweep flags and the launch-wide ones.
__global__ void __launch_bounds__(kThreads) bj_setup(const Prob* __restrict__ probs, const double* __restrict__ A,
                                                     const double* __restrict__ V0T, double* __restrict__ scr,
                                                     int* __restrict__ iscr) {
  const Prob pr = probs[blockIdx.y];
  if (prob_off(pr)) {   // the shared per-sweep flags are still cleared
    if (blockIdx.x == 0 && blockIdx.y == 0)
      for (int s = threadIdx.x; s < kMaxSweeps; s += kThreads) iscr[s] = 0;
    return;
  }
  const bool warm = pr.v0_off >= 0;
  double* W = scr + pr.w_off;
  double* V = scr + pr.v_off;
  const int64_t totW = (int64_t)pr.mp * pr.np, totV = (int64_t)pr.nvp * pr.np;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < totW + totV; e += stride) {
    if (e < totW) {
      const int64_t c = e / pr.mp, r = e % pr.mp;
      double v = 0.0;
      if (!warm && r < pr.m && c < pr.n) v = pr.tr ? A[pr.a_off + c * pr.N + r] : A[pr.a_off + r * pr.N + c];
      W[e] = v;
    } else {
      const int64_t e2 = e - totW;
      const int64_t c = e2 / pr.nvp, r = e2 % pr.nvp;
      double v = 0.0;
      if (r < pr.n && c < pr.n) v = warm ? V0T[pr.v0_off + c * pr.n + r] : (r == c ? 1.0 : 0.0);
      V[e2] = v;
    }
  }
  if (blockIdx.x == 0) {
    for (int s = threadIdx.x; s < kMaxSweeps; s += kThreads) {
      iscr[pr.flag_off + s] = 0;
      if (blockIdx.y == 0) iscr[s] = 0;
    }
  }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Exact column norms of W (squared), one wavefront per column, the rows
// strided over the lanes (coalesced) and summed by a fixed butterfly.
__global__ void __launch_bounds__(kThreads) bj_colnorm(const Prob* __restrict__ probs, double* __restrict__ scr) {
  const Prob pr = probs[blockIdx.y];
  if (prob_off(pr)) return;
  const double* W = scr + pr.w_off;
  double* nrm = scr + pr.sig_off;
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const int nw = gridDim.x * (kThreads / 64);
  for (int j = gw; j < pr.n; j += nw) {
    const double* wj = W + (int64_t)j * pr.mp;
    double s = 0.0;
    const int rows = pr.qr ? pr.n : pr.m;
    for (int i = lane; i < rows; i += 64) s += wj[i] * wj[i];
    s = wave_sum(s);
    if (lane == 0) nrm[j] = s;
  }
}

// Rank of every column by decreasing norm (ties by index), from the norms
// in LDS: pos[rank] = column (sort) or rank[column] = rank with S written
// (final).  n doubles of dynamic LDS.
__device__ inline int norm_rank(const double* sh, int n, int j) {
  const double sj = sh[j];
  int rk = 0;
  for (int i = 0; i < n; ++i) {
    const double si = sh[i];
    rk += (si > sj) || (si == sj && i < j);
  }
  return rk;
}

// Start of a sweep: the columns ordered by decreasing norm (padding columns
// last).  Block b of the sweep is the columns at positions [16b, 16b + 16)
// of this order, so a block holds columns of similar norm.  The norms are
// the exact ones (first sweep) or the diagonals of the panels' rotated Gram
// matrices left by the last round.  The first sweep also records
// ||A||_F^2 = sum of the column norms (fixed order).
__global__ void __launch_bounds__(kThreads) bj_sort(const Prob* __restrict__ probs, double* __restrict__ scr,
                                                    int* __restrict__ iscr, int sweep) {
  extern __shared__ double sh[];
  __shared__ double red[kThreads];
  const Prob pr = probs[blockIdx.x];
  if (prob_off(pr)) return;
  if (sweep > 0 && iscr[pr.flag_off + sweep - 1] == 0) return;
  const double* nrm = scr + pr.sig_off;
  int* pos = iscr + pr.pos_off;
  const int tid = threadIdx.x;
  for (int j = tid; j < pr.n; j += kThreads) sh[j] = nrm[j];
  __syncthreads();
  if (sweep == 0) {
    double s = 0.0;
    for (int j = tid; j < pr.n; j += kThreads) s += sh[j];
    s = block_sum(s, red);
    if (tid == 0) scr[pr.sig_off + pr.n] = s;
  }
  for (int j = tid; j < pr.np; j += kThreads) pos[j < pr.n ? norm_rank(sh, pr.n, j) : j] = j;
}

// One round of a sweep: workgroup = (problem, block pair).  Registers are
// capped for four waves per SIMD (four workgroups per CU with 26 KB of LDS
// each): a round is latency-bound per workgroup, so the CU needs several.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) bj_round(const Prob* __restrict__ probs, const int2* __restrict__ wmap,
                                                     double* __restrict__ scr, int* __restrict__ iscr, int sweep,
                                                     int round) {
  // LDS (26 KB, so that several workgroups share a CU): G's first buffer,
  // and one region that holds the waves' partial Gram tiles until G is
  // formed and then G's second buffer and Q
  __shared__ double g0[kPW][kPW + 1];
  __shared__ double un[2][kPW][kPW + 1];
  __shared__ double r_c[kB], r_s[kB], r_t[kB];   // the step's rotations
  auto Gb = [&](int b) { return b ? un[0] : g0; };   // G's two buffers
  double(*Qm)[kPW + 1] = un[1];
  double(*part)[3][256] = reinterpret_cast<double(*)[3][256]>(&un[0][0][0]);   // [2][3][256]
  static_assert(2 * 3 * 256 <= 2 * kPW * (kPW + 1), "partial tiles must fit the shared region");

  const int2 wm = wmap[blockIdx.x];
  const Prob pr = probs[wm.x];
  if (prob_off(pr)) return;
  if (round >= pr.nb - 1) return;
  int* flags = iscr + pr.flag_off;
  if (sweep > 0 && flags[sweep - 1] == 0) return;   // converged in an earlier sweep
  int P, Q;
  rr_pair(round, wm.y, pr.nb, P, Q);
  const int64_t mp = pr.mp, nvp = pr.nvp;
  double* W = scr + pr.w_off;
  double* V = scr + pr.v_off;
  const double normF2 = scr[pr.sig_off + pr.n];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
  // panel column c < 16 is W column pos[16P + c], c >= 16 is pos[16Q + c - 16]
  const int* posP = iscr + pr.pos_off + P * kB;
  const int* posQ = iscr + pr.pos_off + Q * kB;
  const int64_t myP = posP[lr], myQ = posQ[lr];

  // ---- G = X^T X: three 16 x 16 tiles (PP, PQ, QQ), rows split over the waves
  {
    // Rows go in 8-row groups, dealt round-robin to the waves.  In a group
    // starting at row r, lane (lr, lg) loads rows r + 2lg and r + 2lg + 1 of
    // panel column lr (one 16-byte load per block) and supplies them to two
    // MFMA k-steps: the sum over rows of G = X^T X takes them in any order.
    const double* xp = W + myP * mp + 2 * lg;
    const double* xq = W + myQ * mp + 2 * lg;
    d4 aPP = {0.0, 0.0, 0.0, 0.0}, aPQ = aPP, aQQ = aPP;
    const int ngr = pr.ms >> 3;
    auto ld2 = [](const double* p) { return *reinterpret_cast<const double2*>(p); };
    auto mm = [&](double p, double q) {
      aPP = __builtin_amdgcn_mfma_f64_16x16x4f64(p, p, aPP, 0, 0, 0);
      aPQ = __builtin_amdgcn_mfma_f64_16x16x4f64(p, q, aPQ, 0, 0, 0);
      aQQ = __builtin_amdgcn_mfma_f64_16x16x4f64(q, q, aQQ, 0, 0, 0);
    };
    // two groups per iteration (g, g + 4); the next two are loaded first
    double2 pv[2], qv[2];
    int g = w;
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (g + 4 * h < ngr) {
        pv[h] = ld2(xp + 8 * (g + 4 * h));
        qv[h] = ld2(xq + 8 * (g + 4 * h));
      }
    for (; g < ngr; g += 8) {
      double2 pn[2], qn[2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (g + 8 + 4 * h < ngr) {
          pn[h] = ld2(xp + 8 * (g + 8 + 4 * h));
          qn[h] = ld2(xq + 8 * (g + 8 + 4 * h));
        }
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (g + 4 * h < ngr) {
          mm(pv[h].x, qv[h].x);
          mm(pv[h].y, qv[h].y);
        }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        pv[h] = pn[h];
        qv[h] = qn[h];
      }
    }
    // accumulator element v of lane (lr, lg) is tile entry (lg + 4v, lr).
    // Waves 2 and 3 publish their tiles, waves 0 and 1 add them to theirs
    // (w + 2 into w) and publish the sums.
    if (w >= 2) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int idx = (lg + 4 * v) * 16 + lr;
        part[w - 2][0][idx] = aPP[v];
        part[w - 2][1][idx] = aPQ[v];
        part[w - 2][2][idx] = aQQ[v];
      }
    }
    __syncthreads();
    if (w < 2) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int idx = (lg + 4 * v) * 16 + lr;
        part[w][0][idx] = aPP[v] + part[w][0][idx];
        part[w][1][idx] = aPQ[v] + part[w][1][idx];
        part[w][2][idx] = aQQ[v] + part[w][2][idx];
      }
    }
  }
  __syncthreads();
  // the last two partials, in a fixed order; the diagonal tiles are read
  // from their upper triangle so that G is exactly symmetric
  for (int e = tid; e < kPW * kPW; e += kThreads) {
    const int i = e / kPW, j = e % kPW;
    int t, a, b;
    if (i < kB && j < kB) {
      t = 0;
      a = min(i, j);
      b = max(i, j);
    } else if (i >= kB && j >= kB) {
      t = 2;
      a = min(i, j) - kB;
      b = max(i, j) - kB;
    } else if (i < kB) {
      t = 1;
      a = i;
      b = j - kB;
    } else {
      t = 1;
      a = j;
      b = i - kB;
    }
    const int idx = a * 16 + b;
    g0[i][j] = part[0][t][idx] + part[1][t][idx];   // (wave 0 + wave 2) + (wave 1 + wave 3)
  }
  __syncthreads();

  const double tol = 10.0 * DBL_EPSILON * (double)(pr.qr ? pr.n : pr.m);
  const double abs_tol = pr.qr ? 0.0 : (double)pr.m * DBL_EPSILON * DBL_EPSILON * normF2;
  int need = 0;
  for (int e = tid; e < kPW * kPW; e += kThreads) {
    const int i = e / kPW, j = e % kPW;
    if (i < j && needs_rot(g0[i][j], g0[i][i], g0[j][j], tol, abs_tol)) need = 1;
  }
  // the panel's column norms (diagonal of G; of the rotated G below) are
  // what the next sweep orders the columns by
  double* nrm = scr + pr.sig_off;
  const int64_t dcol = (tid < kB) ? posP[tid] : (tid < kPW ? posQ[tid - kB] : 0);
  if (!__syncthreads_or(need)) {
    if (tid < kPW && dcol < pr.n) nrm[dcol] = g0[tid][tid];
    return;
  }

  // ---- cyclic Jacobi on G (16 disjoint rotations per step), Q accumulated.
  // Threads 0-15 compute the step's rotations from one buffer of G; then
  // thread (a, b) rotates the 2 x 2 block {p_a, q_a} x {p_b, q_b} into the
  // other buffer (a <= b, with its transpose; every entry is rewritten each
  // step) and rows 2b, 2b + 1 of Q's columns p_a, q_a in place.  Two sweeps:
  // the panel's columns need not be fully orthogonalised here, the next
  // outer sweep recomputes G from W.
  for (int e = tid; e < kPW * kPW; e += kThreads) Qm[e / kPW][e % kPW] = (e / kPW == e % kPW) ? 1.0 : 0.0;
  __syncthreads();
  const int ba = tid >> 4, bb = tid & 15;
  int any_rot = 0, cur = 0;
  for (int isw = 0; isw < kInnerSweeps; ++isw) {
    int rot = 0;
    for (int r = 0; r < kPW - 1; ++r) {
      const double(*Gc)[kPW + 1] = Gb(cur);
      double(*Gn)[kPW + 1] = Gb(cur ^ 1);
      // everything this thread reads of G and Q is fixed for the step: it
      // is loaded before the rotations are known, so only c and s cross the
      // barrier
      int p1, q1, p2, q2;
      panel_pair(r, ba, p1, q1);
      panel_pair(r, bb, p2, q2);
      const double x = Gc[p1][p2], y = Gc[p1][q2], z = Gc[q1][p2], u = Gc[q1][q2];
      double qx[2], qy[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        qx[h] = Qm[2 * bb + h][p1];
        qy[h] = Qm[2 * bb + h][q1];
      }
      if (tid < kB) {
        int p, q;
        panel_pair(r, tid, p, q);
        double c, sn, t;
        rotation(Gc[p][p], Gc[q][q], Gc[p][q], tol, abs_tol, c, sn, t);
        rot |= (sn != 0.0);
        r_c[tid] = c;
        r_s[tid] = sn;
        r_t[tid] = t;
      }
      __syncthreads();
      const double c1 = r_c[ba], s1 = r_s[ba];
      if (ba < bb) {
        const double c2 = r_c[bb], s2 = r_s[bb];
        const double rp0 = c1 * x - s1 * z, rp1 = c1 * y - s1 * u;
        const double rq0 = s1 * x + c1 * z, rq1 = s1 * y + c1 * u;
        const double npp = c2 * rp0 - s2 * rp1, npq = s2 * rp0 + c2 * rp1;
        const double nqp = c2 * rq0 - s2 * rq1, nqq = s2 * rq0 + c2 * rq1;
        Gn[p1][p2] = npp;
        Gn[p2][p1] = npp;
        Gn[p1][q2] = npq;
        Gn[q2][p1] = npq;
        Gn[q1][p2] = nqp;
        Gn[p2][q1] = nqp;
        Gn[q1][q2] = nqq;
        Gn[q2][q1] = nqq;
      } else if (ba == bb) {
        // (x, y; z, u) = (a_pp, a_pq; a_qp, a_qq)
        const double t1 = r_t[ba];
        const bool rt = s1 != 0.0;
        Gn[p1][p1] = rt ? x - t1 * y : x;
        Gn[q1][q1] = rt ? u + t1 * y : u;
        Gn[p1][q1] = rt ? 0.0 : y;
        Gn[q1][p1] = rt ? 0.0 : y;
      }
      if (s1 != 0.0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          Qm[2 * bb + h][p1] = c1 * qx[h] - s1 * qy[h];
          Qm[2 * bb + h][q1] = s1 * qx[h] + c1 * qy[h];
        }
      }
      cur ^= 1;
      __syncthreads();
    }
    rot = __syncthreads_or(rot);
    any_rot |= rot;
    if (!rot) break;
  }
  if (tid < kPW && dcol < pr.n) nrm[dcol] = Gb(cur)[tid][tid];
  if (!any_rot) return;
  if (tid == 0) {
    atomicOr(&flags[sweep], 1);
    atomicOr(&iscr[sweep], 1);
  }

  // ---- [X_P X_Q] <- [X_P X_Q] Q for X = W (mp rows) and V (nvp rows),
  // 16-row tiles dealt to the waves; a tile is read whole before it is
  // written, and no other wave touches its rows
  double qb[2][8];
#pragma unroll
  for (int jt = 0; jt < 2; ++jt)
#pragma unroll
    for (int s = 0; s < 8; ++s) qb[jt][s] = Qm[4 * s + lg][16 * jt + lr];
  int64_t gcol[8];   // W column of panel column 4s + lg
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int col = 4 * s + lg;
    gcol[s] = (col < kB) ? posP[col] : posQ[col - kB];
  }
  const int tW = pr.ms >> 4, tV = pr.qr ? 0 : (int)(nvp >> 4);   // QR: V holds V0
  // tile t: rows [16t, 16t + 16) of W (t < tW) or of V
  auto tile = [&](int t, double*& base, int64_t& ld, int64_t& r0) {
    if (t < tW) {
      base = W;
      ld = mp;
      r0 = (int64_t)t * 16;
    } else {
      base = V;
      ld = nvp;
      r0 = (int64_t)(t - tW) * 16;
    }
  };
  double a[8];
  if (w < tW + tV) {
    double* base;
    int64_t ld, r0;
    tile(w, base, ld, r0);
#pragma unroll
    for (int s = 0; s < 8; ++s) a[s] = base[gcol[s] * ld + r0 + lr];
  }
  for (int t = w; t < tW + tV; t += 4) {
    double* base;
    int64_t ld, r0;
    tile(t, base, ld, r0);
    // the next tile's loads go out before this tile's MFMAs
    double an[8];
    if (t + 4 < tW + tV) {
      double* nb;
      int64_t nld, nr0;
      tile(t + 4, nb, nld, nr0);
#pragma unroll
      for (int s = 0; s < 8; ++s) an[s] = nb[gcol[s] * nld + nr0 + lr];
    }
    d4 c0 = {0.0, 0.0, 0.0, 0.0}, c1 = c0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], qb[0][s], c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], qb[1][s], c1, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = r0 + lg + 4 * v;
      base[myP * ld + row] = c0[v];
      base[myQ * ld + row] = c1[v];
    }
    if (t + 4 < tW + tV) {
#pragma unroll
      for (int s = 0; s < 8; ++s) a[s] = an[s];
    }
  }
}

// ---- QR preconditioning ------------------------------------------------
// Panel `panel` (columns kb .. kb + nb, kb = 32 panel) of the Householder QR
// of W, one workgroup per problem: thread t holds rows kb + t + 256u of the
// panel in registers.  Step j: one block reduction gives x.a_c for every
// panel column c, x = column j from row kb + j down; with alpha = -sign(x_j)
// |x| the reflector is H_j = I - tau v v^T, v = x - alpha e_j, tau = 1 /
// (alpha (alpha - x_j)), and v.a_c = x.a_c - alpha a_c[kb + j] gives both the
// update of the later columns and, for the earlier reflectors y_c, the new
// column of T (H_0 .. H_{nb-1} = I - Y T Y^T; T(0:j, j) = -tau_j T(0:j, 0:j)
// Y(:, 0:j)^T v_j, the forward recurrence of LAPACK's dlarft).  Out: R in
// the panel's columns of W (rows >= kb), Y (zero above its diagonal) and T.
template <int PW, int ROWS>
__global__ void __launch_bounds__(kThreads) bj_qr_panel(const Prob* __restrict__ probs, const int* __restrict__ qlist,
                                                        double* __restrict__ scr, int panel) {
  constexpr int TPC = kThreads / PW;   // threads summing one column's partials
  __shared__ double part[PW][kThreads + 1];
  __shared__ double part2[PW][TPC + 1];
  __shared__ double res[PW], row[PW], rdiag[PW];
  __shared__ double Ts[PW][PW + 1];
  const Prob pr = probs[qlist[blockIdx.x]];
  if (prob_off(pr)) return;
  const int kb = panel * PW;
  if (kb >= pr.n) return;
  const int nb = min(PW, pr.n - kb);
  const int m = pr.m, tid = threadIdx.x;
  const int64_t mp = pr.mp;
  double* W = scr + pr.w_off;
  for (int e = tid; e < PW * PW; e += kThreads) Ts[e / PW][e % PW] = 0.0;
  double a[ROWS][PW];
#pragma unroll
  for (int u = 0; u < ROWS; ++u) {
    const int r = kb + tid + kThreads * u;
#pragma unroll
    for (int c = 0; c < PW; ++c) a[u][c] = (r < m && c < nb) ? W[(int64_t)(kb + c) * mp + r] : 0.0;
  }
  // the steps are unrolled (j a constant in each), so the pivot column is a
  // register, not a select over all PW columns, and finished columns have no
  // code (a runtime j cost 239-258 us per 32-column panel)
  tn::static_for<PW>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j >= nb) return;
    const int rj = kb + j;
    double x[ROWS];
#pragma unroll
    for (int u = 0; u < ROWS; ++u) {
      const int r = kb + tid + kThreads * u;
      x[u] = (r >= rj) ? a[u][j] : 0.0;
    }
    // every column's x . a_c: the later ones for their update, the earlier
    // (the reflectors y_c) for the new column of T
    if (tid == j) {   // row rj is row 0 of thread j
#pragma unroll
      for (int c = 0; c < PW; ++c) row[c] = a[0][c];
    }
#pragma unroll
    for (int c = 0; c < PW; ++c) {
      double d = 0.0;
#pragma unroll
      for (int u = 0; u < ROWS; ++u) d = fma(x[u], a[u][c], d);
      part[c][tid] = d;
    }
    __syncthreads();
    {
      const int c = tid / TPC, s0 = (tid % TPC) * (kThreads / TPC);
      double acc = 0.0;
      for (int i = 0; i < kThreads / TPC; ++i) acc += part[c][s0 + i];
      part2[c][tid % TPC] = acc;
    }
    __syncthreads();
    if (tid < PW) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < TPC; ++i) acc += part2[tid][i];
      res[tid] = acc;
    }
    __syncthreads();
    const double xx = res[j], xj = row[j];
    const bool refl = xx > 0.0;
    const double alpha = refl ? -copysign(sqrt(xx), xj) : xj;
    const double tau = refl ? 1.0 / (alpha * (alpha - xj)) : 0.0;
    const double vj = xj - alpha;
#pragma unroll
    for (int c = j + 1; c < PW; ++c) {
      if (c >= nb) continue;
      const double coef = tau * (res[c] - alpha * row[c]);
#pragma unroll
      for (int u = 0; u < ROWS; ++u) {
        const int r = kb + tid + kThreads * u;
        if (r >= rj) a[u][c] -= coef * ((r == rj) ? vj : x[u]);
      }
    }
    // column j becomes the reflector (rows >= rj)
#pragma unroll
    for (int u = 0; u < ROWS; ++u) {
      const int r = kb + tid + kThreads * u;
      if (r == rj) a[u][j] = vj;
    }
    if (tid < j) {
      double sum = 0.0;
      for (int l = tid; l < j; ++l) sum += Ts[tid][l] * (res[l] - alpha * row[l]);
      Ts[tid][j] = -tau * sum;
    }
    if (tid == j) {
      Ts[j][j] = tau;
      rdiag[j] = alpha;
    }
    __syncthreads();
  });
  double* Y = scr + pr.y_off;
#pragma unroll
  for (int u = 0; u < ROWS; ++u) {
    const int r = kb + tid + kThreads * u;
    if (r >= pr.mp) continue;
    const int i = r - kb;
#pragma unroll
    for (int c = 0; c < PW; ++c) {
      if (c >= nb) continue;
      W[(int64_t)(kb + c) * mp + r] = (c > i) ? a[u][c] : (c == i ? rdiag[c] : 0.0);
      Y[(int64_t)c * mp + r] = (i >= c) ? a[u][c] : 0.0;
    }
  }
  for (int e = tid; e < PW * PW; e += kThreads) scr[pr.t_off + e] = Ts[e / PW][e % PW];
}

// Trailing update after panel `panel`: workgroup (tile, problem) applies
// Q_panel^T = I - Y T^T Y^T to the 32 columns c0 .. c0 + 32 of W (rows >= kb):
// Z = Y^T C (rows in 64-row chunks through LDS, a fixed order), Z2 = T^T Z,
// C -= Y Z2.
template <int PW>
__global__ void __launch_bounds__(kThreads) bj_qr_update(const Prob* __restrict__ probs, const int* __restrict__ qlist,
                                                         double* __restrict__ scr, int panel) {
  constexpr int ZQ = PW * 32 / kThreads;   // entries of Z (PW x 32) per thread
  __shared__ double Ys[64][PW + 1], Cs[64][33];
  __shared__ double Zs[PW][33], Z2[PW][33], Tt[PW][PW + 1];
  const Prob pr = probs[qlist[blockIdx.y]];
  if (prob_off(pr)) return;
  const int kb = panel * PW;
  const int c0 = kb + PW + 32 * (int)blockIdx.x;
  if (c0 >= pr.n) return;
  const int ncol = min(32, pr.np - c0);   // a 16-wide panel's tiles may end past W's padded columns
  const int m = pr.m, tid = threadIdx.x;
  const int64_t mp = pr.mp;
  double* W = scr + pr.w_off;
  const double* Y = scr + pr.y_off;
  for (int e = tid; e < PW * PW; e += kThreads) Tt[e / PW][e % PW] = scr[pr.t_off + e];
  const int zi = tid / (32 / ZQ), zc = (tid % (32 / ZQ)) * ZQ;
  double acc[ZQ];
#pragma unroll
  for (int q = 0; q < ZQ; ++q) acc[q] = 0.0;
  for (int r0 = kb; r0 < m; r0 += 64) {
    for (int e = tid; e < 64 * PW; e += kThreads) {
      const int c = e / 64, rr = e % 64, r = r0 + rr;
      Ys[rr][c] = (r < m) ? Y[(int64_t)c * mp + r] : 0.0;
    }
    for (int e = tid; e < 64 * 32; e += kThreads) {
      const int c = e / 64, rr = e % 64, r = r0 + rr;
      Cs[rr][c] = (r < m && c < ncol) ? W[(int64_t)(c0 + c) * mp + r] : 0.0;
    }
    __syncthreads();
    for (int rr = 0; rr < 64; ++rr) {
      const double y = Ys[rr][zi];
#pragma unroll
      for (int q = 0; q < ZQ; ++q) acc[q] = fma(y, Cs[rr][zc + q], acc[q]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < ZQ; ++q) Zs[zi][zc + q] = acc[q];
  __syncthreads();
  // Z2 = T^T Z (T upper triangular)
#pragma unroll
  for (int q = 0; q < ZQ; ++q) {
    double z = 0.0;
    for (int l = 0; l <= zi; ++l) z = fma(Tt[l][zi], Zs[l][zc + q], z);
    acc[q] = z;
  }
#pragma unroll
  for (int q = 0; q < ZQ; ++q) Z2[zi][zc + q] = acc[q];
  __syncthreads();
  // C -= Y Z2: thread (row rr, 8 columns)
  const int rr = tid & 63, cg = (tid >> 6) * 8;
  for (int r = kb + rr; r < m; r += 64) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = (cg + q < ncol) ? W[(int64_t)(c0 + cg + q) * mp + r] : 0.0;
#pragma unroll 4
    for (int i = 0; i < PW; ++i) {
      const double yi = Y[(int64_t)i * mp + r];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = fma(-yi, Z2[i][cg + q], v[q]);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (cg + q < ncol) W[(int64_t)(c0 + cg + q) * mp + r] = v[q];
  }
}

// X = R^T in place: R's strict upper triangle moves to the lower one.
__global__ void __launch_bounds__(kThreads) bj_qr_transpose(const Prob* __restrict__ probs,
                                                            const int* __restrict__ qlist, double* __restrict__ scr) {
  const Prob pr = probs[qlist[blockIdx.y]];
  if (prob_off(pr)) return;
  double* W = scr + pr.w_off;
  const int64_t n = pr.n, mp = pr.mp;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n * n; e += stride) {
    const int64_t i = e / n, k = e % n;
    if (k >= i) continue;
    const double v = W[i * mp + k];   // R[k][i]
    W[k * mp + i] = v;                // X[i][k]
    W[i * mp + k] = 0.0;
  }
}

// QR problems: U_X^T in rank order, row k = column of X / sigma (zero for
// sigma = 0, completed below).
__global__ void __launch_bounds__(kThreads) bj_qr_ut(const Prob* __restrict__ probs, const int* __restrict__ qlist,
                                                     double* __restrict__ scr, const int* __restrict__ iscr) {
  const Prob pr = probs[qlist[blockIdx.y]];
  if (prob_off(pr)) return;
  const double* W = scr + pr.w_off;
  const double* sig = scr + pr.sig_off;
  const int* rank = iscr + pr.rank_off;
  double* Ut = scr + pr.ut_off;
  const int64_t n = pr.n;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n * n; e += stride) {
    const int64_t k = e / n, i = e % n;
    const double sk = sqrt(sig[k]);
    Ut[rank[k] * n + i] = (sk > 0.0) ? W[k * pr.mp + i] / sk : 0.0;
  }
}

// Rows of U_X^T whose singular value is zero (an exactly rank-deficient A)
// are completed to an orthonormal basis: for each, the first unit vector e_i
// that keeps more than half its length after two Gram-Schmidt passes
// against the rows already orthonormal.  One workgroup per problem; returns
// at once when every singular value is positive.
__global__ void __launch_bounds__(kThreads) bj_qr_complete(const Prob* __restrict__ probs,
                                                           const int* __restrict__ qlist, double* __restrict__ scr,
                                                           const double* __restrict__ S) {
  __shared__ double u[kQrMaxM], d[kQrMaxM], red[kThreads];
  const Prob pr = probs[qlist[blockIdx.x]];
  if (prob_off(pr)) return;
  const int n = pr.n, tid = threadIdx.x;
  double* Ut = scr + pr.ut_off;
  int good = n;   // rows [0, good) are orthonormal (S is descending)
  while (good > 0 && !(S[pr.s_off + good - 1] > 0.0)) --good;
  for (int k = good; k < n; ++k) {
    for (int cand = 0; cand < n; ++cand) {
      for (int i = tid; i < n; i += kThreads) u[i] = (i == cand) ? 1.0 : 0.0;
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass) {
        for (int j = tid; j < k; j += kThreads) {
          double s = 0.0;
          for (int i = 0; i < n; ++i) s = fma(Ut[(int64_t)j * n + i], u[i], s);
          d[j] = s;
        }
        __syncthreads();
        for (int i = tid; i < n; i += kThreads) {
          double v = u[i];
          for (int j = 0; j < k; ++j) v = fma(-d[j], Ut[(int64_t)j * n + i], v);
          u[i] = v;
        }
        __syncthreads();
      }
      double s = 0.0;
      for (int i = tid; i < n; i += kThreads) s = fma(u[i], u[i], s);
      s = block_sum(s, red);
      if (s > 0.25) {
        const double inv = 1.0 / sqrt(s);
        for (int i = tid; i < n; i += kThreads) Ut[(int64_t)k * n + i] = u[i] * inv;
        __syncthreads();
        break;
      }
    }
  }
}

// The Newton-Schulz step's last part: VT = (3 VT - H) / 2, H = (VT VT^T) VT
// (in W's storage, leading dimension n).  U_X is orthogonal only to the
// sweeps' tolerance; this makes VT's rows orthonormal to rounding, so a warm
// start does not carry the error from one call to the next.
__global__ void __launch_bounds__(kThreads) bj_qr_polish(const Prob* __restrict__ probs, const int* __restrict__ qlist,
                                                         const double* __restrict__ scr, double* __restrict__ VT) {
  const Prob pr = probs[qlist[blockIdx.y]];
  if (prob_off(pr)) return;
  const double* H = scr + pr.w_off;
  const int64_t nn = (int64_t)pr.n * pr.n;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < nn; e += stride)
    VT[pr.vt_off + e] = 1.5 * VT[pr.vt_off + e] - 0.5 * H[e];
}

// Singular values (square roots of the exact column norms, descending; ties
// by column index), their ranks, and the convergence status.
__global__ void __launch_bounds__(kThreads) bj_sigma(const Prob* __restrict__ probs, double* __restrict__ scr,
                                                     int* __restrict__ iscr, double* __restrict__ S, int nsweeps,
                                                     int* __restrict__ status) {
  extern __shared__ double sh[];
  const Prob pr = probs[blockIdx.x];
  if (prob_off(pr)) return;
  const double* nrm = scr + pr.sig_off;
  int* rank = iscr + pr.rank_off;
  const int tid = threadIdx.x;
  for (int j = tid; j < pr.n; j += kThreads) sh[j] = nrm[j];
  __syncthreads();
  for (int j = tid; j < pr.n; j += kThreads) {
    const int rk = norm_rank(sh, pr.n, j);
    rank[j] = rk;
    S[pr.s_off + rk] = sqrt(sh[j]);
  }
  if (tid == 0) {
    int sw = -1;
    for (int s = 0; s < nsweeps; ++s)
      if (iscr[pr.flag_off + s] == 0) {
        sw = s + 1;
        break;
      }
    if (sw < 0) {
      atomicAdd(status, 1);
      sw = nsweeps;
    }
    atomicMax(status + 1, sw);
  }
}

// U and VT from the normalised columns of W and from V, in rank order.
__global__ void __launch_bounds__(kThreads) bj_vectors(const Prob* __restrict__ probs, const double* __restrict__ scr,
                                                       const int* __restrict__ iscr, double* __restrict__ U,
                                                       double* __restrict__ VT) {
  const Prob pr = probs[blockIdx.y];
  if (prob_off(pr)) return;
  if (pr.qr) return;   // VT by the U_X^T V0^T product
  const double* W = scr + pr.w_off;
  const double* V = scr + pr.v_off;
  const double* sig = scr + pr.sig_off;
  const int* rank = iscr + pr.rank_off;
  const int64_t m = pr.m, n = pr.n;
  const int64_t totU = m * n, totV = n * n;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < totU + totV; e += stride) {
    if (e < totU) {
      const int64_t j = e / m, i = e % m;
      const double sj = sqrt(sig[j]);
      const double v = (sj > 0.0) ? W[j * pr.mp + i] / sj : 0.0;
      const int rk = rank[j];
      if (!pr.tr) {
        if (pr.u_off >= 0) U[pr.u_off + i * n + rk] = v;   // U is M x n
      } else {
        VT[pr.vt_off + rk * pr.N + i] = v;   // VT is n x N
      }
    } else {
      const int64_t e2 = e - totU;
      const int64_t j = e2 / n, i = e2 % n;
      const int rk = rank[j];
      const double vv = V[j * pr.nvp + i];
      if (!pr.tr)
        VT[pr.vt_off + rk * n + i] = vv;     // VT is n x n
      else if (pr.u_off >= 0)
        U[pr.u_off + i * n + rk] = vv;       // U is M x M
    }
  }
}

}  // namespace

struct Plan {
  std::vector<Prob> probs;
  Prob* d_probs = nullptr;
  int2* d_map = nullptr;
  int64_t pairs = 0;
  int max_nb = 0;
  size_t lds = 0;                 // dynamic LDS of the sort: the largest n
  double* d_scr = nullptr;
  int* d_iscr = nullptr;
  tn_gemm_plan warm = nullptr;    // W = A V0 of the warm-started problems
  std::vector<int> gd_prob, qd_prob, qg_prob, qh_prob;   // problem of each group of warm / qr_vt / qr_g / qr_h
  int max_sw = kMaxSweeps;        // sweep cap (tn_svd_set_max_sweeps)
  int* d_qlist = nullptr;         // QR-preconditioned problems: 32-wide panels first, then 16-, then 8-wide
  int nqr = 0, nqr32 = 0, nqr16 = 0, max_n32 = 0, max_n16 = 0, max_n8 = 0;
  tn_gemm_plan qr_vt = nullptr;   // their VT = U_X^T V0^T
  tn_gemm_plan qr_g = nullptr;    // G = VT VT^T
  tn_gemm_plan qr_h = nullptr;    // H = G VT (then VT = (3 VT - H) / 2)
};

static int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

void destroy(Plan* P) {
  if (!P) return;
  if (P->warm) tn_gemm_plan_destroy(P->warm);
  if (P->qr_vt) tn_gemm_plan_destroy(P->qr_vt);
  if (P->qr_g) tn_gemm_plan_destroy(P->qr_g);
  if (P->qr_h) tn_gemm_plan_destroy(P->qr_h);
  tn::pool_free(P->d_qlist);
  tn::pool_free(P->d_probs);
  tn::pool_free(P->d_map);
  tn::pool_free(P->d_scr);
  tn::pool_free(P->d_iscr);
  delete P;
}

bool empty(const Plan* P) { return P == nullptr || P->probs.empty(); }

int set_mask(Plan* P, const int* mask, const std::vector<int64_t>& idx) {
  if (P == nullptr || P->probs.empty()) return TN_OK;
  TN_REQUIRE(idx.size() == P->probs.size(), TN_EINVAL, "block SVD mask: %zu indices for %zu problems", idx.size(),
             P->probs.size());
  for (size_t k = 0; k < P->probs.size(); ++k) {
    P->probs[k].mask = mask;
    P->probs[k].mask_idx = mask ? idx[k] : -1;
  }
  int rc = tn_memcpy_h2d(P->d_probs, P->probs.data(), sizeof(Prob) * P->probs.size(), nullptr);
  auto sub = [&](tn_gemm_plan g, const std::vector<int>& prob) {
    if (rc != TN_OK || g == nullptr) return;
    std::vector<int32_t> gi(prob.size());
    for (size_t k = 0; k < prob.size(); ++k) gi[k] = (int32_t)idx[prob[k]];
    rc = tn_gemm_plan_set_mask(g, mask, gi.data());
  };
  sub(P->warm, P->gd_prob);
  sub(P->qr_vt, P->qd_prob);
  sub(P->qr_g, P->qg_prob);
  sub(P->qr_h, P->qh_prob);
  return rc;
}

int build(Plan** out, const std::vector<Spec>& specs, int max_sweeps) {
  auto* P = new Plan();
  P->max_sw = max_sweeps > 0 ? std::min(max_sweeps, kMaxSweeps) : kMaxSweeps;
  int64_t scr = 0, iscr = kMaxSweeps;
  std::vector<int2> map;
  std::vector<tn_gemm_desc_ex> gd, qd, qg, qh;
  std::vector<int> qlist, qlist16, qlist8;
  for (size_t i = 0; i < specs.size(); ++i) {
    const Spec& s = specs[i];
    Prob p{};
    p.M = s.M;
    p.N = s.N;
    p.a_off = s.a_off;
    p.u_off = s.u_off;
    p.s_off = s.s_off;
    p.vt_off = s.vt_off;
    p.v0_off = s.v0_off;
    p.tr = s.M < s.N;
    const int64_t m = std::max(s.M, s.N), n = std::min(s.M, s.N);
    const int64_t np = round_up(n, kPW), mp = round_up(m, 16), nvp = round_up(n, 16);
    if (p.tr && s.v0_off >= 0) {
      destroy(P);
      tn::set_error("svd: a warm start needs M >= N");
      return TN_EINVAL;
    }
    if (mp * np >= (int64_t(1) << 31) || n > kMaxN) {
      destroy(P);
      tn::set_error("svd: %lld x %lld is too large", (long long)s.M, (long long)s.N);
      return TN_EUNSUPP;
    }
    p.m = (int32_t)m;
    p.n = (int32_t)n;
    p.mp = (int32_t)mp;
    p.nvp = (int32_t)nvp;
    p.np = (int32_t)np;
    p.nb = (int32_t)(np / kB);
    p.w_off = scr;
    scr += mp * np;
    p.v_off = scr;
    scr += nvp * np;
    p.sig_off = scr;
    scr += round_up(n + 1, 2);
    p.flag_off = (int32_t)iscr;
    iscr += kMaxSweeps;
    p.rank_off = (int32_t)iscr;
    iscr += n;
    p.pos_off = (int32_t)iscr;
    iscr += np;
    // only S and VT wanted: QR-preconditioned sweeps (W held by a panel
    // workgroup's registers up to kQrMaxM rows)
    p.qr = (s.u_off < 0 && !p.tr && m <= kQrMaxM) ? 1 : 0;
    p.qpw = (m <= kQrWideMaxM) ? kPW : (m <= kQrMidMaxM ? kPW / 2 : kPW / 4);
    p.ms = p.qr ? (int32_t)round_up(n, 16) : p.mp;
    p.y_off = p.t_off = p.ut_off = -1;
    if (p.qr) {
      p.y_off = scr;
      scr += mp * kPW;
      p.t_off = scr;
      scr += kPW * kPW;
      p.ut_off = scr;
      scr += n * n;
      (p.qpw == kPW ? qlist : (p.qpw == kPW / 2 ? qlist16 : qlist8)).push_back((int)i);
      if (p.qpw == kPW)
        P->max_n32 = std::max(P->max_n32, (int)n);
      else if (p.qpw == kPW / 2)
        P->max_n16 = std::max(P->max_n16, (int)n);
      else
        P->max_n8 = std::max(P->max_n8, (int)n);
      // VT(k, r) = sum_i Ut(k, i) V0T(i, r), V0T kept in V (row stride nvp)
      tn_gemm_desc_ex g{};
      g.M = n;
      g.N = n;
      g.K = n;
      g.a_off = p.ut_off;
      g.lda = n;
      g.b_off = p.v_off;
      g.ldb = nvp;
      g.c_off = s.vt_off;
      g.ldc = n;
      qd.push_back(g);
      P->qd_prob.push_back((int)i);
      // Newton-Schulz step on VT: G = VT VT^T into Ut's storage, H = G VT
      // into W's (both free by then)
      tn_gemm_desc_ex h{};
      h.M = n;
      h.N = n;
      h.K = n;
      h.a_off = s.vt_off;
      h.lda = n;
      h.b_off = s.vt_off;
      h.b_row = tn_index_map{1, 0, {n, 0, 0, 0}, {1, 0, 0, 0}};
      h.b_col = tn_index_map{1, 0, {n, 0, 0, 0}, {n, 0, 0, 0}};
      h.c_off = p.ut_off;
      h.ldc = n;
      qg.push_back(h);
      P->qg_prob.push_back((int)i);
      tn_gemm_desc_ex f{};
      f.M = n;
      f.N = n;
      f.K = n;
      f.a_off = p.ut_off;
      f.lda = n;
      f.b_off = s.vt_off;
      f.ldb = n;
      f.c_off = p.w_off;
      f.ldc = n;
      qh.push_back(f);
      P->qh_prob.push_back((int)i);
    }
    for (int k = 0; k < p.nb / 2; ++k) map.push_back(make_int2((int)i, k));
    P->max_nb = std::max(P->max_nb, p.nb);
    P->lds = std::max(P->lds, sizeof(double) * (size_t)n);
    if (s.v0_off >= 0) {
      // W(i, j) = sum_k A(i, k) V0T(j, k), W column-major with leading dimension mp
      tn_gemm_desc_ex g{};
      g.M = m;
      g.N = n;
      g.K = n;
      g.a_off = s.a_off;
      g.lda = s.N;
      g.b_off = s.v0_off;
      g.c_off = p.w_off;
      g.b_row = tn_index_map{1, 0, {n, 0, 0, 0}, {1, 0, 0, 0}};
      g.b_col = tn_index_map{1, 0, {n, 0, 0, 0}, {n, 0, 0, 0}};
      g.c_row = tn_index_map{1, 0, {m, 0, 0, 0}, {1, 0, 0, 0}};
      g.c_col = tn_index_map{1, 0, {n, 0, 0, 0}, {mp, 0, 0, 0}};
      g.ldb = n;
      g.ldc = mp;
      gd.push_back(g);
      P->gd_prob.push_back((int)i);
    }
    P->probs.push_back(p);
  }
  P->pairs = (int64_t)map.size();
  if (P->probs.empty()) {
    *out = P;
    return TN_OK;
  }
  int rc = tn::pool_alloc((void**)&P->d_probs, sizeof(Prob) * P->probs.size());
  if (rc == TN_OK) rc = tn::pool_alloc((void**)&P->d_map, sizeof(int2) * map.size());
  if (rc == TN_OK) rc = tn::pool_alloc((void**)&P->d_scr, sizeof(double) * scr);
  if (rc == TN_OK) rc = tn::pool_alloc((void**)&P->d_iscr, sizeof(int) * iscr);
  if (rc == TN_OK) rc = tn_memcpy_h2d(P->d_probs, P->probs.data(), sizeof(Prob) * P->probs.size(), nullptr);
  if (rc == TN_OK) rc = tn_memcpy_h2d(P->d_map, map.data(), sizeof(int2) * map.size(), nullptr);
  if (rc == TN_OK && !gd.empty()) rc = tn_gemm_plan_create_ex(&P->warm, TN_GEMM_F64, (int64_t)gd.size(), gd.data());
  P->nqr32 = (int)qlist.size();
  P->nqr16 = (int)qlist16.size();
  qlist.insert(qlist.end(), qlist16.begin(), qlist16.end());
  qlist.insert(qlist.end(), qlist8.begin(), qlist8.end());
  P->nqr = (int)qlist.size();
  if (rc == TN_OK && P->nqr > 0) rc = tn::pool_alloc((void**)&P->d_qlist, sizeof(int) * qlist.size());
  if (rc == TN_OK && P->nqr > 0) rc = tn_memcpy_h2d(P->d_qlist, qlist.data(), sizeof(int) * qlist.size(), nullptr);
  if (rc == TN_OK && P->nqr > 0) rc = tn_gemm_plan_create_ex(&P->qr_vt, TN_GEMM_F64, (int64_t)qd.size(), qd.data());
  if (rc == TN_OK && P->nqr > 0) rc = tn_gemm_plan_create_ex(&P->qr_g, TN_GEMM_F64, (int64_t)qg.size(), qg.data());
  if (rc == TN_OK && P->nqr > 0) rc = tn_gemm_plan_create_ex(&P->qr_h, TN_GEMM_F64, (int64_t)qh.size(), qh.data());
  if (rc != TN_OK) {
    destroy(P);
    return rc;
  }
  *out = P;
  return TN_OK;
}

int launch(Plan* P, const double* A, const double* V0T, double* U, double* S, double* VT, int* status,
           hipStream_t st) {
  if (empty(P)) return TN_OK;
  const unsigned np = (unsigned)P->probs.size();
  hipLaunchKernelGGL(bj_setup, dim3(kSetupBlocks, np), dim3(kThreads), 0, st, P->d_probs, A, V0T, P->d_scr,
                     P->d_iscr);
  TN_LAUNCH_CHECK();
  if (P->warm) {
    const int rc = tn_gemm_exec(P->warm, A, V0T, P->d_scr, st);
    if (rc != TN_OK) return rc;
  }
  if (P->nqr > 0) {
    const unsigned nq = (unsigned)P->nqr;
    // panels of the 32-wide problems (list head), then of the 16- and 8-wide ones
    const int n32 = P->nqr32, n16 = P->nqr16, n8 = P->nqr - P->nqr32 - P->nqr16;
    for (int p = 0; n32 > 0 && p < (P->max_n32 + kPW - 1) / kPW; ++p) {
      hipLaunchKernelGGL((bj_qr_panel<kPW, 4>), dim3((unsigned)n32), dim3(kThreads), 0, st, P->d_probs, P->d_qlist,
                         P->d_scr, p);
      TN_LAUNCH_CHECK();
      const int tiles = (P->max_n32 - kPW * (p + 1) + 31) / 32;
      if (tiles > 0) {
        hipLaunchKernelGGL(bj_qr_update<kPW>, dim3((unsigned)tiles, (unsigned)n32), dim3(kThreads), 0, st, P->d_probs,
                           P->d_qlist, P->d_scr, p);
        TN_LAUNCH_CHECK();
      }
    }
    constexpr int kPN = kPW / 2;
    for (int p = 0; n16 > 0 && p < (P->max_n16 + kPN - 1) / kPN; ++p) {
      hipLaunchKernelGGL((bj_qr_panel<kPN, 8>), dim3((unsigned)n16), dim3(kThreads), 0, st, P->d_probs,
                         P->d_qlist + n32, P->d_scr, p);
      TN_LAUNCH_CHECK();
      const int tiles = (P->max_n16 - kPN * (p + 1) + 31) / 32;
      if (tiles > 0) {
        hipLaunchKernelGGL(bj_qr_update<kPN>, dim3((unsigned)tiles, (unsigned)n16), dim3(kThreads), 0, st, P->d_probs,
                           P->d_qlist + n32, P->d_scr, p);
        TN_LAUNCH_CHECK();
      }
    }
    constexpr int kP8 = kPW / 4;
    for (int p = 0; n8 > 0 && p < (P->max_n8 + kP8 - 1) / kP8; ++p) {
      hipLaunchKernelGGL((bj_qr_panel<kP8, 16>), dim3((unsigned)n8), dim3(kThreads), 0, st, P->d_probs,
                         P->d_qlist + n32 + n16, P->d_scr, p);
      TN_LAUNCH_CHECK();
      const int tiles = (P->max_n8 - kP8 * (p + 1) + 31) / 32;
      if (tiles > 0) {
        hipLaunchKernelGGL(bj_qr_update<kP8>, dim3((unsigned)tiles, (unsigned)n8), dim3(kThreads), 0, st, P->d_probs,
                           P->d_qlist + n32 + n16, P->d_scr, p);
        TN_LAUNCH_CHECK();
      }
    }
    hipLaunchKernelGGL(bj_qr_transpose, dim3(kSetupBlocks, nq), dim3(kThreads), 0, st, P->d_probs, P->d_qlist,
                       P->d_scr);
    TN_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(bj_colnorm, dim3(kNormBlocks, np), dim3(kThreads), 0, st, P->d_probs, P->d_scr);
  TN_LAUNCH_CHECK();
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  TN_HIP_CHECK(hipStreamIsCapturing(st, &cap));
  const bool capturing = cap != hipStreamCaptureStatusNone;
  // Under capture every sweep is launched (a converged problem's workgroups
  // return at once); otherwise the host stops after a sweep that rotated
  // nothing in any problem.
  const int max_sw = std::min(capturing ? kMaxSweepsCaptured : kMaxSweeps, P->max_sw);
  int ran = 0;
  for (int s = 0; s < max_sw; ++s) {
    hipLaunchKernelGGL(bj_sort, dim3(np), dim3(kThreads), P->lds, st, P->d_probs, P->d_scr, P->d_iscr, s);
    TN_LAUNCH_CHECK();
    for (int r = 0; r < P->max_nb - 1; ++r) {
      hipLaunchKernelGGL(bj_round, dim3((unsigned)P->pairs), dim3(kThreads), 0, st, P->d_probs, P->d_map, P->d_scr,
                         P->d_iscr, s, r);
      TN_LAUNCH_CHECK();
    }
    ran = s + 1;
    if (!capturing && s >= 1) {
      int any = 1;
      TN_HIP_CHECK(hipMemcpyAsync(&any, P->d_iscr + s, sizeof(int), hipMemcpyDeviceToHost, st));
      TN_HIP_CHECK(hipStreamSynchronize(st));
      if (!any) break;
    }
  }
  hipLaunchKernelGGL(bj_colnorm, dim3(kNormBlocks, np), dim3(kThreads), 0, st, P->d_probs, P->d_scr);
  TN_LAUNCH_CHECK();
  hipLaunchKernelGGL(bj_sigma, dim3(np), dim3(kThreads), P->lds, st, P->d_probs, P->d_scr, P->d_iscr, S, ran,
                     status);
  TN_LAUNCH_CHECK();
  hipLaunchKernelGGL(bj_vectors, dim3(kSetupBlocks, np), dim3(kThreads), 0, st, P->d_probs, P->d_scr, P->d_iscr, U,
                     VT);
  TN_LAUNCH_CHECK();
  if (P->nqr > 0) {
    const unsigned nq = (unsigned)P->nqr;
    hipLaunchKernelGGL(bj_qr_ut, dim3(kSetupBlocks, nq), dim3(kThreads), 0, st, P->d_probs, P->d_qlist, P->d_scr,
                       P->d_iscr);
    TN_LAUNCH_CHECK();
    hipLaunchKernelGGL(bj_qr_complete, dim3(nq), dim3(kThreads), 0, st, P->d_probs, P->d_qlist, P->d_scr, S);
    TN_LAUNCH_CHECK();
    int rc = tn_gemm_exec(P->qr_vt, P->d_scr, P->d_scr, VT, st);
    if (rc == TN_OK) rc = tn_gemm_exec(P->qr_g, VT, VT, P->d_scr, st);
    if (rc == TN_OK) rc = tn_gemm_exec(P->qr_h, P->d_scr, VT, P->d_scr, st);
    if (rc != TN_OK) return rc;
    hipLaunchKernelGGL(bj_qr_polish, dim3(kSetupBlocks, nq), dim3(kThreads), 0, st, P->d_probs, P->d_qlist, P->d_scr,
                       VT);
    TN_LAUNCH_CHECK();
  }
  return TN_OK;
}

}  // namespace bj
}  // namespace tn
