// Grouped sector GEMM on gfx950 matrix cores.
//
// Replaces the per-Qnum loop of contract (UniTensorTools.cpp:191-203), where
// every symmetry sector was one serial matrixMul -> cublasDgemm with a fresh
// handle (uni10_lapack_gpu.cu:42-152).  Here every sector of a contraction
// (and of many contractions) is cut into output tiles; one launch runs all
// tiles of all sectors.  Tiles are laid out so that neighbouring tiles of a
// sector land on the same XCD (blockIdx -> XCD is round-robin over 8 XCDs),
// which keeps their shared A rows / B columns in that XCD's L2.
//
// Operand addressing: every element address is (row offset) + (column
// offset), each from an index map (a leading dimension, or up to four
// (extent, stride) levels decoded with 32-bit magic-number division), so a
// contraction reads tensors in their stored bond order and writes its result
// in any order with no separate permute pass and no memory indirection.
// Each thread keeps the offsets of its fixed rows/columns in registers and
// decodes the K-side offsets once per K step.
//
// Arithmetic modes (storage is always FP64, row-major):
//   F64     v_mfma_f64_16x16x4f64, FP64 accumulate            (plain drop-in)
//   F32     v_mfma_f32_16x16x4f32 on inputs rounded to FP32, FP32 accumulate
//   BF16X3  inputs split x = h + m (+ rest), each BF16; products
//           hh + hm + mh on v_mfma_f32_16x16x32_bf16, FP32 accumulate
//   BF16X6  x = h + m + l; products hh+hm+mh+hl+mm+lh, FP32 accumulate
// Every output element is computed by exactly one workgroup, in a K order
// fixed per arithmetic mode (not per tile shape): results do not depend on
// the tile shape or on what else is in the group (no split-K, no atomics).
//
// MFMA operand layout used (16x16 output tile, wave of 64 lanes):
//   16x16x4 (f64/f32): lane l supplies A[l%16][l/16] and B[l/16][l%16];
//     accumulator element i of lane l is C[(l/16)+4i][l%16] for f64 and
//     C[4*(l/16)+i][l%16] for f32 (measured with a one-instruction probe).
//   16x16x32 bf16: lane l supplies A[l%16][8*(l/16)+i] and B[8*(l/16)+i][l%16]
//     for i = 0..7 (any k order works as long as A and B use the same one);
//     accumulator layout as f32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <climits>
#include <cstring>
#include <vector>

#include "tn_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct TileDev {
  int32_t group;   // -1: empty slot (padding for the XCD remap)
  int32_t tm, tn;
  int32_t pad;
};

// Index map on the device: always TN_MAX_MAP_LEVELS levels, the real ones
// right-aligned (last = fastest) and the rest of extent 1 and stride 0, so
// decoding is branch-free; the plain layout is the single level
// (range, leading dimension or 1).  Each extent carries its 32-bit division
// magic: q = (umulhi(i, magic) + i) >> shift, exact for i < 2^31.
struct MapDev {
  uint32_t ext[TN_MAX_MAP_LEVELS];
  uint32_t magic[TN_MAX_MAP_LEVELS];
  uint32_t shift[TN_MAX_MAP_LEVELS];
  int64_t stride[TN_MAX_MAP_LEVELS];
};

struct GroupDev {
  int64_t M, N, K;
  int64_t a_off, b_off, c_off;
  int64_t lda, ldb, ldc;
  MapDev a_row, a_col, b_row, b_col, c_row, c_col;
  int64_t ea_off, eb_off;     // exact mode: offsets of the A-row / B-column scale exponents
  const int* mask;            // with mask_idx >= 0: the group is skipped while mask[mask_idx] == 0
  int64_t mask_idx;
  int64_t d_off, rs_off;      // with rs_div > 0: D(m,n) = C(m,n) * S[rs_off + m / rs_div] at d_off
  int64_t rs_div;
  int64_t sa_off, sb_off, kp; // exact mode: the group's A / B slices in the plan's slice buffer, padded K
};

// Write-back of one output element: C, and the row-scaled copy when the
// group asks for it (the stored value times the scale, one rounding).
__device__ __forceinline__ void put_c(double* __restrict__ C, const GroupDev& g, int64_t off, int64_t m, double v) {
  C[g.c_off + off] = v;
  if (g.rs_div > 0) C[g.d_off + off] = v * C[g.rs_off + m / g.rs_div];
}

__device__ __forceinline__ bool group_off(const GroupDev& g) {
  return g.mask != nullptr && g.mask_idx >= 0 && g.mask[g.mask_idx] == 0;
}

__device__ __forceinline__ int64_t map_off(const MapDev& m, int64_t i64) {
  uint32_t i = (uint32_t)i64;
  int64_t off = 0;
#pragma unroll
  for (int l = TN_MAX_MAP_LEVELS - 1; l >= 0; --l) {
    const uint32_t q = (__umulhi(i, m.magic[l]) + i) >> m.shift[l];
    off += (int64_t)(i - q * m.ext[l]) * m.stride[l];
    i = q;
  }
  return off;
}

// ----------------------------------------------------------------------------
// FP64 / FP32 kernel.  256 threads = 2x2 waves; the block owns a BM x BN
// output tile, each wave a (BM/2) x (BN/2) sub-tile of TM x TN MFMA tiles.
//
// K order (the same for every tile shape, so results never depend on it):
// K runs in 32-wide chunks; in MFMA step s (0..7) of chunk c, lane group
// g = lane/16 supplies k = 32c + 8g + s, and step s accumulates into chain
// s mod 2; C = chain0 + chain1 at the end.  Two independent accumulation
// chains keep the matrix core busy when a wave owns a single 16x16 tile.
//
// Pipeline: a stage is KC (a multiple of 32) consecutive k.  Global loads
// of stage t+D are in flight (registers, D slots) while stage t is
// multiplied out of LDS; LDS is double-buffered with one barrier per stage,
// so a deeper stage amortises the barrier and the LDS read latency over
// more MFMAs.  Each thread loads element pairs along the contiguous axis:
// A pairs (k, k+1), B pairs (n, n+1); VEC = one 16-byte load per pair (the
// host checks that every pair is contiguous and 16-byte aligned).
constexpr int kKTab = 2048;   // longest K whose offsets are tabulated in LDS

// NT = 512: two waves per 16x16 sub-tile, one per K chain (steps s even /
// odd of every chunk), so each SIMD holds two waves; the chains are summed
// in the same order as with one wave holding both.
//
// PK: every group reads A's columns and B's rows in the plain layout
// (offset k and k * ldb), so the K-side offsets are one multiply and the
// block builds no LDS tables (less LDS, no table pass before the first
// stage).
template <typename T, int BM, int BN, int KC, int D, bool VEC, bool KT, int NT = 256, bool PK = false>
__global__ void __launch_bounds__(NT) gemm_mfma_k32(const double* __restrict__ A,
                                                     const double* __restrict__ B,
                                                     double* __restrict__ C,
                                                     const GroupDev* __restrict__ groups,
                                                     const TileDev* __restrict__ tiles) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int NPA = KC / 2;                 // A pairs per row of a stage
  constexpr int AROWS = NT / NPA;             // A rows per pass of the block
  constexpr int PA = BM / AROWS;              // A pairs per thread per stage
  constexpr int NPB = BN / 2;                 // B pairs per k row
  constexpr int BROWS = NT / NPB;             // k rows per pass of the block
  constexpr int PB = KC / BROWS;              // B pairs per thread per stage
  constexpr int PAD = 16 / (int)sizeof(T);    // keeps rows 16-byte aligned
  static_assert(KC % 32 == 0 && TM >= 1 && TN >= 1 && NT % NPA == 0 && BM % AROWS == 0 &&
                    NT % NPB == 0 && KC % BROWS == 0 && (NT == 256 || NT == 512),
                "tile shape");
  constexpr int NCH = (NT == 512) ? 1 : 2;    // K chains held by one wave
  typedef typename std::conditional<sizeof(T) == 8, d4, f4>::type acc_t;
  typedef T t2 __attribute__((ext_vector_type(2)));
  __shared__ __attribute__((aligned(16))) T As[2][BM][KC + PAD];
  __shared__ __attribute__((aligned(16))) T Bs[2][KC][BN + PAD];
  constexpr bool TAB = KT && !PK;
  __shared__ int32_t tabA[TAB ? kKTab : 1], tabB[TAB ? kKTab : 1];

  const TileDev tile = tiles[blockIdx.x];
  if (tile.group < 0) return;
  const GroupDev& g = groups[tile.group];
  if (group_off(g)) return;
  const int64_t M = g.M, N = g.N, K = g.K;
  const int64_t m0 = (int64_t)tile.tm * BM, n0 = (int64_t)tile.tn * BN;
  const double* __restrict__ Ab = A + g.a_off;
  const double* __restrict__ Bb = B + g.b_off;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lg = lane >> 4, lr = lane & 15;
  const int tw = wave & 3;                    // the wave's 16x16 sub-tiles
  const int ch = (NT == 512) ? (wave >> 2) : 0;   // its K chain when NCH == 1
  const int wm0 = (tw >> 1) * WM, wn0 = (tw & 1) * WN;

  // Out-of-range rows, columns and k are clamped to the last valid index for
  // addressing (every load reads valid memory, with no branches) and their
  // values are zeroed when the stage is written to LDS.
  const int a_kp = tid % NPA;                 // pair (2*a_kp, 2*a_kp+1) of the stage
  const int a_r0 = tid / NPA;                 // rows a_r0 + AROWS*i
  int64_t a_roff[PA];
  bool a_rv[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) {
    const int64_t m = m0 + a_r0 + AROWS * i;
    a_rv[i] = m < M;
    a_roff[i] = map_off(g.a_row, a_rv[i] ? m : M - 1);
  }
  const int b_np = tid % NPB;
  const int b_kr0 = tid / NPB;                // k rows b_kr0 + BROWS*i
  const int64_t bn = n0 + 2 * b_np;
  const bool b_ok0 = bn < N, b_ok1 = bn + 1 < N;
  const int64_t b_coff0 = map_off(g.b_col, std::min<int64_t>(bn, VEC ? N - 2 : N - 1));
  const int64_t b_coff1 = VEC ? 0 : map_off(g.b_col, std::min<int64_t>(bn + 1, N - 1));
  const int64_t a_kmax = VEC ? K - 2 : K - 1;

  // K-side offsets: with KT (K <= kKTab, offsets below 2^31) the block
  // tabulates a_col(k) and b_row(k) in LDS once, so a stage costs one LDS
  // read per offset instead of a multi-level decode.
  const MapDev a_col = g.a_col, b_row = g.b_row;
  const int64_t a_ks = a_col.stride[TN_MAX_MAP_LEVELS - 1], b_ks = b_row.stride[TN_MAX_MAP_LEVELS - 1];
  if constexpr (TAB) {
    for (int k = tid; k < K; k += NT) {
      tabA[k] = (int32_t)map_off(a_col, k);
      tabB[k] = (int32_t)map_off(b_row, k);
    }
    __syncthreads();
  }
  auto koff_a = [&](int64_t k) -> int64_t { return PK ? k * a_ks : (TAB ? (int64_t)tabA[k] : map_off(a_col, k)); };
  auto koff_b = [&](int64_t k) -> int64_t { return PK ? k * b_ks : (TAB ? (int64_t)tabB[k] : map_off(b_row, k)); };

  acc_t acc[NCH][TM][TN];
#pragma unroll
  for (int h = 0; h < NCH; ++h)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[h][i][j] = acc_t{0, 0, 0, 0};

  // Every stage issues the same loads, so the waits count them exactly; the
  // zeroing happens at the LDS write, so no wait follows a load.
  double ra[D][PA][2], rb[D][PB][2];
  auto load = [&](int d, int64_t t) {
    const int64_t ka = t * KC + 2 * a_kp;
    const int64_t ao0 = koff_a(std::min<int64_t>(ka, a_kmax));
    if constexpr (VEC) {
#pragma unroll
      for (int i = 0; i < PA; ++i) {
        const double2 v = *reinterpret_cast<const double2*>(Ab + a_roff[i] + ao0);
        ra[d][i][0] = v.x;
        ra[d][i][1] = v.y;
      }
    } else {
      const int64_t ao1 = koff_a(std::min<int64_t>(ka + 1, K - 1));
#pragma unroll
      for (int i = 0; i < PA; ++i) {
        ra[d][i][0] = Ab[a_roff[i] + ao0];
        ra[d][i][1] = Ab[a_roff[i] + ao1];
      }
    }
#pragma unroll
    for (int i = 0; i < PB; ++i) {
      const int64_t kb = t * KC + b_kr0 + BROWS * i;
      const int64_t bo = koff_b(std::min<int64_t>(kb, K - 1));
      if constexpr (VEC) {
        const double2 v = *reinterpret_cast<const double2*>(Bb + bo + b_coff0);
        rb[d][i][0] = v.x;
        rb[d][i][1] = v.y;
      } else {
        rb[d][i][0] = Bb[bo + b_coff0];
        rb[d][i][1] = Bb[bo + b_coff1];
      }
    }
  };
  // write register slot d (holding stage t) to LDS buffer st, zeroing the
  // elements outside the matrix
  auto store = [&](int d, int st, int64_t t) {
    const int64_t ka = t * KC + 2 * a_kp;
    const bool kv0 = ka < K, kv1 = ka + 1 < K;
#pragma unroll
    for (int i = 0; i < PA; ++i)
      *reinterpret_cast<t2*>(&As[st][a_r0 + AROWS * i][2 * a_kp]) =
          t2{(T)(a_rv[i] && kv0 ? ra[d][i][0] : 0.0), (T)(a_rv[i] && kv1 ? ra[d][i][1] : 0.0)};
#pragma unroll
    for (int i = 0; i < PB; ++i) {
      const bool kv = t * KC + b_kr0 + BROWS * i < K;
      *reinterpret_cast<t2*>(&Bs[st][b_kr0 + BROWS * i][2 * b_np]) =
          t2{(T)(kv && b_ok0 ? rb[d][i][0] : 0.0), (T)(kv && b_ok1 ? rb[d][i][1] : 0.0)};
    }
  };
  // Operands of one 32-k chunk: A (TM x 8 values) and B (TN x 8 values) per
  // lane.  The next chunk's operands are read from LDS into separate
  // registers before the current chunk's MFMAs issue, so LDS latency hides
  // behind the matrix core and no read waits on an MFMA's source registers.
  // (step s = 2q + h of a chunk accumulates into chain h)
  struct Frag {
    T a[4][NCH][TM];
    T b[4][NCH][TN];
  };
  auto read_frag = [&](Frag& f, int st, int u) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if constexpr (NCH == 2) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const t2 v = *reinterpret_cast<const t2*>(&As[st][wm0 + 16 * i + lr][32 * u + 8 * lg + 2 * q]);
          f.a[q][0][i] = v.x;
          f.a[q][1][i] = v.y;
        }
      } else {
#pragma unroll
        for (int i = 0; i < TM; ++i) f.a[q][0][i] = As[st][wm0 + 16 * i + lr][32 * u + 8 * lg + 2 * q + ch];
      }
#pragma unroll
      for (int h = 0; h < NCH; ++h)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          f.b[q][h][j] = Bs[st][32 * u + 8 * lg + 2 * q + (NCH == 2 ? h : ch)][wn0 + 16 * j + lr];
    }
  };
  auto mfma_frag = [&](const Frag& f) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int h = 0; h < NCH; ++h)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            if constexpr (sizeof(T) == 8)
              acc[h][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(f.a[q][h][i], f.b[q][h][j], acc[h][i][j], 0, 0, 0);
            else
              acc[h][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(f.a[q][h][i], f.b[q][h][j], acc[h][i][j], 0, 0, 0);
          }
  };
  // (sched_barrier keeps the scheduler from sinking the reads back next to
  // their MFMAs, which would serialise every MFMA behind an LDS round trip.)
  auto compute = [&](int st) {
    Frag f[2];
    read_frag(f[0], st, 0);
#pragma unroll
    for (int u = 0; u < KC / 32; ++u) {
      if (u + 1 < KC / 32) read_frag(f[(u + 1) & 1], st, u + 1);
      __builtin_amdgcn_sched_barrier(0);
      mfma_frag(f[u & 1]);
      __builtin_amdgcn_sched_barrier(0);
    }
  };

  const int64_t ns = (K + KC - 1) / KC;
#pragma unroll
  for (int d = 0; d < D; ++d) load(d, d);
  store(0, 0, 0);
  load(0, D);
  __syncthreads();
  for (int64_t tb = 0; tb < ns; tb += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int64_t t = tb + d;
      // Every iteration before the last multiply stores stage t+1 to the
      // other LDS buffer (k past K zeroed) and loads stage t+1+D (clamped
      // addresses past the end) into the freed register slot, so those
      // iterations issue the same memory operations and the waits count
      // them exactly; from the last multiply on, the iterations of the
      // D-group only skip (the loads still in flight are never waited for).
      // One exit (the loop condition) keeps the accumulators in the same
      // registers on every path (an exit after the multiply of each slot
      // made the compiler copy them out of the accumulation registers).
      if (t < ns) compute((int)(t & 1));
      if (t + 1 < ns) {
        store((d + 1) % D, (int)((t + 1) & 1), t + 1);
        load((d + 1) % D, t + 1 + D);
        __syncthreads();
      }
    }
  }

  // write back: C = chain0 + chain1.  With NCH == 1 the chain-1 waves hand
  // their accumulators to the chain-0 waves through LDS (the stage buffers
  // are free once every wave has passed the barrier below, i.e. finished
  // its last multiply).
  acc_t other[TM][TN];
  if constexpr (NCH == 1) {
    __syncthreads();
    acc_t* xch = reinterpret_cast<acc_t*>(&As[0][0][0]);
    static_assert(sizeof(acc_t) * 4 * 64 * TM * TN <= sizeof(As), "exchange fits the A stages");
    if (ch == 1) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) xch[((tw * TM + i) * TN + j) * 64 + lane] = acc[0][i][j];
    }
    __syncthreads();
    if (ch == 1) return;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) other[i][j] = xch[((tw * TM + i) * TN + j) * 64 + lane];
  } else {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) other[i][j] = acc[NCH - 1][i][j];
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gn = n0 + wn0 + j * 16 + lr;
    if (gn >= N) continue;
    const int64_t coff = map_off(g.c_col, gn);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // accumulator row: f64 16x16x4 -> (l/16) + 4v ; f32 16x16x4 -> 4(l/16) + v
        const int row = (sizeof(T) == 8) ? (lg + 4 * v) : (4 * lg + v);
        const int64_t gm = m0 + wm0 + i * 16 + row;
        if (gm < M) put_c(C, g, map_off(g.c_row, gm) + coff, gm, (double)acc[0][i][j][v] + (double)other[i][j][v]);
      }
  }
}

// ----------------------------------------------------------------------------
// Split-BF16 kernel: BM x BN tile, BK deep (multiple of 32), pieces NP = 2
// (BF16X3) or 3 (BF16X6).
__device__ inline void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  float r = x - (float)h;
  m = (__bf16)r;
  float r2 = r - (float)m;
  l = (__bf16)r2;
}

template <int BM, int BN, int NP, int BK>
__global__ void __launch_bounds__(256) gemm_split_bf16(const double* __restrict__ A,
                                                       const double* __restrict__ B,
                                                       double* __restrict__ C,
                                                       const GroupDev* __restrict__ groups,
                                                       const TileDev* __restrict__ tiles) {
  static_assert(BK % 32 == 0 && 256 % BK == 0 && 256 % BN == 0, "tile shape");
  constexpr int KP = BK + 8;                       // padded row (bf16 elements)
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int A_PER = BM * BK / 256;
  constexpr int B_PER = BN * BK / 256;
  constexpr int A_RSTEP = 256 / BK;
  constexpr int B_KSTEP = 256 / BN;
  __shared__ __attribute__((aligned(16))) __bf16 As[NP][BM][KP];   // [piece][m][k]
  __shared__ __attribute__((aligned(16))) __bf16 Bs[NP][BN][KP];   // [piece][n][k]

  const TileDev tile = tiles[blockIdx.x];
  if (tile.group < 0) return;
  const GroupDev& g = groups[tile.group];
  if (group_off(g)) return;
  const int64_t M = g.M, N = g.N, K = g.K;
  const int64_t m0 = (int64_t)tile.tm * BM, n0 = (int64_t)tile.tn * BN;
  const double* __restrict__ Ab = A + g.a_off;
  const double* __restrict__ Bb = B + g.b_off;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;

  const int a_kk = tid % BK;
  const int a_mm0 = tid / BK;
  int64_t a_roff[A_PER];
#pragma unroll
  for (int i = 0; i < A_PER; ++i) {
    const int64_t m = m0 + a_mm0 + i * A_RSTEP;
    a_roff[i] = (m < M) ? map_off(g.a_row, m) : -1;
  }
  const int b_nn = tid % BN;
  const int b_kk0 = tid / BN;
  const int64_t b_coff = (n0 + b_nn < N) ? map_off(g.b_col, n0 + b_nn) : -1;

  f4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0, 0, 0, 0};

  float ra[A_PER], rb[B_PER];
  auto load_tiles = [&](int64_t k0) {
    const int64_t ka = k0 + a_kk;
    const int64_t a_coff = (ka < K) ? map_off(g.a_col, ka) : -1;
#pragma unroll
    for (int i = 0; i < A_PER; ++i)
      ra[i] = (a_roff[i] >= 0 && a_coff >= 0) ? (float)Ab[a_roff[i] + a_coff] : 0.0f;
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int64_t kb = k0 + b_kk0 + i * B_KSTEP;
      const int64_t b_roff = (kb < K) ? map_off(g.b_row, kb) : -1;
      rb[i] = (b_roff >= 0 && b_coff >= 0) ? (float)Bb[b_roff + b_coff] : 0.0f;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      __bf16 h, m, l;
      split3(ra[i], h, m, l);
      const int mm = a_mm0 + i * A_RSTEP;
      As[0][mm][a_kk] = h;
      As[1][mm][a_kk] = m;
      if constexpr (NP > 2) As[2][mm][a_kk] = l;
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      __bf16 h, m, l;
      split3(rb[i], h, m, l);
      const int kk = b_kk0 + i * B_KSTEP;
      Bs[0][b_nn][kk] = h;
      Bs[1][b_nn][kk] = m;
      if constexpr (NP > 2) Bs[2][b_nn][kk] = l;
    }
  };

  const int64_t nk = (K + BK - 1) / BK;
  load_tiles(0);
  for (int64_t kt = 0; kt < nk; ++kt) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (kt + 1 < nk) load_tiles((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      const int kq = ks + 8 * (lane >> 4);
      bf16x8 a[NP][TM], b[NP][TN];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
          a[p][i] = *reinterpret_cast<const bf16x8*>(&As[p][wm0 + i * 16 + (lane & 15)][kq]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          b[p][j] = *reinterpret_cast<const bf16x8*>(&Bs[p][wn0 + j * 16 + (lane & 15)][kq]);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          // smallest terms first
          if constexpr (NP > 2) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], b[0][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[1][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[2][j], acc[i][j], 0, 0, 0);
          }
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[1][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gn = n0 + wn0 + j * 16 + (lane & 15);
    if (gn >= N) continue;
    const int64_t coff = map_off(g.c_col, gn);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gm = m0 + wm0 + i * 16 + 4 * (lane >> 4) + v;
        if (gm < M) put_c(C, g, map_off(g.c_row, gm) + coff, gm, (double)acc[i][j][v]);
      }
  }
}

// ----------------------------------------------------------------------------
// Exact split-BF16 mode (TN_GEMM_BF16_EXACT): FP64-grade products from BF16
// MFMAs with no rounding in the slice products or their sums (fixed-point
// slicing after K. Ozaki, T. Ogita, S. Oishi, S. M. Rump, Numer. Algorithms
// 59 (2012); integer slices on low-precision matrix units as in H. Ootomo,
// K. Ozaki, R. Yokota, IJHPCA (2024)).
//   * Row m of A is scaled by 2^-ea[m] and column n of B by 2^-eb[n], with
//     ea, eb the exponents that bring the row / column maximum into
//     [1/2, 1) (exact: powers of two).
//   * A scaled value x in (-1, 1) is cut into kExactSlices integer slices of
//     7 bits: y = 128 x, q = rint(y), x = y - q, repeated; q_1 is in
//     [-128, 128], later slices in [-64, 64], all exact in BF16 (8
//     significant bits), and x = sum_s q_s 2^-7s + r with |r| <= 2^-7S / 2.
//   * The slice products q_s(A) q_t(B) with s + t <= S + 1 are accumulated
//     per diagonal d = s + t on v_mfma_f32_16x16x32_bf16: every product is
//     an integer of magnitude at most 2^14 and a diagonal gains at most
//     6 * 32 * 2^14 per 32 k, so after 128 k its FP32 sum is still an integer
//     below 2^24, i.e. exact; it is then folded into one of two FP64
//     accumulators and cleared: G1 += 2^14 acc_0 + 2^7 acc_1 + acc_2 and
//     G2 += 2^14 acc_3 + 2^7 acc_4 + acc_5, integers below 2^39 per 128 k,
//     so both sums stay exact for K below 2^21.
//   * C(m, n) = 2^(ea[m] + eb[n]) (2^-49 G2 + 2^-28 G1): the sum of all six
//     diagonals, sum_d 2^-7(d+2) acc_d, rounded once.
//   * The slices are formed once per launch (exact_prep_kernel, 12 bytes
//     per operand element in the plan's slice buffer) and the tiles copy
//     them; slicing in every tile cost each A element one slicing per tile
//     column (the kernel ran at 9 TFLOP/s FP64-equivalent at 4096^3).
// The error is the dropped slice products (s + t > S + 1) and remainders:
// relative to max_k |A(m, k)| max_k |B(k, n)|, at most about K 2^-7(S+1).
constexpr int kExactSlices = 6;

constexpr int kSliceK = 64;   // K padding of the slices: a multiple of every exact tile's k per stage

// The slices of eight consecutive k of one row (A) or column (B), already
// scaled by its power of two: six 7-bit integers per element, x = sum_s q_s
// 2^-7(s+1) + r, each exact in BF16; slice q goes to dst + q * stride.
__device__ __forceinline__ void store_slices(double (&x)[8], __bf16* __restrict__ dst, int64_t stride) {
#pragma unroll
  for (int q = 0; q < kExactSlices; ++q) {
    bf16x8 v;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double y = 128.0 * x[u], t = rint(y);
      x[u] = y - t;
      // t is an integer of magnitude at most 128: through int it converts
      // exactly in two instructions (a direct double -> BF16 conversion
      // brings a round-to-odd sequence of ten)
      v[u] = (__bf16)(float)(int)t;
    }
    *reinterpret_cast<bf16x8*>(dst + (int64_t)q * stride) = v;
  }
}

// The operands' scale exponents and slices, once per launch, one unit per
// block of kPrepWaves waves from a list built with the plan: kPrepRows rows
// of A or 64 columns of B.
// Every wave issues all its loads at once (one memory round trip) when the
// padded K is at most kPrepKp, keeping the values in registers from the
// largest magnitude to the slicing; longer rows and columns loop.
//   A: 16 lanes per row, four rows per wave; a lane takes 8-k chunks
//      c = j, j + 16, ... (j = lane % 16), a 16-lane max gives the row's
//      exponent, the lane slices its chunks.
//   B: one column per lane; wave w takes chunks w, w + kPrepWaves, ...; the
//      waves' maxima meet in LDS, then each thread slices its chunks.
// Exponents at ea_off + m / eb_off + n, slices as [slice][row][kp] (A) and
// [slice][column][kp] (B), k fastest, zero past K.  It replaced a scales
// kernel and a slices kernel, each launched over (largest group) x (groups)
// blocks and converting the slice digits through a round-to-odd double ->
// BF16 sequence: 27.5 us for the U(1) chi = 1024 sector group against 15
// us, most of which is the first touch of A and B after a cache flush.
// PK: every group is plain on the K side (offset k * stride, no index map).
struct PrepUnit {
  int32_t group, kind, first, count;   // kind 0: A rows [first, first + count); 1: B columns
};
constexpr int kPrepThreads = 512;
constexpr int kPrepWaves = kPrepThreads / 64;
constexpr int kPrepRows = 4 * kPrepWaves;
constexpr int kPrepKp = 256;           // K padding held in registers (A: 2 chunks of 8 per lane, B: 4)
template <bool PK>
__global__ void __launch_bounds__(kPrepThreads) exact_prep_kernel(const double* __restrict__ A, const double* __restrict__ B,
                                                        const GroupDev* __restrict__ groups,
                                                        const PrepUnit* __restrict__ units, int32_t* __restrict__ ex,
                                                        __bf16* __restrict__ sl) {
  __shared__ double part[kPrepWaves][64];
  __shared__ int scl[64];
  const PrepUnit un = units[blockIdx.x];
  const GroupDev& g = groups[un.group];
  if (group_off(g)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t K = g.K, kp = g.kp;
  const bool fits = kp <= kPrepKp;
  if (un.kind == 0) {
    constexpr int CH = kPrepKp / (8 * 16);
    const int64_t a_ks = g.a_col.stride[TN_MAX_MAP_LEVELS - 1];
    const int j = lane & 15, rw = 4 * wave + (lane >> 4);
    const bool live = rw < un.count;
    const int64_t m = un.first + (live ? rw : 0);
    const int64_t ro = g.a_off + map_off(g.a_row, m);
    auto at = [&](int64_t k) { return A[ro + (PK ? k * a_ks : map_off(g.a_col, k))]; };
    // the 16 lanes of a row: max over lanes j ^ 1 .. j ^ 8 (every lane ends
    // with the row's maximum)
    auto row_max = [&](double v) {
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
      return v;
    };
    if (fits) {
      double x[CH][8];
      double mx = 0.0;
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t k = 8 * (j + 16 * c) + u;
          x[c][u] = (live && k < K) ? at(k) : 0.0;
        }
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int u = 0; u < 8; ++u) mx = fmax(mx, fabs(x[c][u]));
      mx = row_max(mx);
      const int sc = (mx > 0.0) ? ilogb(mx) + 1 : 0;
      if (live && j == 0) ex[g.ea_off + m] = sc;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int64_t k0 = 8 * (j + 16 * c);
        if (live && k0 < kp) {
#pragma unroll
          for (int u = 0; u < 8; ++u) x[c][u] = ldexp(x[c][u], -sc);
          store_slices(x[c], sl + g.sa_off + m * kp + k0, g.M * kp);
        }
      }
    } else {
      double mx = 0.0;
      for (int64_t k = j; live && k < K; k += 16) mx = fmax(mx, fabs(at(k)));
      mx = row_max(mx);
      const int sc = (mx > 0.0) ? ilogb(mx) + 1 : 0;
      if (live && j == 0) ex[g.ea_off + m] = sc;
      for (int64_t k0 = 8 * j; live && k0 < kp; k0 += 128) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = (k0 + u < K) ? ldexp(at(k0 + u), -sc) : 0.0;
        store_slices(x, sl + g.sa_off + m * kp + k0, g.M * kp);
      }
    }
  } else {
    constexpr int CH = kPrepKp / (8 * kPrepWaves);
    const int64_t b_ks = g.b_row.stride[TN_MAX_MAP_LEVELS - 1];
    const int64_t n = un.first + lane;
    const bool live = lane < un.count;
    const int64_t co = g.b_off + (live ? map_off(g.b_col, n) : 0);
    auto at = [&](int64_t k) { return B[co + (PK ? k * b_ks : map_off(g.b_row, k))]; };
    double x[CH][8];
    double mx = 0.0;
    if (fits) {
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t k = 8 * (wave + kPrepWaves * c) + u;
          x[c][u] = (live && k < K) ? at(k) : 0.0;
        }
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int u = 0; u < 8; ++u) mx = fmax(mx, fabs(x[c][u]));
    } else {
      for (int64_t k = wave; live && k < K; k += kPrepWaves) mx = fmax(mx, fabs(at(k)));
    }
    part[wave][lane] = mx;
    __syncthreads();
    if (wave == 0) {
      double m4 = part[0][lane];
#pragma unroll
      for (int w = 1; w < kPrepWaves; ++w) m4 = fmax(m4, part[w][lane]);
      const int sc = (m4 > 0.0) ? ilogb(m4) + 1 : 0;
      scl[lane] = sc;
      if (live) ex[g.eb_off + n] = sc;
    }
    __syncthreads();
    const int sc = scl[lane];
    if (fits) {
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int64_t k0 = 8 * (wave + kPrepWaves * c);
        if (live && k0 < kp) {
#pragma unroll
          for (int u = 0; u < 8; ++u) x[c][u] = ldexp(x[c][u], -sc);
          store_slices(x[c], sl + g.sb_off + n * kp + k0, g.N * kp);
        }
      }
    } else {
      for (int64_t k0 = 8 * wave; live && k0 < kp; k0 += 8 * kPrepWaves) {
        double y[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) y[u] = (k0 + u < K) ? ldexp(at(k0 + u), -sc) : 0.0;
        store_slices(y, sl + g.sb_off + n * kp + k0, g.N * kp);
      }
    }
  }
}

template <int BM, int BN, int BK>
__global__ void __launch_bounds__(256) gemm_exact_bf16(const __bf16* __restrict__ sl, double* __restrict__ C,
                                                       const GroupDev* __restrict__ groups,
                                                       const TileDev* __restrict__ tiles,
                                                       const int32_t* __restrict__ ex) {
  constexpr int S = kExactSlices;
  static_assert(BK % 32 == 0 && 256 % BK == 0 && 256 % BN == 0, "tile shape");
  constexpr int KP = BK + 8;                       // padded row (bf16 elements)
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int kFlush = 4;                        // 32-k steps between FP64 flushes (128 k)
  __shared__ __attribute__((aligned(16))) __bf16 As[S][BM][KP];    // [slice][m][k]
  __shared__ __attribute__((aligned(16))) __bf16 Bs[S][BN][KP];    // [slice][n][k]

  const TileDev tile = tiles[blockIdx.x];
  if (tile.group < 0) return;
  const GroupDev& g = groups[tile.group];
  if (group_off(g)) return;
  const int64_t M = g.M, N = g.N, K = g.K;
  const int64_t m0 = (int64_t)tile.tm * BM, n0 = (int64_t)tile.tn * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;

  static_assert(S == 6, "the FP64 folding takes the diagonals three at a time");
  f4 acc[S][TM][TN];
  double gacc[2][TM][TN][4];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[d][i][j] = f4{0, 0, 0, 0};
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) gacc[h][i][j][v] = 0.0;
  // every term is an integer below 2^53: the additions are exact
  auto flush = [&]() {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            gacc[h][i][j][v] += 16384.0 * (double)acc[3 * h][i][j][v] + 128.0 * (double)acc[3 * h + 1][i][j][v] +
                                (double)acc[3 * h + 2][i][j][v];
#pragma unroll
          for (int d = 3 * h; d < 3 * h + 3; ++d) acc[d][i][j] = f4{0, 0, 0, 0};
        }
  };

  // A stage of the tile is S x (BM + BN) x BK slices, read as 16-byte
  // vectors of eight k (rows past M or N read as zeros)
  constexpr int KV = BK / 8;                       // vectors per row of a stage
  constexpr int AV = (BM * KV + 255) / 256, BV = (BN * KV + 255) / 256;
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const u4* __restrict__ sa = reinterpret_cast<const u4*>(sl + g.sa_off);
  const u4* __restrict__ sb = reinterpret_cast<const u4*>(sl + g.sb_off);
  const int64_t kpv = g.kp / 8;                    // vectors per slice row
  u4 ra[S][AV], rb[S][BV];
  auto load_tiles = [&](int64_t k0) {
    const int64_t kv0 = k0 / 8;
#pragma unroll
    for (int i = 0; i < AV; ++i) {
      const int v = tid + 256 * i, row = v / KV, kv = v % KV;
      const bool ok = v < BM * KV && m0 + row < M;
      const int64_t o = (m0 + row) * kpv + kv0 + kv;
#pragma unroll
      for (int q = 0; q < S; ++q) ra[q][i] = ok ? sa[(int64_t)q * M * kpv + o] : u4{0, 0, 0, 0};
    }
#pragma unroll
    for (int i = 0; i < BV; ++i) {
      const int v = tid + 256 * i, row = v / KV, kv = v % KV;
      const bool ok = v < BN * KV && n0 + row < N;
      const int64_t o = (n0 + row) * kpv + kv0 + kv;
#pragma unroll
      for (int q = 0; q < S; ++q) rb[q][i] = ok ? sb[(int64_t)q * N * kpv + o] : u4{0, 0, 0, 0};
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int i = 0; i < AV; ++i) {
      const int v = tid + 256 * i;
      if (v < BM * KV) {
#pragma unroll
        for (int q = 0; q < S; ++q) *reinterpret_cast<u4*>(&As[q][v / KV][8 * (v % KV)]) = ra[q][i];
      }
    }
#pragma unroll
    for (int i = 0; i < BV; ++i) {
      const int v = tid + 256 * i;
      if (v < BN * KV) {
#pragma unroll
        for (int q = 0; q < S; ++q) *reinterpret_cast<u4*>(&Bs[q][v / KV][8 * (v % KV)]) = rb[q][i];
      }
    }
  };

  const int64_t nk = (K + BK - 1) / BK;
  int steps = 0;
  load_tiles(0);
  for (int64_t kt = 0; kt < nk; ++kt) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (kt + 1 < nk) load_tiles((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      const int kq = ks + 8 * (lane >> 4);
      bf16x8 a[S][TM], b[S][TN];
#pragma unroll
      for (int sl = 0; sl < S; ++sl) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
          a[sl][i] = *reinterpret_cast<const bf16x8*>(&As[sl][wm0 + i * 16 + (lane & 15)][kq]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          b[sl][j] = *reinterpret_cast<const bf16x8*>(&Bs[sl][wn0 + j * 16 + (lane & 15)][kq]);
      }
#pragma unroll
      for (int d = 0; d < S; ++d)
#pragma unroll
        for (int sa = 0; sa <= d; ++sa)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[d][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[sa][i], b[d - sa][j], acc[d][i][j], 0, 0, 0);
      if (++steps == kFlush) {
        flush();
        steps = 0;
      }
    }
  }
  flush();
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gn = n0 + wn0 + j * 16 + (lane & 15);
    if (gn >= N) continue;
    const int64_t coff = map_off(g.c_col, gn);
    const int en = ex[g.eb_off + gn];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gm = m0 + wm0 + i * 16 + 4 * (lane >> 4) + v;
        if (gm >= M) continue;
        const double val = ldexp(gacc[1][i][j][v], -49) + ldexp(gacc[0][i][j][v], -28);
        put_c(C, g, map_off(g.c_row, gm) + coff, gm, ldexp(val, ex[g.ea_off + gm] + en));
      }
  }
}

// ----------------------------------------------------------------------------
// Tile configurations (autotuner ids), per arithmetic mode.
struct TileCfg {
  int bm, bn;
};
// FP64/FP32: five 256-thread shapes, the 32x32 tile with 512 threads (two
// waves per sub-tile) at stage depths 32 (id 5), 64 (id 6) and 128 (id 7):
// a deeper stage means fewer barriers per K and a longer prefetch distance
// in time; and the 32x32 tile with two stages of loads in flight instead of
// three (id 8: fewer registers, four blocks per CU, for short K).
// Split-BF16: five shapes.
constexpr int kMaxCfg = 9;
constexpr TileCfg kCfgMfma[kMaxCfg] = {{32, 32}, {64, 64}, {128, 64}, {64, 128}, {64, 32},
                                       {32, 32}, {32, 32}, {32, 32}, {32, 32}};
constexpr TileCfg kCfgBf16[kMaxCfg] = {{32, 32},   {64, 64},   {128, 64},  {64, 128}, {128, 128},
                                       {128, 128}, {128, 128}, {128, 128}, {128, 128}};

// exact split-BF16: 32x32 at 64 and 32 k per stage, 64x32, 32x64 (a 64x64
// tile, one wave per SIMD, was measured slower at 4096^3: the stage is
// bound by the LDS reads of the slice fragments, not by reloading slices)
constexpr TileCfg kCfgExact[kMaxCfg] = {{32, 32}, {32, 32}, {64, 32}, {32, 64}, {32, 32},
                                        {32, 32}, {32, 32}, {32, 32}, {32, 32}};
bool mfma_mode(int mode) { return mode == TN_GEMM_F64 || mode == TN_GEMM_F32; }
int num_cfg(int mode) { return mfma_mode(mode) ? 9 : (mode == TN_GEMM_BF16_EXACT ? 4 : 5); }
int cfg_threads(int mode, int cfg) { return (mfma_mode(mode) && cfg >= 5 && cfg <= 7) ? 512 : 256; }
// k values per pipeline stage of a tile shape (the placement weighs tiles
// by their number of stages)
int cfg_kc(int mode, int cfg) {
  if (mfma_mode(mode)) return (cfg == 1 || cfg == 4 || cfg == 6) ? 64 : (cfg == 7 ? 128 : 32);
  return 32;
}

TileCfg cfg_shape(int mode, int cfg) {
  return mfma_mode(mode) ? kCfgMfma[cfg] : (mode == TN_GEMM_BF16_EXACT ? kCfgExact[cfg] : kCfgBf16[cfg]);
}

typedef void (*ExactFn)(const __bf16*, double*, const GroupDev*, const TileDev*, const int32_t*);
ExactFn pick_exact(int cfg) {
  switch (cfg) {
    case 0: return gemm_exact_bf16<32, 32, 64>;
    case 1: return gemm_exact_bf16<32, 32, 32>;
    case 2: return gemm_exact_bf16<64, 32, 32>;
    case 3: return gemm_exact_bf16<32, 64, 32>;
  }
  return nullptr;
}

typedef void (*KernFn)(const double*, const double*, double*, const GroupDev*, const TileDev*);

template <typename T, bool VEC, bool KT, bool PK>
KernFn pick_mfma(int cfg) {
  switch (cfg) {
    case 0: return gemm_mfma_k32<T, 32, 32, 32, 3, VEC, KT, 256, PK>;
    case 1: return gemm_mfma_k32<T, 64, 64, 64, 2, VEC, KT, 256, PK>;
    case 2: return gemm_mfma_k32<T, 128, 64, 32, 2, VEC, KT, 256, PK>;
    case 3: return gemm_mfma_k32<T, 64, 128, 32, 2, VEC, KT, 256, PK>;
    case 4: return gemm_mfma_k32<T, 64, 32, 64, 2, VEC, KT, 256, PK>;
    case 5: return gemm_mfma_k32<T, 32, 32, 32, 3, VEC, KT, 512, PK>;
    case 6: return gemm_mfma_k32<T, 32, 32, 64, 4, VEC, KT, 512, PK>;
    case 7: return gemm_mfma_k32<T, 32, 32, 128, 2, VEC, KT, 512, PK>;
    case 8: return gemm_mfma_k32<T, 32, 32, 32, 2, VEC, KT, 256, PK>;
  }
  return nullptr;
}

// pk: plain K-side layouts (the tables are not needed then)
template <typename T>
KernFn pick_mfma(int cfg, bool vec, bool kt, bool pk) {
  if (pk) return vec ? pick_mfma<T, true, false, true>(cfg) : pick_mfma<T, false, false, true>(cfg);
  if (vec) return kt ? pick_mfma<T, true, true, false>(cfg) : pick_mfma<T, true, false, false>(cfg);
  return kt ? pick_mfma<T, false, true, false>(cfg) : pick_mfma<T, false, false, false>(cfg);
}

KernFn pick_kernel(int mode, int cfg, bool vec, bool kt, bool pk) {
  switch (mode) {
    case TN_GEMM_F64: return pick_mfma<double>(cfg, vec, kt, pk);
    case TN_GEMM_F32: return pick_mfma<float>(cfg, vec, kt, pk);
    case TN_GEMM_BF16X3:
      switch (cfg) {
        case 0: return gemm_split_bf16<32, 32, 2, 128>;
        case 1: return gemm_split_bf16<64, 64, 2, 64>;
        case 2: return gemm_split_bf16<128, 64, 2, 32>;
        case 3: return gemm_split_bf16<64, 128, 2, 32>;
        case 4: return gemm_split_bf16<128, 128, 2, 32>;
      }
      break;
    case TN_GEMM_BF16X6:
      switch (cfg) {
        case 0: return gemm_split_bf16<32, 32, 3, 128>;
        case 1: return gemm_split_bf16<64, 64, 3, 64>;
        case 2: return gemm_split_bf16<128, 64, 3, 32>;
        case 3: return gemm_split_bf16<64, 128, 3, 32>;
        case 4: return gemm_split_bf16<128, 128, 3, 32>;
      }
      break;
  }
  return nullptr;
}

// Every (k, k+1) pair of A and (n, n+1) pair of B that the FP64/FP32 kernel
// loads is contiguous and at an even element offset: then one 16-byte load
// reads it (given 16-byte aligned base pointers, checked at launch).
bool map_pairs_ok(const tn_index_map& m, int64_t range) {
  if (m.n == 0) return range % 2 == 0;
  if (m.stride[m.n - 1] != 1 || m.ext[m.n - 1] % 2 != 0) return false;
  for (int l = 0; l + 1 < m.n; ++l)
    if (m.ext[l] > 1 && m.stride[l] % 2 != 0) return false;
  return true;
}
bool map_even(const tn_index_map& m, int64_t ld, int64_t range) {
  if (m.n == 0) return range <= 1 || ld % 2 == 0;
  for (int l = 0; l < m.n; ++l)
    if (m.ext[l] > 1 && m.stride[l] % 2 != 0) return false;
  return true;
}
// Largest offset an index map produces over [0, range).
int64_t map_max_off(const tn_index_map& m, int64_t range, int64_t plain_stride) {
  if (range <= 0) return 0;
  if (m.n == 0) return (range - 1) * plain_stride;
  int64_t o = 0;
  for (int l = 0; l < m.n; ++l) o += (m.ext[l] - 1) * m.stride[l];
  return o;
}
bool desc_vec_ok(const tn_gemm_desc_ex& x) {
  return x.a_off % 2 == 0 && x.b_off % 2 == 0 && map_pairs_ok(x.a_col, x.K) && map_even(x.a_row, x.lda, x.M) &&
         map_pairs_ok(x.b_col, x.N) && map_even(x.b_row, x.ldb, x.K);
}

// Host: 32-bit division magic for divisor d >= 1 (exact for dividends < 2^31).
void div_magic(uint32_t d, uint32_t* magic, uint32_t* shift) {
  uint32_t l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  *shift = l;
  *magic = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
}

bool make_map(const tn_index_map& in, int64_t range, int64_t plain_stride, MapDev* out) {
  std::memset(out, 0, sizeof(*out));
  int64_t ext[TN_MAX_MAP_LEVELS], str[TN_MAX_MAP_LEVELS];
  int n = in.n;
  if (n < 0 || n > TN_MAX_MAP_LEVELS) return false;
  if (range == 0) n = 0;
  if (n == 0) {
    n = 1;
    ext[0] = std::max<int64_t>(range, 1);
    str[0] = plain_stride;
  } else {
    int64_t prod = 1;
    for (int l = 0; l < n; ++l) {
      if (in.ext[l] < 1 || in.ext[l] >= (int64_t(1) << 31) || in.stride[l] < 0) return false;
      prod *= in.ext[l];
      if (prod >= (int64_t(1) << 31)) return false;
      ext[l] = in.ext[l];
      str[l] = in.stride[l];
    }
    if (prod != range) return false;
  }
  if (ext[0] >= (int64_t(1) << 31)) return false;
  const int pad = TN_MAX_MAP_LEVELS - n;
  for (int l = 0; l < TN_MAX_MAP_LEVELS; ++l) {
    const bool real = l >= pad;
    out->ext[l] = real ? (uint32_t)ext[l - pad] : 1u;
    out->stride[l] = real ? str[l - pad] : 0;
    div_magic(out->ext[l], &out->magic[l], &out->shift[l]);
  }
  return true;
}

}  // namespace

// The tile list of one tile shape: built on first use and kept until the
// plan is destroyed, so a launch still queued with one shape never reads a
// list freed by a later switch (set_tile, autotune) to another.
struct TileSet {
  TileDev* d_tiles = nullptr;
  int64_t ntiles = 0;     // real tiles
  int64_t nslots = 0;     // launched blocks (with XCD padding)
  bool built = false;
};

struct tn_gemm_plan_s {
  int mode = 0;
  std::atomic<int> cfg{-1};   // current tile shape
  std::mutex mu;              // guards tiles[]
  TileSet tiles[kMaxCfg];
  std::vector<GroupDev> groups;
  GroupDev* d_groups = nullptr;
  int32_t* d_exp = nullptr;   // exact mode: scale exponents (A rows, then B columns, per group)
  __bf16* d_sl = nullptr;     // exact mode: the operands' BF16 slices (per group: A, then B)
  int64_t sl_elems = 0;       // their number
  PrepUnit* d_units = nullptr;   // exact mode: the preparation kernel's work list
  int64_t nunits = 0;
  double flops = 0;
  bool vec_ok = true;     // every group admits 16-byte pair loads
  int64_t max_k = 0;      // longest K over the groups
  bool koff32 = true;     // every K-side offset fits in int32 (LDS tables)
  bool plain_k = true;    // every group reads A columns / B rows in the plain layout
};

// Builds (once) the tile list of shape cfg; the plan is unchanged on
// failure.
static int ensure_tiles(tn_gemm_plan_s* P, int cfg) {
  std::lock_guard<std::mutex> g(P->mu);
  if (P->tiles[cfg].built) return TN_OK;
  const TileCfg sh = cfg_shape(P->mode, cfg);
  const int bm = sh.bm, bn = sh.bn;
  const int64_t kc = cfg_kc(P->mode, cfg);
  // Each group's tile grid is cut into px x py compact regions (px * py =
  // 8, the split that minimises the rows of A plus columns of B one region
  // reads); regions are the unit of XCD placement, so tiles that share A
  // rows and B columns share one XCD's L2.
  struct Region {
    std::vector<TileDev> tiles;
    int64_t work = 0;   // tiles x (K stages + 2): the descriptor, first-stage and write-back latency count
  };
  std::vector<Region> regions;
  for (size_t gi = 0; gi < P->groups.size(); ++gi) {
    const GroupDev& g = P->groups[gi];
    if (g.M == 0 || g.N == 0 || g.K == 0) continue;
    int64_t tm = (g.M + bm - 1) / bm, tn_ = (g.N + bn - 1) / bn;
    TN_REQUIRE(tm < (1 << 30) && tn_ < (1 << 30), TN_EUNSUPP, "GEMM too large");
    int64_t px = 1, py = 8, best = INT64_MAX;
    for (int64_t cx : {1, 2, 4, 8}) {
      const int64_t cy = 8 / cx;
      const int64_t cost = ((tm + cx - 1) / cx) * bm + ((tn_ + cy - 1) / cy) * bn;
      if (cost < best) {
        best = cost;
        px = cx;
        py = cy;
      }
    }
    const int64_t rm = (tm + px - 1) / px, rn = (tn_ + py - 1) / py;
    const int64_t w = (g.K + kc - 1) / kc + 2;
    for (int64_t rx = 0; rx < px; ++rx)
      for (int64_t ry = 0; ry < py; ++ry) {
        Region r;
        for (int64_t b = ry * rn; b < std::min(tn_, (ry + 1) * rn); ++b)
          for (int64_t a = rx * rm; a < std::min(tm, (rx + 1) * rm); ++a) r.tiles.push_back(TileDev{(int32_t)gi, (int32_t)a, (int32_t)b, 0});
        r.work = w * (int64_t)r.tiles.size();
        if (!r.tiles.empty()) regions.push_back(std::move(r));
      }
  }
  // XCD-aware placement (the hardware deals block ids round-robin over the
  // 8 XCDs: block id = slot * 8 + xcd).  Regions go to XCDs heaviest first,
  // each to the least loaded XCD (longest-processing-time rule), so every
  // XCD gets about the same number of k stages.  Within an XCD the tiles
  // run heaviest first, dealt to its CUs in rounds of one tile per CU with
  // every other round reversed, so no CU collects the heavy tiles of
  // several rounds.
  const int64_t kX = 8;
  const int64_t cus = std::max<int64_t>(1, tn::cu_count() / kX);
  std::stable_sort(regions.begin(), regions.end(), [](const Region& a, const Region& b) { return a.work > b.work; });
  std::vector<std::vector<TileDev>> xcd(kX);
  int64_t load[kX] = {0};
  for (const Region& r : regions) {
    int x = 0;
    for (int i = 1; i < kX; ++i)
      if (load[i] < load[x]) x = i;
    load[x] += r.work;
    xcd[x].insert(xcd[x].end(), r.tiles.begin(), r.tiles.end());
  }
  int64_t n = 0, per = 0;
  for (auto& v : xcd) {
    std::stable_sort(v.begin(), v.end(), [&](const TileDev& a, const TileDev& b) {
      return P->groups[a.group].K > P->groups[b.group].K;
    });
    for (size_t r = 1; (int64_t)r * cus < (int64_t)v.size(); r += 2)
      std::reverse(v.begin() + r * cus, v.begin() + std::min<int64_t>(v.size(), (r + 1) * cus));
    n += (int64_t)v.size();
    per = std::max<int64_t>(per, (int64_t)v.size());
  }
  const int64_t slots = per * kX;
  std::vector<TileDev> lay(slots, TileDev{-1, 0, 0, 0});
  for (int64_t x = 0; x < kX; ++x)
    for (size_t i = 0; i < xcd[x].size(); ++i) lay[i * kX + x] = xcd[x][i];
  TN_REQUIRE(slots < (int64_t(1) << 31), TN_EUNSUPP, "too many GEMM tiles");
  TileSet t;
  if (slots > 0) {
    int rc = tn::pool_alloc((void**)&t.d_tiles, sizeof(TileDev) * slots);
    if (rc == TN_OK) rc = tn_memcpy_h2d(t.d_tiles, lay.data(), sizeof(TileDev) * slots, nullptr);
    if (rc != TN_OK) {
      tn::pool_free(t.d_tiles);
      return rc;
    }
  }
  t.ntiles = n;
  t.nslots = slots;
  t.built = true;
  P->tiles[cfg] = t;
  return TN_OK;
}

static TileSet tiles_of(tn_gemm_plan_s* P, int cfg) {
  std::lock_guard<std::mutex> g(P->mu);
  return P->tiles[cfg];
}

// Makes cfg the plan's tile shape (building its list first).
static int select_cfg(tn_gemm_plan_s* P, int cfg) {
  int rc = ensure_tiles(P, cfg);
  if (rc == TN_OK) P->cfg.store(cfg);
  return rc;
}

// Default tile: the largest tile that still gives every CU at least one
// tile; otherwise the smallest tile.
static int default_cfg(const tn_gemm_plan_s* P) {
  const int nc = num_cfg(P->mode);
  int order[kMaxCfg];
  for (int c = 0; c < nc; ++c) order[c] = c;
  std::sort(order, order + nc, [&](int a, int b) {
    const TileCfg x = cfg_shape(P->mode, a), y = cfg_shape(P->mode, b);
    return x.bm * x.bn > y.bm * y.bn || (x.bm * x.bn == y.bm * y.bn && a < b);
  });
  const int64_t want = tn::cu_count();
  for (int i = 0; i < nc; ++i) {
    const int c = order[i];
    const TileCfg sh = cfg_shape(P->mode, c);
    int64_t t = 0;
    for (const auto& g : P->groups)
      t += ((g.M + sh.bm - 1) / sh.bm) * ((g.N + sh.bn - 1) / sh.bn);
    if (t >= want) return c;
  }
  return order[nc - 1];
}

static void destroy_plan(tn_gemm_plan_s* P) {
  tn::pool_free(P->d_groups);
  tn::pool_free(P->d_exp);
  tn::pool_free(P->d_sl);
  tn::pool_free(P->d_units);
  for (TileSet& t : P->tiles) tn::pool_free(t.d_tiles);
  delete P;
}

static int exec_cfg(tn_gemm_plan_s* P, int cfg, const double* A, const double* B, double* C, void* stream);

extern "C" {

int tn_gemm_plan_create_ex(tn_gemm_plan* plan, int mode, int64_t ngroups, const tn_gemm_desc_ex* d) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(plan != nullptr, TN_EINVAL, "tn_gemm_plan_create: NULL out pointer");
  TN_REQUIRE(mode >= TN_GEMM_F64 && mode <= TN_GEMM_BF16_EXACT, TN_EINVAL, "unknown GEMM mode %d", mode);
  TN_REQUIRE(ngroups >= 0 && (ngroups == 0 || d), TN_EINVAL, "tn_gemm_plan_create: NULL descs");
  auto* P = new tn_gemm_plan_s();
  P->mode = mode;
  std::vector<PrepUnit> units;
  for (int64_t i = 0; i < ngroups; ++i) {
    const tn_gemm_desc_ex& x = d[i];
    GroupDev G;
    std::memset(&G, 0, sizeof(G));
    G.M = x.M;
    G.N = x.N;
    G.K = x.K;
    G.a_off = x.a_off;
    G.b_off = x.b_off;
    G.c_off = x.c_off;
    G.lda = x.lda;
    G.ldb = x.ldb;
    G.ldc = x.ldc;
    G.d_off = x.d_off;
    G.rs_off = x.rs_off;
    G.rs_div = x.rs_div;
    const bool ld_ok = (x.a_row.n || x.M == 0 || x.K == 0 || x.lda >= x.K) &&
                       (x.b_row.n || x.K == 0 || x.N == 0 || x.ldb >= x.N) &&
                       (x.c_row.n || x.M == 0 || x.N == 0 || x.ldc >= x.N);
    const bool range_ok = x.M < (int64_t(1) << 31) && x.N < (int64_t(1) << 31) && x.K < (int64_t(1) << 31);
    const bool maps_ok = make_map(x.a_row, x.M, x.lda, &G.a_row) && make_map(x.a_col, x.K, 1, &G.a_col) &&
                         make_map(x.b_row, x.K, x.ldb, &G.b_row) && make_map(x.b_col, x.N, 1, &G.b_col) &&
                         make_map(x.c_row, x.M, x.ldc, &G.c_row) && make_map(x.c_col, x.N, 1, &G.c_col);
    const bool rs_ok = x.rs_div == 0 || (x.rs_div > 0 && x.d_off >= 0 && x.rs_off >= 0);
    if (x.M < 0 || x.N < 0 || x.K < 0 || x.a_off < 0 || x.b_off < 0 || x.c_off < 0 || !ld_ok || !range_ok ||
        !maps_ok || !rs_ok) {
      delete P;
      tn::set_error("GEMM group %lld has an invalid shape, leading dimension or index map", (long long)i);
      return TN_EINVAL;
    }
    int64_t nexp = 0;
    for (const GroupDev& h : P->groups) nexp += h.M + h.N;
    G.ea_off = nexp;
    G.eb_off = nexp + x.M;
    if (mode == TN_GEMM_BF16_EXACT && x.M > 0 && x.N > 0 && x.K > 0) {
      G.kp = (x.K + kSliceK - 1) / kSliceK * kSliceK;
      G.sa_off = P->sl_elems;
      G.sb_off = G.sa_off + kExactSlices * x.M * G.kp;
      P->sl_elems = G.sb_off + kExactSlices * x.N * G.kp;
      const int32_t gi = (int32_t)P->groups.size();
      for (int64_t f = 0; f < x.M; f += kPrepRows)
        units.push_back({gi, 0, (int32_t)f, (int32_t)std::min<int64_t>(kPrepRows, x.M - f)});
      for (int64_t f = 0; f < x.N; f += 64) units.push_back({gi, 1, (int32_t)f, (int32_t)std::min<int64_t>(64, x.N - f)});
    }
    P->groups.push_back(G);
    P->flops += 2.0 * (double)x.M * (double)x.N * (double)x.K;
    if (x.M > 0 && x.N > 0 && x.K > 0 && !desc_vec_ok(x)) P->vec_ok = false;
    if (x.M > 0 && x.N > 0 && x.K > 0 && (x.a_col.n != 0 || x.b_row.n != 0)) P->plain_k = false;
    if (x.M > 0 && x.N > 0) {
      P->max_k = std::max(P->max_k, x.K);
      if (map_max_off(x.a_col, x.K, 1) >= (int64_t(1) << 31) || map_max_off(x.b_row, x.K, x.ldb) >= (int64_t(1) << 31))
        P->koff32 = false;
    }
  }
  int rc = TN_OK;
  if (!P->groups.empty()) {
    rc = tn::pool_alloc((void**)&P->d_groups, sizeof(GroupDev) * P->groups.size());
    if (rc == TN_OK)
      rc = tn_memcpy_h2d(P->d_groups, P->groups.data(), sizeof(GroupDev) * P->groups.size(), nullptr);
    if (rc == TN_OK && mode == TN_GEMM_BF16_EXACT) {
      const GroupDev& last = P->groups.back();
      rc = tn::pool_alloc((void**)&P->d_exp, sizeof(int32_t) * std::max<int64_t>(1, last.eb_off + last.N));
      if (rc == TN_OK) rc = tn::pool_alloc((void**)&P->d_sl, sizeof(__bf16) * std::max<int64_t>(8, P->sl_elems));
      P->nunits = (int64_t)units.size();
      if (rc == TN_OK && !units.empty()) rc = tn::pool_alloc((void**)&P->d_units, sizeof(PrepUnit) * units.size());
      if (rc == TN_OK && !units.empty())
        rc = tn_memcpy_h2d(P->d_units, units.data(), sizeof(PrepUnit) * units.size(), nullptr);
    }
  }
  if (rc == TN_OK) rc = select_cfg(P, default_cfg(P));
  if (rc != TN_OK) {
    destroy_plan(P);
    return rc;
  }
  *plan = P;
  return TN_OK;
}

int tn_gemm_plan_create(tn_gemm_plan* plan, int mode, int64_t ngroups, const tn_gemm_desc* d) {
  TN_REQUIRE(ngroups >= 0 && (ngroups == 0 || d), TN_EINVAL, "tn_gemm_plan_create: NULL descs");
  std::vector<tn_gemm_desc_ex> ex(ngroups);
  for (int64_t i = 0; i < ngroups; ++i) {
    const tn_gemm_desc& x = d[i];
    std::memset(&ex[i], 0, sizeof(tn_gemm_desc_ex));
    ex[i].M = x.M;
    ex[i].N = x.N;
    ex[i].K = x.K;
    ex[i].a_off = x.a_off;
    ex[i].b_off = x.b_off;
    ex[i].c_off = x.c_off;
    ex[i].lda = x.lda;
    ex[i].ldb = x.ldb;
    ex[i].ldc = x.ldc;
  }
  return tn_gemm_plan_create_ex(plan, mode, ngroups, ex.data());
}

int tn_gemm_plan_set_tile(tn_gemm_plan P, int tile_id) {
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_plan_set_tile: NULL plan");
  TN_REQUIRE(tile_id >= 0 && tile_id < num_cfg(P->mode), TN_EINVAL, "tile id %d outside [0,%d)", tile_id,
             num_cfg(P->mode));
  return select_cfg(P, tile_id);
}

int tn_gemm_plan_autotune(tn_gemm_plan P, const double* A, const double* B, double* C, int reps,
                          void* stream, float* best_ms) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_plan_autotune: NULL plan");
  if (P->groups.empty()) return TN_OK;
  if (reps < 1) reps = 1;
  hipStream_t s = tn::stream_of(stream);
  hipEvent_t e0, e1;
  TN_HIP_CHECK(hipEventCreate(&e0));
  TN_HIP_CHECK(hipEventCreate(&e1));
  // Each shape is timed through its own tile list; callers on other threads
  // keep launching the current shape until the best one is selected.  The
  // shapes are timed in interleaved rounds and each keeps its fastest round,
  // so a clock ramp or a neighbour's burst during one shape's turn does not
  // decide the choice; after the first round only the shapes within 25% of
  // the best are timed again.
  constexpr int kRounds = 3;
  const int ncfg = num_cfg(P->mode);
  std::vector<float> t_cfg(ncfg, 1e30f);
  int rc = TN_OK;
  for (int c = 0; c < ncfg && rc == TN_OK; ++c) {
    rc = ensure_tiles(P, c);
    if (rc == TN_OK) rc = exec_cfg(P, c, A, B, C, stream);   // warm-up
  }
  for (int round = 0; round < kRounds && rc == TN_OK; ++round) {
    const float lead = *std::min_element(t_cfg.begin(), t_cfg.end());
    for (int c = 0; c < ncfg && rc == TN_OK; ++c) {
      if (round > 0 && t_cfg[c] > 1.25f * lead) continue;
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < reps && rc == TN_OK; ++r) rc = exec_cfg(P, c, A, B, C, stream);
      (void)hipEventRecord(e1, s);
      if (rc != TN_OK) break;
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      t_cfg[c] = std::min(t_cfg[c], ms / reps);
    }
  }
  int best = P->cfg.load();
  float best_t = 1e30f;
  for (int c = 0; c < ncfg; ++c)
    if (t_cfg[c] < best_t) {
      best_t = t_cfg[c];
      best = c;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != TN_OK) return rc;
  if (best_ms) *best_ms = best_t;
  return select_cfg(P, best);
}

int tn_gemm_plan_set_mask(tn_gemm_plan P, const int* mask, const int32_t* group_idx) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_plan_set_mask: NULL plan");
  TN_REQUIRE(mask == nullptr || group_idx != nullptr || P->groups.empty(), TN_EINVAL,
             "tn_gemm_plan_set_mask: NULL group indices");
  for (size_t g = 0; g < P->groups.size(); ++g) {
    P->groups[g].mask = mask;
    P->groups[g].mask_idx = mask ? (int64_t)group_idx[g] : -1;
  }
  if (P->groups.empty()) return TN_OK;
  int rc = tn_device_sync();   // launches queued with the old descriptors finish first
  if (rc == TN_OK)
    rc = tn_memcpy_h2d(P->d_groups, P->groups.data(), sizeof(GroupDev) * P->groups.size(), nullptr);
  return rc;
}

int tn_gemm_plan_info(tn_gemm_plan P, int64_t* tiles, double* flops, int* tile_id) {
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_plan_info: NULL plan");
  const int c = P->cfg.load();
  if (tiles) *tiles = tiles_of(P, c).ntiles;
  if (flops) *flops = P->flops;
  if (tile_id) *tile_id = c;
  return TN_OK;
}

// Fill kernel for sectors with K == 0 (dgemm with beta = 0 writes zeros).
__global__ void zero_c_kernel(double* __restrict__ C, const GroupDev* __restrict__ gp) {
  const GroupDev& g = *gp;
  if (group_off(g)) return;
  const int64_t total = g.M * g.N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t m = e / g.N, n = e % g.N;
    put_c(C, g, map_off(g.c_row, m) + map_off(g.c_col, n), m, 0.0);
  }
}

int tn_gemm_exec(tn_gemm_plan P, const double* A, const double* B, double* C, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_exec: NULL plan");
  return exec_cfg(P, P->cfg.load(), A, B, C, stream);
}

}  // extern "C"

static int exec_cfg(tn_gemm_plan_s* P, int cfg, const double* A, const double* B, double* C, void* stream) {
  hipStream_t s = tn::stream_of(stream);
  const TileSet t = tiles_of(P, cfg);
  for (size_t gi = 0; gi < P->groups.size(); ++gi) {
    const GroupDev& g = P->groups[gi];
    if (g.K == 0 && g.M > 0 && g.N > 0) {
      TN_REQUIRE(C != nullptr, TN_EINVAL, "tn_gemm_exec: NULL C");
      int64_t blocks = std::min<int64_t>((g.M * g.N + 255) / 256, 1024);
      hipLaunchKernelGGL(zero_c_kernel, dim3((unsigned)blocks), dim3(256), 0, s, C, P->d_groups + gi);
      TN_LAUNCH_CHECK();
    }
  }
  if (t.nslots == 0) return TN_OK;
  TN_REQUIRE(A && B && C, TN_EINVAL, "tn_gemm_exec: NULL operand");
  if (P->mode == TN_GEMM_BF16_EXACT) {
    ExactFn k = pick_exact(cfg);
    TN_REQUIRE(k != nullptr, TN_EINVAL, "no exact kernel for tile %d", cfg);
    // every operand element scaled and sliced once per launch (not once per
    // tile that reads it)
    if (P->nunits > 0) {
      hipLaunchKernelGGL(P->plain_k ? exact_prep_kernel<true> : exact_prep_kernel<false>, dim3((unsigned)P->nunits),
                         dim3(kPrepThreads), 0, s, A, B, P->d_groups, P->d_units, P->d_exp, P->d_sl);
      TN_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k, dim3((unsigned)t.nslots), dim3(256), 0, s, P->d_sl, C, P->d_groups, t.d_tiles, P->d_exp);
    TN_LAUNCH_CHECK();
    return TN_OK;
  }
  const bool vec = P->vec_ok && ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15) == 0;
  KernFn k = pick_kernel(P->mode, cfg, vec, P->koff32 && P->max_k <= kKTab, P->plain_k);
  TN_REQUIRE(k != nullptr, TN_EINVAL, "no kernel for mode %d tile %d", P->mode, cfg);
  hipLaunchKernelGGL(k, dim3((unsigned)t.nslots), dim3(cfg_threads(P->mode, cfg)), 0, s, A, B, C, P->d_groups,
                     t.d_tiles);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

extern "C" {

int tn_gemm_plan_destroy(tn_gemm_plan P) {
  if (!P) return TN_OK;
  // the plan's descriptors may still be read by queued launches
  hipError_t e = hipDeviceSynchronize();
  destroy_plan(P);
  TN_REQUIRE(e == hipSuccess, TN_EHIP, "hipDeviceSynchronize: %s", hipGetErrorString(e));
  return TN_OK;
}

int tn_matmul_f64(const double* A, const double* B, int64_t M, int64_t N, int64_t K, double* C, int mode,
                  void* stream) {
  tn_gemm_desc d{M, N, K, 0, 0, 0, std::max<int64_t>(K, 1), std::max<int64_t>(N, 1), std::max<int64_t>(N, 1)};
  tn_gemm_plan P = nullptr;
  int rc = tn_gemm_plan_create(&P, mode, 1, &d);
  if (rc != TN_OK) return rc;
  rc = tn_gemm_exec(P, A, B, C, stream);
  if (rc == TN_OK) {
    hipError_t e = hipStreamSynchronize(tn::stream_of(stream));
    if (e != hipSuccess) {
      tn::set_error("hipStreamSynchronize: %s", hipGetErrorString(e));
      rc = TN_EHIP;
    }
  }
  tn_gemm_plan_destroy(P);
  return rc;
}

}  // extern "C"
