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
// Every output element is computed by exactly one workgroup, accumulating
// over K in increasing order, one MFMA K-step at a time, for every tile
// shape: results do not depend on the tile shape or on what else is in the
// group (no split-K, no atomics).
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

// Index map on the device: n levels (0 = plain), extents with their 32-bit
// division magic (q = (umulhi(i, magic) + i) >> shift, exact for i < 2^31).
struct MapDev {
  int32_t n;
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
};

__device__ __forceinline__ int64_t map_off(const MapDev& m, int64_t i64, int64_t plain_mul) {
  if (m.n == 0) return i64 * plain_mul;
  uint32_t i = (uint32_t)i64;
  int64_t off = 0;
#pragma unroll
  for (int l = TN_MAX_MAP_LEVELS - 1; l >= 0; --l) {
    if (l < m.n) {
      const uint32_t q = (__umulhi(i, m.magic[l]) + i) >> m.shift[l];
      off += (int64_t)(i - q * m.ext[l]) * m.stride[l];
      i = q;
    }
  }
  return off;
}

// ----------------------------------------------------------------------------
// FP64 / FP32 kernel: BM x BN block tile, BK deep, 256 threads = 2x2 waves.
template <typename T, int BM, int BN, int BK>
__global__ void __launch_bounds__(256) gemm_mfma_16x16x4(const double* __restrict__ A,
                                                         const double* __restrict__ B,
                                                         double* __restrict__ C,
                                                         const GroupDev* __restrict__ groups,
                                                         const TileDev* __restrict__ tiles) {
  static_assert(256 % BK == 0 && 256 % BN == 0, "tile shape");
  constexpr int WM = BM / 2, WN = BN / 2;          // per-wave tile
  constexpr int TM = WM / 16, TN = WN / 16;        // MFMA tiles per wave
  constexpr int PAD = (sizeof(T) == 8) ? 1 : 2;
  constexpr int A_PER = BM * BK / 256;             // A elements per thread per K step
  constexpr int B_PER = BN * BK / 256;
  constexpr int A_RSTEP = 256 / BK;                // A rows covered per pass
  constexpr int B_KSTEP = 256 / BN;                // B k-rows covered per pass
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];

  const TileDev tile = tiles[blockIdx.x];
  if (tile.group < 0) return;
  const GroupDev& g = groups[tile.group];
  const int64_t M = g.M, N = g.N, K = g.K;
  const int64_t m0 = (int64_t)tile.tm * BM, n0 = (int64_t)tile.tn * BN;
  const double* __restrict__ Ab = A + g.a_off;
  const double* __restrict__ Bb = B + g.b_off;
  double* __restrict__ Cb = C + g.c_off;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;

  // fixed per-thread offsets
  const int a_kk = tid % BK;
  const int a_mm0 = tid / BK;
  int64_t a_roff[A_PER];
#pragma unroll
  for (int i = 0; i < A_PER; ++i) {
    const int64_t m = m0 + a_mm0 + i * A_RSTEP;
    a_roff[i] = (m < M) ? map_off(g.a_row, m, g.lda) : -1;
  }
  const int b_nn = tid % BN;
  const int b_kk0 = tid / BN;
  const int64_t b_coff = (n0 + b_nn < N) ? map_off(g.b_col, n0 + b_nn, 1) : -1;

  typedef typename std::conditional<sizeof(T) == 8, d4, f4>::type acc_t;
  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  T ra[A_PER], rb[B_PER];
  auto load_tiles = [&](int64_t k0) {
    const int64_t ka = k0 + a_kk;
    const int64_t a_coff = (ka < K) ? map_off(g.a_col, ka, 1) : -1;
#pragma unroll
    for (int i = 0; i < A_PER; ++i)
      ra[i] = (a_roff[i] >= 0 && a_coff >= 0) ? (T)Ab[a_roff[i] + a_coff] : (T)0;
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int64_t kb = k0 + b_kk0 + i * B_KSTEP;
      const int64_t b_roff = (kb < K) ? map_off(g.b_row, kb, g.ldb) : -1;
      rb[i] = (b_roff >= 0 && b_coff >= 0) ? (T)Bb[b_roff + b_coff] : (T)0;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) As[a_kk][a_mm0 + i * A_RSTEP] = ra[i];
#pragma unroll
    for (int i = 0; i < B_PER; ++i) Bs[b_kk0 + i * B_KSTEP][b_nn] = rb[i];
  };

  const int64_t nk = (K + BK - 1) / BK;
  load_tiles(0);
  for (int64_t kt = 0; kt < nk; ++kt) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (kt + 1 < nk) load_tiles((kt + 1) * BK);  // next global load overlaps the MFMAs
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      T a[TM], b[TN];
      const int kr = ks + (lane >> 4);
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[kr][wm0 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[kr][wn0 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          if constexpr (sizeof(T) == 8)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
  }
  // write back
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gn = n0 + wn0 + j * 16 + (lane & 15);
    if (gn >= N) continue;
    const int64_t coff = map_off(g.c_col, gn, 1);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // accumulator row: f64 16x16x4 -> (l/16) + 4v ; f32 16x16x4 -> 4(l/16) + v
        const int row = (sizeof(T) == 8) ? ((lane >> 4) + 4 * v) : (4 * (lane >> 4) + v);
        const int64_t gm = m0 + wm0 + i * 16 + row;
        if (gm < M) Cb[map_off(g.c_row, gm, g.ldc) + coff] = (double)acc[i][j][v];
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
  const int64_t M = g.M, N = g.N, K = g.K;
  const int64_t m0 = (int64_t)tile.tm * BM, n0 = (int64_t)tile.tn * BN;
  const double* __restrict__ Ab = A + g.a_off;
  const double* __restrict__ Bb = B + g.b_off;
  double* __restrict__ Cb = C + g.c_off;

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
    a_roff[i] = (m < M) ? map_off(g.a_row, m, g.lda) : -1;
  }
  const int b_nn = tid % BN;
  const int b_kk0 = tid / BN;
  const int64_t b_coff = (n0 + b_nn < N) ? map_off(g.b_col, n0 + b_nn, 1) : -1;

  f4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0, 0, 0, 0};

  float ra[A_PER], rb[B_PER];
  auto load_tiles = [&](int64_t k0) {
    const int64_t ka = k0 + a_kk;
    const int64_t a_coff = (ka < K) ? map_off(g.a_col, ka, 1) : -1;
#pragma unroll
    for (int i = 0; i < A_PER; ++i)
      ra[i] = (a_roff[i] >= 0 && a_coff >= 0) ? (float)Ab[a_roff[i] + a_coff] : 0.0f;
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int64_t kb = k0 + b_kk0 + i * B_KSTEP;
      const int64_t b_roff = (kb < K) ? map_off(g.b_row, kb, g.ldb) : -1;
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
    const int64_t coff = map_off(g.c_col, gn, 1);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gm = m0 + wm0 + i * 16 + 4 * (lane >> 4) + v;
        if (gm < M) Cb[map_off(g.c_row, gm, g.ldc) + coff] = (double)acc[i][j][v];
      }
  }
}

// ----------------------------------------------------------------------------
// Tile configurations (autotuner ids).
struct TileCfg {
  int bm, bn;
};
constexpr TileCfg kCfgs[] = {{32, 32}, {64, 64}, {128, 64}, {64, 128}, {128, 128}};
constexpr int kNumCfg = sizeof(kCfgs) / sizeof(kCfgs[0]);

typedef void (*KernFn)(const double*, const double*, double*, const GroupDev*, const TileDev*);

KernFn pick_kernel(int mode, int cfg) {
  switch (mode) {
    case TN_GEMM_F64:
      switch (cfg) {
        case 0: return gemm_mfma_16x16x4<double, 32, 32, 128>;
        case 1: return gemm_mfma_16x16x4<double, 64, 64, 64>;
        case 2: return gemm_mfma_16x16x4<double, 128, 64, 32>;
        case 3: return gemm_mfma_16x16x4<double, 64, 128, 32>;
        case 4: return gemm_mfma_16x16x4<double, 128, 128, 16>;
      }
      break;
    case TN_GEMM_F32:
      switch (cfg) {
        case 0: return gemm_mfma_16x16x4<float, 32, 32, 128>;
        case 1: return gemm_mfma_16x16x4<float, 64, 64, 64>;
        case 2: return gemm_mfma_16x16x4<float, 128, 64, 32>;
        case 3: return gemm_mfma_16x16x4<float, 64, 128, 32>;
        case 4: return gemm_mfma_16x16x4<float, 128, 128, 16>;
      }
      break;
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

// Host: 32-bit division magic for divisor d >= 1 (exact for dividends < 2^31).
void div_magic(uint32_t d, uint32_t* magic, uint32_t* shift) {
  uint32_t l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  *shift = l;
  *magic = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
}

bool make_map(const tn_index_map& in, int64_t range, MapDev* out) {
  std::memset(out, 0, sizeof(*out));
  if (in.n == 0 || range == 0) return true;
  if (in.n < 0 || in.n > TN_MAX_MAP_LEVELS) return false;
  int64_t prod = 1;
  for (int l = 0; l < in.n; ++l) {
    if (in.ext[l] < 1 || in.ext[l] >= (int64_t(1) << 31) || in.stride[l] < 0) return false;
    prod *= in.ext[l];
    if (prod >= (int64_t(1) << 31)) return false;
  }
  if (prod != range) return false;
  out->n = in.n;
  for (int l = 0; l < in.n; ++l) {
    out->ext[l] = (uint32_t)in.ext[l];
    div_magic(out->ext[l], &out->magic[l], &out->shift[l]);
    out->stride[l] = in.stride[l];
  }
  return true;
}

}  // namespace

struct tn_gemm_plan_s {
  int mode = 0;
  int cfg = -1;
  std::vector<GroupDev> groups;
  GroupDev* d_groups = nullptr;
  TileDev* d_tiles = nullptr;
  int64_t ntiles = 0;     // real tiles
  int64_t nslots = 0;     // launched blocks (with XCD padding)
  double flops = 0;
};

static int build_tiles(tn_gemm_plan_s* P, int cfg) {
  const int bm = kCfgs[cfg].bm, bn = kCfgs[cfg].bn;
  std::vector<TileDev> nat;
  for (size_t gi = 0; gi < P->groups.size(); ++gi) {
    const GroupDev& g = P->groups[gi];
    if (g.M == 0 || g.N == 0 || g.K == 0) continue;
    int64_t tm = (g.M + bm - 1) / bm, tn_ = (g.N + bn - 1) / bn;
    TN_REQUIRE(tm < (1 << 30) && tn_ < (1 << 30), TN_EUNSUPP, "GEMM too large");
    for (int64_t b = 0; b < tn_; ++b)
      for (int64_t a = 0; a < tm; ++a) nat.push_back(TileDev{(int32_t)gi, (int32_t)a, (int32_t)b, 0});
  }
  const int64_t n = (int64_t)nat.size();
  // XCD-aware placement: natural tile i goes to XCD i / per, slot i % per,
  // block id = slot * 8 + xcd (the hardware deals block ids round-robin).
  const int64_t kX = 8;
  const int64_t per = (n + kX - 1) / kX;
  const int64_t slots = per * kX;
  std::vector<TileDev> lay(slots, TileDev{-1, 0, 0, 0});
  for (int64_t i = 0; i < n; ++i) lay[(i % per) * kX + i / per] = nat[i];
  TN_REQUIRE(slots < (int64_t(1) << 31), TN_EUNSUPP, "too many GEMM tiles");
  if (P->d_tiles) tn::pool_free(P->d_tiles);
  P->d_tiles = nullptr;
  if (slots > 0) {
    int rc = tn::pool_alloc((void**)&P->d_tiles, sizeof(TileDev) * slots);
    if (rc != TN_OK) return rc;
    rc = tn_memcpy_h2d(P->d_tiles, lay.data(), sizeof(TileDev) * slots, nullptr);
    if (rc != TN_OK) return rc;
  }
  P->ntiles = n;
  P->nslots = slots;
  P->cfg = cfg;
  return TN_OK;
}

// Default tile: the largest tile that still gives every CU at least one
// tile; otherwise the smallest tile.
static int default_cfg(const tn_gemm_plan_s* P) {
  const int order[] = {4, 2, 1, 0};
  const int64_t want = tn::cu_count();
  for (int c : order) {
    int64_t t = 0;
    for (const auto& g : P->groups)
      t += ((g.M + kCfgs[c].bm - 1) / kCfgs[c].bm) * ((g.N + kCfgs[c].bn - 1) / kCfgs[c].bn);
    if (t >= want) return c;
  }
  return 0;
}

static void destroy_plan(tn_gemm_plan_s* P) {
  tn::pool_free(P->d_groups);
  tn::pool_free(P->d_tiles);
  delete P;
}

extern "C" {

int tn_gemm_plan_create_ex(tn_gemm_plan* plan, int mode, int64_t ngroups, const tn_gemm_desc_ex* d) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(plan != nullptr, TN_EINVAL, "tn_gemm_plan_create: NULL out pointer");
  TN_REQUIRE(mode >= TN_GEMM_F64 && mode <= TN_GEMM_BF16X6, TN_EINVAL, "unknown GEMM mode %d", mode);
  TN_REQUIRE(ngroups >= 0 && (ngroups == 0 || d), TN_EINVAL, "tn_gemm_plan_create: NULL descs");
  auto* P = new tn_gemm_plan_s();
  P->mode = mode;
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
    const bool ld_ok = (x.a_row.n || x.M == 0 || x.K == 0 || x.lda >= x.K) &&
                       (x.b_row.n || x.K == 0 || x.N == 0 || x.ldb >= x.N) &&
                       (x.c_row.n || x.M == 0 || x.N == 0 || x.ldc >= x.N);
    const bool range_ok = x.M < (int64_t(1) << 31) && x.N < (int64_t(1) << 31) && x.K < (int64_t(1) << 31);
    const bool maps_ok = make_map(x.a_row, x.M, &G.a_row) && make_map(x.a_col, x.K, &G.a_col) &&
                         make_map(x.b_row, x.K, &G.b_row) && make_map(x.b_col, x.N, &G.b_col) &&
                         make_map(x.c_row, x.M, &G.c_row) && make_map(x.c_col, x.N, &G.c_col);
    if (x.M < 0 || x.N < 0 || x.K < 0 || x.a_off < 0 || x.b_off < 0 || x.c_off < 0 || !ld_ok || !range_ok ||
        !maps_ok) {
      delete P;
      tn::set_error("GEMM group %lld has an invalid shape, leading dimension or index map", (long long)i);
      return TN_EINVAL;
    }
    P->groups.push_back(G);
    P->flops += 2.0 * (double)x.M * (double)x.N * (double)x.K;
  }
  int rc = TN_OK;
  if (!P->groups.empty()) {
    rc = tn::pool_alloc((void**)&P->d_groups, sizeof(GroupDev) * P->groups.size());
    if (rc == TN_OK)
      rc = tn_memcpy_h2d(P->d_groups, P->groups.data(), sizeof(GroupDev) * P->groups.size(), nullptr);
  }
  if (rc == TN_OK) rc = build_tiles(P, default_cfg(P));
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
  TN_REQUIRE(tile_id >= 0 && tile_id < kNumCfg, TN_EINVAL, "tile id %d outside [0,%d)", tile_id, kNumCfg);
  if (tile_id == P->cfg) return TN_OK;
  return build_tiles(P, tile_id);
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
  int best = P->cfg;
  float best_t = 1e30f;
  int rc = TN_OK;
  for (int c = 0; c < kNumCfg && rc == TN_OK; ++c) {
    rc = build_tiles(P, c);
    if (rc != TN_OK) break;
    rc = tn_gemm_exec(P, A, B, C, stream);   // warm-up
    if (rc != TN_OK) break;
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps && rc == TN_OK; ++r) rc = tn_gemm_exec(P, A, B, C, stream);
    (void)hipEventRecord(e1, s);
    if (rc != TN_OK) break;
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= reps;
    if (ms < best_t) {
      best_t = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != TN_OK) return rc;
  if (best_ms) *best_ms = best_t;
  return build_tiles(P, best);
}

int tn_gemm_plan_info(tn_gemm_plan P, int64_t* tiles, double* flops, int* tile_id) {
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_plan_info: NULL plan");
  if (tiles) *tiles = P->ntiles;
  if (flops) *flops = P->flops;
  if (tile_id) *tile_id = P->cfg;
  return TN_OK;
}

// Fill kernel for sectors with K == 0 (dgemm with beta = 0 writes zeros).
__global__ void zero_c_kernel(double* __restrict__ C, const GroupDev* __restrict__ gp) {
  const GroupDev& g = *gp;
  const int64_t total = g.M * g.N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t m = e / g.N, n = e % g.N;
    C[g.c_off + map_off(g.c_row, m, g.ldc) + map_off(g.c_col, n, 1)] = 0.0;
  }
}

int tn_gemm_exec(tn_gemm_plan P, const double* A, const double* B, double* C, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_exec: NULL plan");
  hipStream_t s = tn::stream_of(stream);
  for (size_t gi = 0; gi < P->groups.size(); ++gi) {
    const GroupDev& g = P->groups[gi];
    if (g.K == 0 && g.M > 0 && g.N > 0) {
      TN_REQUIRE(C != nullptr, TN_EINVAL, "tn_gemm_exec: NULL C");
      int64_t blocks = std::min<int64_t>((g.M * g.N + 255) / 256, 1024);
      hipLaunchKernelGGL(zero_c_kernel, dim3((unsigned)blocks), dim3(256), 0, s, C, P->d_groups + gi);
      TN_LAUNCH_CHECK();
    }
  }
  if (P->nslots == 0) return TN_OK;
  TN_REQUIRE(A && B && C, TN_EINVAL, "tn_gemm_exec: NULL operand");
  KernFn k = pick_kernel(P->mode, P->cfg);
  TN_REQUIRE(k != nullptr, TN_EINVAL, "no kernel for mode %d tile %d", P->mode, P->cfg);
  hipLaunchKernelGGL(k, dim3((unsigned)P->nslots), dim3(256), 0, s, A, B, C, P->d_groups, P->d_tiles);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

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
