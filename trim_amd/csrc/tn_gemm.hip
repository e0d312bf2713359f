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
// Arithmetic modes (storage is always FP64, row-major):
//   F64     v_mfma_f64_16x16x4f64, FP64 accumulate            (plain drop-in)
//   F32     v_mfma_f32_16x16x4f32 on inputs rounded to FP32, FP32 accumulate
//   BF16X3  inputs split x = h + m (+ rest), each BF16; products
//           hh + hm + mh on v_mfma_f32_16x16x32_bf16, FP32 accumulate
//   BF16X6  x = h + m + l; products hh+hm+mh+hl+mm+lh, FP32 accumulate
// Every output element is computed by exactly one workgroup, accumulating
// over K in a fixed order: results do not depend on what else is in the
// group (no split-K, no atomics).
//
// MFMA operand layout used (16x16 output tile, wave of 64 lanes):
//   16x16x4 (f64/f32): lane l supplies A[l%16][l/16] and B[l/16][l%16];
//     accumulator element i of lane l is C[(l/16)+4i][l%16] for f64 and
//     C[4*(l/16)+i][l%16] for f32 (measured with a one-instruction probe).
//   16x16x32 bf16: lane l supplies A[l%16][8*(l/16)+i] and B[8*(l/16)+i][l%16]
//     for i = 0..7; same accumulator layout.
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

struct GroupDev {
  int64_t M, N, K;
  int64_t a_off, b_off, c_off;
  int64_t lda, ldb, ldc;
};

// ----------------------------------------------------------------------------
// FP64 / FP32 kernel: BM x BN block tile, BK = 16, 256 threads = 2x2 waves.
template <typename T, int BM, int BN>
__global__ void __launch_bounds__(256) gemm_mfma_16x16x4(const double* __restrict__ A,
                                                         const double* __restrict__ B,
                                                         double* __restrict__ C,
                                                         const GroupDev* __restrict__ groups,
                                                         const TileDev* __restrict__ tiles) {
  constexpr int BK = 16;
  constexpr int WM = BM / 2, WN = BN / 2;          // per-wave tile
  constexpr int TM = WM / 16, TN = WN / 16;        // MFMA tiles per wave
  constexpr int PAD = (sizeof(T) == 8) ? 1 : 2;
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];

  const TileDev tile = tiles[blockIdx.x];
  if (tile.group < 0) return;
  const GroupDev g = groups[tile.group];
  const int64_t m0 = (int64_t)tile.tm * BM, n0 = (int64_t)tile.tn * BN;
  const double* __restrict__ Ab = A + g.a_off;
  const double* __restrict__ Bb = B + g.b_off;
  double* __restrict__ Cb = C + g.c_off;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;

  typedef typename std::conditional<sizeof(T) == 8, d4, f4>::type acc_t;
  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  constexpr int A_PER = BM * BK / 256;  // elements per thread
  constexpr int B_PER = BN * BK / 256;
  T ra[A_PER], rb[B_PER];

  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      int e = tid + i * 256;
      int kk = e % BK, mm = e / BK;
      int64_t gm = m0 + mm, gk = k0 + kk;
      double v = (gm < g.M && gk < g.K) ? Ab[gm * g.lda + gk] : 0.0;
      ra[i] = (T)v;
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int e = tid + i * 256;
      int nn = e % BN, kk = e / BN;
      int64_t gk = k0 + kk, gn = n0 + nn;
      double v = (gk < g.K && gn < g.N) ? Bb[gk * g.ldb + gn] : 0.0;
      rb[i] = (T)v;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      int e = tid + i * 256;
      As[e % BK][e / BK] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int e = tid + i * 256;
      Bs[e / BN][e % BN] = rb[i];
    }
  };

  const int64_t nk = (g.K + BK - 1) / BK;
  load_tiles(0);
  for (int64_t kt = 0; kt < nk; ++kt) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (kt + 1 < nk) load_tiles((kt + 1) * BK);  // overlap next global load with MFMAs
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
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int64_t gn = n0 + wn0 + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // accumulator row: f64 16x16x4 -> (l/16) + 4v ; f32 16x16x4 -> 4(l/16) + v
        const int row = (sizeof(T) == 8) ? ((lane >> 4) + 4 * v) : (4 * (lane >> 4) + v);
        const int64_t gm = m0 + wm0 + i * 16 + row;
        if (gm < g.M && gn < g.N) Cb[gm * g.ldc + gn] = (double)acc[i][j][v];
      }
    }
}

// ----------------------------------------------------------------------------
// Split-BF16 kernel: BM x BN tile, BK = 32, pieces NP = 2 (BF16X3) or 3 (BF16X6).
__device__ inline void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  float r = x - (float)h;
  m = (__bf16)r;
  float r2 = r - (float)m;
  l = (__bf16)r2;
}

template <int BM, int BN, int NP>
__global__ void __launch_bounds__(256) gemm_split_bf16(const double* __restrict__ A,
                                                       const double* __restrict__ B,
                                                       double* __restrict__ C,
                                                       const GroupDev* __restrict__ groups,
                                                       const TileDev* __restrict__ tiles) {
  constexpr int BK = 32;
  constexpr int KP = BK + 8;                       // padded row (bf16 elements)
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  __shared__ __attribute__((aligned(16))) __bf16 As[NP][BM][KP];   // [piece][m][k]
  __shared__ __attribute__((aligned(16))) __bf16 Bs[NP][BN][KP];   // [piece][n][k]

  const TileDev tile = tiles[blockIdx.x];
  if (tile.group < 0) return;
  const GroupDev g = groups[tile.group];
  const int64_t m0 = (int64_t)tile.tm * BM, n0 = (int64_t)tile.tn * BN;
  const double* __restrict__ Ab = A + g.a_off;
  const double* __restrict__ Bb = B + g.b_off;
  double* __restrict__ Cb = C + g.c_off;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;

  f4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0, 0, 0, 0};

  constexpr int A_PER = BM * BK / 256;
  constexpr int B_PER = BN * BK / 256;
  float ra[A_PER], rb[B_PER];

  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      int e = tid + i * 256;
      int kk = e % BK, mm = e / BK;
      int64_t gm = m0 + mm, gk = k0 + kk;
      ra[i] = (gm < g.M && gk < g.K) ? (float)Ab[gm * g.lda + gk] : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int e = tid + i * 256;
      int nn = e % BN, kk = e / BN;
      int64_t gk = k0 + kk, gn = n0 + nn;
      rb[i] = (gk < g.K && gn < g.N) ? (float)Bb[gk * g.ldb + gn] : 0.0f;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      int e = tid + i * 256;
      __bf16 h, m, l;
      split3(ra[i], h, m, l);
      As[0][e / BK][e % BK] = h;
      As[1][e / BK][e % BK] = m;
      if constexpr (NP > 2) As[2][e / BK][e % BK] = l;
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      int e = tid + i * 256;
      __bf16 h, m, l;
      split3(rb[i], h, m, l);
      Bs[0][e % BN][e / BN] = h;
      Bs[1][e % BN][e / BN] = m;
      if constexpr (NP > 2) Bs[2][e % BN][e / BN] = l;
    }
  };

  const int64_t nk = (g.K + BK - 1) / BK;
  load_tiles(0);
  const int kq = 8 * (lane >> 4);
  for (int64_t kt = 0; kt < nk; ++kt) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (kt + 1 < nk) load_tiles((kt + 1) * BK);
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
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int64_t gn = n0 + wn0 + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gm = m0 + wm0 + i * 16 + 4 * (lane >> 4) + v;
        if (gm < g.M && gn < g.N) Cb[gm * g.ldc + gn] = (double)acc[i][j][v];
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
#define TN_CFG_CASES(MAKE)                       \
  switch (cfg) {                                 \
    case 0: return MAKE(32, 32);                 \
    case 1: return MAKE(64, 64);                 \
    case 2: return MAKE(128, 64);                \
    case 3: return MAKE(64, 128);                \
    case 4: return MAKE(128, 128);               \
  }
#define MK_F64(a, b) gemm_mfma_16x16x4<double, a, b>
#define MK_F32(a, b) gemm_mfma_16x16x4<float, a, b>
#define MK_B3(a, b) gemm_split_bf16<a, b, 2>
#define MK_B6(a, b) gemm_split_bf16<a, b, 3>
  switch (mode) {
    case TN_GEMM_F64: TN_CFG_CASES(MK_F64); break;
    case TN_GEMM_F32: TN_CFG_CASES(MK_F32); break;
    case TN_GEMM_BF16X3: TN_CFG_CASES(MK_B3); break;
    case TN_GEMM_BF16X6: TN_CFG_CASES(MK_B6); break;
  }
  return nullptr;
#undef MK_F64
#undef MK_F32
#undef MK_B3
#undef MK_B6
#undef TN_CFG_CASES
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
    if (g.M == 0 || g.N == 0) continue;
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
  for (int64_t i = 0; i < n; ++i) {
    int64_t x = i / std::max<int64_t>(per, 1), s = i % std::max<int64_t>(per, 1);
    lay[s * kX + x] = nat[i];
  }
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

// Default tile choice: the largest tile that still gives every CU at least
// one tile; otherwise the smallest tile.
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

extern "C" {

int tn_gemm_plan_create(tn_gemm_plan* plan, int mode, int64_t ngroups, const tn_gemm_desc* d) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(plan != nullptr, TN_EINVAL, "tn_gemm_plan_create: NULL out pointer");
  TN_REQUIRE(mode >= TN_GEMM_F64 && mode <= TN_GEMM_BF16X6, TN_EINVAL, "unknown GEMM mode %d",
             mode);
  TN_REQUIRE(ngroups >= 0 && (ngroups == 0 || d), TN_EINVAL, "tn_gemm_plan_create: NULL descs");
  auto* P = new tn_gemm_plan_s();
  P->mode = mode;
  for (int64_t i = 0; i < ngroups; ++i) {
    const tn_gemm_desc& x = d[i];
    if (x.M < 0 || x.N < 0 || x.K < 0 || (x.M > 0 && x.K > 0 && x.lda < x.K) ||
        (x.K > 0 && x.N > 0 && x.ldb < x.N) || (x.M > 0 && x.N > 0 && x.ldc < x.N) ||
        x.a_off < 0 || x.b_off < 0 || x.c_off < 0) {
      delete P;
      tn::set_error("GEMM group %lld has an invalid shape or leading dimension", (long long)i);
      return TN_EINVAL;
    }
    P->groups.push_back(GroupDev{x.M, x.N, x.K, x.a_off, x.b_off, x.c_off, x.lda, x.ldb, x.ldc});
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
    tn::pool_free(P->d_groups);
    tn::pool_free(P->d_tiles);
    delete P;
    return rc;
  }
  *plan = P;
  return TN_OK;
}

int tn_gemm_plan_set_tile(tn_gemm_plan P, int tile_id) {
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_plan_set_tile: NULL plan");
  TN_REQUIRE(tile_id >= 0 && tile_id < kNumCfg, TN_EINVAL, "tile id %d outside [0,%d)", tile_id,
             kNumCfg);
  if (tile_id == P->cfg) return TN_OK;
  return build_tiles(P, tile_id);
}

int tn_gemm_plan_info(tn_gemm_plan P, int64_t* tiles, double* flops, int* tile_id) {
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_plan_info: NULL plan");
  if (tiles) *tiles = P->ntiles;
  if (flops) *flops = P->flops;
  if (tile_id) *tile_id = P->cfg;
  return TN_OK;
}

int tn_gemm_exec(tn_gemm_plan P, const double* A, const double* B, double* C, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_gemm_exec: NULL plan");
  hipStream_t s = tn::stream_of(stream);
  // Sectors with K == 0 produce zeros (dgemm with beta = 0).
  for (const auto& g : P->groups)
    if (g.K == 0 && g.M > 0 && g.N > 0) {
      for (int64_t r = 0; r < g.M; ++r) {
        int rc = tn_zero(C + g.c_off + r * g.ldc, sizeof(double) * g.N, stream);
        if (rc != TN_OK) return rc;
      }
    }
  if (P->nslots == 0) return TN_OK;
  TN_REQUIRE(A && B && C, TN_EINVAL, "tn_gemm_exec: NULL operand");
  KernFn k = pick_kernel(P->mode, P->cfg);
  TN_REQUIRE(k != nullptr, TN_EINVAL, "no kernel for mode %d tile %d", P->mode, P->cfg);
  hipLaunchKernelGGL(k, dim3((unsigned)P->nslots), dim3(256), 0, s, A, B, C, P->d_groups,
                     P->d_tiles);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_gemm_plan_destroy(tn_gemm_plan P) {
  if (!P) return TN_OK;
  tn::pool_free(P->d_groups);
  tn::pool_free(P->d_tiles);
  delete P;
  return TN_OK;
}

int tn_matmul_f64(const double* A, const double* B, int64_t M, int64_t N, int64_t K, double* C,
                  int mode, void* stream) {
  tn_gemm_desc d{M, N, K, 0, 0, 0, K, N, N};
  tn_gemm_plan P = nullptr;
  int rc = tn_gemm_plan_create(&P, mode, 1, &d);
  if (rc != TN_OK) return rc;
  rc = tn_gemm_exec(P, A, B, C, stream);
  // the plan's device tables are read by the launch: keep them alive until it ends
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
