// Index-remap kernels: dense permute (reshapeElem), symmetric block permute.
//
// Reference behaviour being replaced:
//  * _reshapeElem (uni10/src/uni10/tools/lib/uni10_tools_gpu.cu:319-339): one
//    thread per element, r integer div/mod each, coalesced reads and
//    scattered writes, a leaked cudaMalloc'd stride table per call.
//  * the host odometer of UniTensorReal.cpp:601-646 and the per-Qidx host
//    sub-block loop of UniTensorReal.cpp:683-756 (symmetric tensors).
//
// Design (gfx950):
//  * The host planner drops unit axes and merges axes that stay adjacent,
//    then picks one of two kernels:
//      - row copy, when the innermost axis stays innermost: every row of the
//        merged innermost extent is a contiguous run in both tensors; 16-byte
//        loads/stores along it, one division chain per row;
//      - tiled transpose otherwise: a tile is (source-inner run) x
//        (destination-inner run), each grown over trailing axes to ~64
//        elements (the outermost grown axis is chunked).  A workgroup reads
//        the tile with runs contiguous in the source, stages it in LDS
//        (padded rows), and writes it with runs contiguous in the
//        destination.  Per-row offsets come from two small tables built once
//        per workgroup, so the element loops do no division.
//  * Block permute: the host turns the Qidx list into moves; moves are cut
//    into chunks of at most 4096 elements, one workgroup per chunk, all in a
//    single launch; threads walk the destination in order (coalesced
//    writes) and gather from the source sub-block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "tn_internal.h"

namespace {

constexpr int kMaxRank = 16;

// ----------------------------------------------------------------------------
// Row-copy kernel: rows of length L, contiguous in both source and destination.
// Source row r = mixed-radix digits over the outer source axes; destination
// row offset = sum digit_k * dstride_k.
struct RowCopyParams {
  int64_t L;          // contiguous run length (elements)
  int64_t rows;       // number of runs
  int nouter;         // number of outer axes
  int64_t odim[kMaxRank];     // outer axes extents (source order, outermost first)
  int64_t odst[kMaxRank];     // destination stride of each outer axis
};

template <bool kVec2>
__global__ void __launch_bounds__(256) row_copy_kernel(const double* __restrict__ src,
                                                       double* __restrict__ dst,
                                                       RowCopyParams p, int rows_per_block) {
  const int64_t row0 = (int64_t)blockIdx.x * rows_per_block;
  for (int rr = 0; rr < rows_per_block; ++rr) {
    int64_t row = row0 + rr;
    if (row >= p.rows) return;
    int64_t rem = row, doff = 0;
    for (int k = p.nouter - 1; k >= 0; --k) {
      int64_t d = p.odim[k];
      int64_t q = rem / d;
      doff += (rem - q * d) * p.odst[k];
      rem = q;
    }
    const double* s = src + row * p.L;
    double* t = dst + doff * 1;
    if (kVec2) {
      const int64_t n2 = p.L >> 1;
      const double2* s2 = reinterpret_cast<const double2*>(s);
      double2* t2 = reinterpret_cast<double2*>(t);
      for (int64_t i = threadIdx.x; i < n2; i += blockDim.x) t2[i] = s2[i];
    } else {
      for (int64_t i = threadIdx.x; i < p.L; i += blockDim.x) t[i] = s[i];
    }
  }
}

// ----------------------------------------------------------------------------
// Tiled transpose.  Tile axes: S group (source-inner, trailing source axes,
// outermost S axis chunked) and D group (destination-inner axes, outermost D
// axis chunked).  Remaining axes are "outer" (one per grid coordinate).
constexpr int kTileS = 64;
constexpr int kTileD = 64;

struct TileParams {
  // S group: axes listed innermost first.  Extent of S tile = nS.
  int ns_axes;
  int64_t s_dim[kMaxRank];      // extent of each S axis inside the tile (full or chunk)
  int64_t s_full[kMaxRank];     // full extent of the axis (for bounds)
  int64_t s_dst[kMaxRank];      // destination stride of that axis
  int nd_axes;
  int64_t d_dim[kMaxRank];
  int64_t d_full[kMaxRank];
  int64_t d_src[kMaxRank];      // source stride of that axis
  // chunked axes (the outermost S / D axis): source and destination strides
  int64_t s_chunk_src, s_chunk_dst, s_chunk_full, s_chunk_len, s_nchunks;
  int64_t d_chunk_src, d_chunk_dst, d_chunk_full, d_chunk_len, d_nchunks;
  int64_t s_inner;              // product of full inner S axes (source contiguous run unit)
  int64_t d_inner;              // product of full inner D axes
  int nS, nD;                   // tile extents (elements) in S and D
  int ws, wd;                   // lanes along S (load phase) / along D (store phase)
  // outer axes
  int nouter;
  int64_t o_dim[kMaxRank];
  int64_t o_src[kMaxRank];
  int64_t o_dst[kMaxRank];
};

typedef double dbl2 __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ dbl2 load2(const double* p) {
  const dbl2* q = reinterpret_cast<const dbl2*>(p);
  if constexpr (NT) return __builtin_nontemporal_load(q);
  return *q;
}
template <bool NT>
__device__ __forceinline__ void store2(double* p, dbl2 v) {
  dbl2* q = reinterpret_cast<dbl2*>(p);
  if constexpr (NT) __builtin_nontemporal_store(v, q);
  else *q = v;
}

// VL: rows are read as 16-byte pairs along s; VS: columns are written as
// 16-byte pairs along d (the planner checks that every pair is in range and
// aligned).  TD: the largest D extent of a tile (64, or 32 for the tiles of
// tensors larger than the caches: half the LDS, twice the workgroups per CU).
// NT: streaming (non-temporal) loads and stores, for tensors larger than the
// caches.  Workgroups run destination chunk fastest, so the workgroups in
// flight together write neighbouring runs of the same destination rows.
template <int TD, bool VL, bool VS, bool NT>
__global__ void __launch_bounds__(256) tiled_permute_kernel(const double* __restrict__ src,
                                                            double* __restrict__ dst,
                                                            TileParams p) {
  __shared__ double tile[TD][kTileS + 1];
  __shared__ int64_t srow_src[TD];       // source offset of D-row d (relative)
  __shared__ int64_t scol_dst[kTileS];   // destination offset of S-col s (relative)
  __shared__ int s_valid_sh, d_valid_sh;
  __shared__ int64_t base_src_sh, base_dst_sh;

  const int tid = threadIdx.x;
  if (tid == 0) {
    // decompose blockIdx.x into (outer digits..., s chunk, d chunk), d chunk fastest
    int64_t b = blockIdx.x;
    int64_t dc = b % p.d_nchunks;
    b /= p.d_nchunks;
    int64_t sc = b % p.s_nchunks;
    b /= p.s_nchunks;
    int64_t bs = 0, bd = 0;
    for (int k = p.nouter - 1; k >= 0; --k) {
      int64_t d = p.o_dim[k];
      int64_t q = b / d;
      int64_t c = b - q * d;
      bs += c * p.o_src[k];
      bd += c * p.o_dst[k];
      b = q;
    }
    int64_t s0 = sc * p.s_chunk_len, d0 = dc * p.d_chunk_len;
    bs += s0 * p.s_chunk_src + d0 * p.d_chunk_src;
    bd += s0 * p.s_chunk_dst + d0 * p.d_chunk_dst;
    base_src_sh = bs;
    base_dst_sh = bd;
    int64_t sv = min(p.s_chunk_len, p.s_chunk_full - s0) * p.s_inner;
    int64_t dv = min(p.d_chunk_len, p.d_chunk_full - d0) * p.d_inner;
    s_valid_sh = (int)sv;
    d_valid_sh = (int)dv;
  }
  // Offset tables: for S index s, its destination offset; for D index d, its
  // source offset.  Digits innermost first.
  if (tid < p.nS) {
    int64_t rem = tid, off = 0;
    for (int k = 0; k < p.ns_axes; ++k) {
      int64_t d = p.s_dim[k];
      int64_t q = rem / d;
      off += (rem - q * d) * p.s_dst[k];
      rem = q;
    }
    scol_dst[tid] = off;
  }
  if (tid >= 128 && tid - 128 < p.nD) {
    int t = tid - 128;
    int64_t rem = t, off = 0;
    for (int k = 0; k < p.nd_axes; ++k) {
      int64_t d = p.d_dim[k];
      int64_t q = rem / d;
      off += (rem - q * d) * p.d_src[k];
      rem = q;
    }
    srow_src[t] = off;
  }
  __syncthreads();
  const int sv = s_valid_sh, dv = d_valid_sh;
  const double* sbase = src + base_src_sh;
  double* dbase = dst + base_dst_sh;
  // load: lanes run along s (contiguous in source)
  if constexpr (VL) {
    const int ws = p.ws / 2;
    const int s = 2 * (tid % ws);
    for (int d = tid / ws; d < dv; d += 256 / ws) {
      if (s < sv) {
        const dbl2 v = load2<NT>(sbase + srow_src[d] + s);
        tile[d][s] = v.x;
        tile[d][s + 1] = v.y;
      }
    }
  } else {
    const int ws = p.ws;
    const int s = tid % ws;
    for (int d = tid / ws; d < dv; d += 256 / ws) {
      if (s < sv) tile[d][s] = sbase[srow_src[d] + s];
    }
  }
  __syncthreads();
  // store: lanes run along d (contiguous in destination)
  if constexpr (VS) {
    const int wd = p.wd / 2;
    const int d = 2 * (tid % wd);
    for (int s = tid / wd; s < sv; s += 256 / wd) {
      if (d < dv) store2<NT>(dbase + scol_dst[s] + d, dbl2{tile[d][s], tile[d + 1][s]});
    }
  } else {
    const int wd = p.wd;
    const int d = tid % wd;
    for (int s = tid / wd; s < sv; s += 256 / wd) {
      if (d < dv) dbase[scol_dst[s] + d] = tile[d][s];
    }
  }
}

// ----------------------------------------------------------------------------
struct Plan {
  enum Kind { kEmpty, kCopy, kRows, kTile } kind = kEmpty;
  int64_t total = 0;
  RowCopyParams rows{};
  bool vec2 = false;
  TileParams tile{};
  int64_t grid = 0;
  bool tile_vl = false, tile_vs = false;   // 16-byte pairs along s / d
  bool stream = false;                     // larger than the caches: 64 x 32 tiles, streaming accesses
};

// Tile of a tiled transpose for target run lengths tgt (source side: 64, 32
// or 16) and tgt_d (destination side, at most tgt).
static int build_tile(const std::vector<int64_t>& dims, const std::vector<int>& perm,
                      const std::vector<int64_t>& sstr, const std::vector<int64_t>& dstr, int tgt, int tgt_d,
                      TileParams& T, int64_t* grid_out) {
  const int r = (int)dims.size();
  std::memset(&T, 0, sizeof(T));
  std::vector<int> in_s(r, 0), in_d(r, 0);
  // S group: trailing source axes (innermost first), stop before the
  // destination-inner axis.
  const int dinner = perm[r - 1];
  int64_t prod = 1;
  int s_chunk_axis = -1;
  for (int i = r - 1; i >= 0; --i) {
    if (i == dinner) break;
    if (prod * dims[i] >= tgt) {
      s_chunk_axis = i;
      break;
    }
    in_s[i] = 1;
    T.s_dim[T.ns_axes] = dims[i];
    T.s_full[T.ns_axes] = dims[i];
    T.s_dst[T.ns_axes] = dstr[i];
    T.ns_axes++;
    prod *= dims[i];
  }
  T.s_inner = prod;
  if (s_chunk_axis >= 0) {
    int64_t len = std::max<int64_t>(1, tgt / prod);
    len = std::min(len, dims[s_chunk_axis]);
    in_s[s_chunk_axis] = 2;
    T.s_chunk_src = sstr[s_chunk_axis];
    T.s_chunk_dst = dstr[s_chunk_axis];
    T.s_chunk_full = dims[s_chunk_axis];
    T.s_chunk_len = len;
    T.s_nchunks = (dims[s_chunk_axis] + len - 1) / len;
    T.s_dim[T.ns_axes] = len;
    T.s_full[T.ns_axes] = dims[s_chunk_axis];
    T.s_dst[T.ns_axes] = dstr[s_chunk_axis];
    T.ns_axes++;
    T.nS = (int)(prod * len);
  } else {
    // every trailing axis up to the destination-inner one fits: treat the
    // outermost included S axis as a single full chunk.
    T.s_chunk_src = 0;
    T.s_chunk_dst = 0;
    T.s_chunk_full = 1;
    T.s_chunk_len = 1;
    T.s_nchunks = 1;
    T.nS = (int)prod;
  }
  // D group: trailing output axes (innermost first), not in S.
  prod = 1;
  int d_chunk_axis = -1;
  for (int j = r - 1; j >= 0; --j) {
    int i = perm[j];
    if (in_s[i]) break;
    if (prod * dims[i] >= tgt_d) {
      d_chunk_axis = i;
      break;
    }
    in_d[i] = 1;
    T.d_dim[T.nd_axes] = dims[i];
    T.d_full[T.nd_axes] = dims[i];
    T.d_src[T.nd_axes] = sstr[i];
    T.nd_axes++;
    prod *= dims[i];
  }
  T.d_inner = prod;
  if (d_chunk_axis >= 0) {
    int64_t len = std::max<int64_t>(1, tgt_d / prod);
    len = std::min(len, dims[d_chunk_axis]);
    in_d[d_chunk_axis] = 2;
    T.d_chunk_src = sstr[d_chunk_axis];
    T.d_chunk_dst = dstr[d_chunk_axis];
    T.d_chunk_full = dims[d_chunk_axis];
    T.d_chunk_len = len;
    T.d_nchunks = (dims[d_chunk_axis] + len - 1) / len;
    T.d_dim[T.nd_axes] = len;
    T.d_full[T.nd_axes] = dims[d_chunk_axis];
    T.d_src[T.nd_axes] = sstr[d_chunk_axis];
    T.nd_axes++;
    T.nD = (int)(prod * len);
  } else {
    T.d_chunk_src = 0;
    T.d_chunk_dst = 0;
    T.d_chunk_full = 1;
    T.d_chunk_len = 1;
    T.d_nchunks = 1;
    T.nD = (int)prod;
  }
  // Outer axes (source order, outermost first).
  for (int i = 0; i < r; ++i)
    if (!in_s[i] && !in_d[i]) {
      T.o_dim[T.nouter] = dims[i];
      T.o_src[T.nouter] = sstr[i];
      T.o_dst[T.nouter] = dstr[i];
      T.nouter++;
    }
  int64_t grid = T.s_nchunks * T.d_nchunks;
  for (int k = 0; k < T.nouter; ++k) grid *= T.o_dim[k];
  *grid_out = grid;
  T.ws = tgt;
  T.wd = tgt_d;
  TN_REQUIRE(T.nS >= 1 && T.nS <= tgt && T.nD >= 1 && T.nD <= tgt_d, TN_EUNSUPP,
             "permute planner produced tile %d x %d", T.nD, T.nS);
  TN_REQUIRE(grid < (int64_t(1) << 31), TN_EUNSUPP, "permute grid too large (%lld)",
             (long long)grid);
  return TN_OK;
}

int make_plan(int rank, const int64_t* dims_in, const int* perm_in, Plan* out) {
  TN_REQUIRE(rank >= 0 && rank <= kMaxRank, TN_EUNSUPP, "permute: rank %d outside [0,%d]", rank,
             kMaxRank);
  int64_t total = 1;
  std::vector<int> seen(rank, 0);
  for (int i = 0; i < rank; ++i) {
    TN_REQUIRE(dims_in[i] >= 0, TN_EINVAL, "permute: negative extent on axis %d", i);
    total *= dims_in[i];
    TN_REQUIRE(perm_in[i] >= 0 && perm_in[i] < rank && !seen[perm_in[i]], TN_EINVAL,
               "permute: perm is not a permutation of 0..%d", rank - 1);
    seen[perm_in[i]] = 1;
  }
  out->total = total;
  if (total == 0) {
    out->kind = Plan::kEmpty;
    return TN_OK;
  }
  // Drop unit axes.
  std::vector<int> keep_map(rank, -1);
  std::vector<int64_t> dims;
  for (int i = 0; i < rank; ++i)
    if (dims_in[i] != 1) {
      keep_map[i] = (int)dims.size();
      dims.push_back(dims_in[i]);
    }
  std::vector<int> perm;
  for (int j = 0; j < rank; ++j)
    if (keep_map[perm_in[j]] >= 0) perm.push_back(keep_map[perm_in[j]]);
  // Merge source axes i, i+1 that are adjacent (in order) in the output.
  bool merged = true;
  while (merged && perm.size() > 1) {
    merged = false;
    for (size_t j = 0; j + 1 < perm.size(); ++j) {
      int a = perm[j];
      if (perm[j + 1] == a + 1) {
        dims[a] *= dims[a + 1];
        dims.erase(dims.begin() + a + 1);
        perm.erase(perm.begin() + j + 1);
        for (auto& q : perm)
          if (q > a) --q;
        merged = true;
        break;
      }
    }
  }
  const int r = (int)dims.size();
  if (r <= 1) {
    out->kind = Plan::kCopy;
    return TN_OK;
  }
  std::vector<int64_t> sstr(r), ostr(r), dstr(r);
  sstr[r - 1] = 1;
  for (int i = r - 1; i > 0; --i) sstr[i - 1] = sstr[i] * dims[i];
  ostr[r - 1] = 1;
  for (int j = r - 1; j > 0; --j) ostr[j - 1] = ostr[j] * dims[perm[j]];
  for (int j = 0; j < r; ++j) dstr[perm[j]] = ostr[j];

  if (perm[r - 1] == r - 1) {
    out->kind = Plan::kRows;
    RowCopyParams& P = out->rows;
    P.L = dims[r - 1];
    P.rows = total / P.L;
    P.nouter = r - 1;
    for (int i = 0; i < r - 1; ++i) {
      P.odim[i] = dims[i];
      P.odst[i] = dstr[i];
    }
    out->vec2 = (P.L % 2 == 0);
    return TN_OK;
  }
  // Tiled transpose: the largest tile that still gives every CU two
  // workgroups; small tensors get small tiles.  A tensor larger than the
  // caches takes 64 x 32 tiles (512-byte source runs, 256-byte destination
  // runs) with streaming accesses: measured 5.8 TB/s at (4096,2,2,4096)
  // against 5.1-5.3 TB/s for cached 64 x 64 tiles.
  out->kind = Plan::kTile;
  const int64_t want = 2 * (int64_t)tn::cu_count();
  out->stream = (size_t)total * sizeof(double) >= tn::kStreamingBytes;
  for (int tgt : {64, 32, 16}) {
    int64_t grid = 0;
    int rc = build_tile(dims, perm, sstr, dstr, tgt, (out->stream && tgt == 64) ? 32 : tgt, out->tile, &grid);
    if (rc != TN_OK) return rc;
    out->grid = grid;
    if (grid >= want) break;
  }
  out->stream = out->stream && out->tile.wd <= 32;
  // Pairs along s stay inside a row when the S group's full extent and every
  // chunk's extent are even (then each row start and valid length is even);
  // likewise along d.
  const TileParams& T = out->tile;
  out->tile_vl = (T.s_inner * T.s_chunk_full) % 2 == 0 && (T.s_inner * T.s_chunk_len) % 2 == 0;
  out->tile_vs = (T.d_inner * T.d_chunk_full) % 2 == 0 && (T.d_inner * T.d_chunk_len) % 2 == 0;
  return TN_OK;
}

int run_plan(const Plan& P, const double* src, double* dst, hipStream_t s) {
  switch (P.kind) {
    case Plan::kEmpty:
      return TN_OK;
    case Plan::kCopy:
      return tn::copy_d2d(dst, src, P.total * sizeof(double), s);
    case Plan::kRows: {
      const RowCopyParams& R = P.rows;
      bool vec = P.vec2 && (((uintptr_t)src | (uintptr_t)dst) & 15) == 0;
      // rows per block: aim at >= 2048 elements per block
      int rpb = (int)std::max<int64_t>(1, std::min<int64_t>(64, 2048 / std::max<int64_t>(1, R.L)));
      // keep at least two workgroups per CU when the tensor allows it
      while (rpb > 1 && (R.rows + rpb - 1) / rpb < 2 * (int64_t)tn::cu_count()) rpb /= 2;
      int64_t grid = (R.rows + rpb - 1) / rpb;
      TN_REQUIRE(grid < (int64_t(1) << 31), TN_EUNSUPP, "row-copy grid too large");
      if (vec) {
        // destination row offsets are multiples of L (even) -> 16-byte aligned
        hipLaunchKernelGGL(row_copy_kernel<true>, dim3((unsigned)grid), dim3(256), 0, s, src, dst,
                           R, rpb);
      } else {
        hipLaunchKernelGGL(row_copy_kernel<false>, dim3((unsigned)grid), dim3(256), 0, s, src, dst,
                           R, rpb);
      }
      TN_LAUNCH_CHECK();
      return TN_OK;
    }
    case Plan::kTile: {
      const bool aligned = (((uintptr_t)src | (uintptr_t)dst) & 15) == 0;
      const bool vl = P.tile_vl && aligned, vs = P.tile_vs && aligned;
      // streaming kernels need both pair paths (16-byte accesses throughout)
      auto k = (P.stream && vl && vs) ? tiled_permute_kernel<32, true, true, true>
               : vl ? (vs ? tiled_permute_kernel<kTileD, true, true, false> : tiled_permute_kernel<kTileD, true, false, false>)
                    : (vs ? tiled_permute_kernel<kTileD, false, true, false> : tiled_permute_kernel<kTileD, false, false, false>);
      hipLaunchKernelGGL(k, dim3((unsigned)P.grid), dim3(256), 0, s, src, dst, P.tile);
      TN_LAUNCH_CHECK();
      return TN_OK;
    }
  }
  return TN_EINVAL;
}

// ----------------------------------------------------------------------------
// Block permute.
struct MoveDev {
  int64_t src_off, bin_c, sbin_c;
  int64_t dst_off, bot_c, sbot_c;
  int64_t total;
  double sign;
  int64_t odim[kMaxRank];   // destination-order extents (sub-block, per output bond)
  int64_t osrc[kMaxRank];   // source linear stride of each output-order axis
};

struct ChunkDev {
  int64_t move;
  int64_t start;
  int64_t count;
};

__global__ void __launch_bounds__(256) block_permute_kernel(const double* __restrict__ src,
                                                            double* __restrict__ dst,
                                                            const MoveDev* __restrict__ moves,
                                                            const ChunkDev* __restrict__ chunks,
                                                            int rank) {
  const ChunkDev c = chunks[blockIdx.x];
  const MoveDev& m = moves[c.move];
  const double sign = m.sign;
  for (int64_t t = threadIdx.x; t < c.count; t += blockDim.x) {
    int64_t f = c.start + t;          // destination linear index within the sub-block
    int64_t rem = f, e = 0;
    for (int k = rank - 1; k >= 0; --k) {
      int64_t d = m.odim[k];
      int64_t q = rem / d;
      e += (rem - q * d) * m.osrc[k];
      rem = q;
    }
    int64_t sr = e / m.sbin_c, sc = e - sr * m.sbin_c;
    int64_t dr = f / m.sbot_c, dc = f - dr * m.sbot_c;
    dst[m.dst_off + dr * m.bot_c + dc] = sign * src[m.src_off + sr * m.bin_c + sc];
  }
}

// In-place scaling of rectangular sub-blocks (the swap-gate signs of
// addGate): chunk c covers elements [start, start + count) of rectangle
// `rect`, row-major over its cols.
struct RectDev {
  int64_t off, rows, cols, pitch;
};

__global__ void __launch_bounds__(256) rect_scale_kernel(double* __restrict__ base,
                                                         const RectDev* __restrict__ rects,
                                                         const ChunkDev* __restrict__ chunks, double a) {
  const ChunkDev c = chunks[blockIdx.x];
  const RectDev r = rects[c.move];
  for (int64_t t = threadIdx.x; t < c.count; t += blockDim.x) {
    const int64_t f = c.start + t;
    const int64_t row = f / r.cols, col = f - row * r.cols;
    double* p = base + r.off + row * r.pitch + col;
    *p = a * *p;
  }
}

}  // namespace

struct tn_blockperm_plan_s {
  int rank = 0;
  int64_t nmoves = 0, nchunks = 0, dst_elems = 0, covered = 0;
  MoveDev* d_moves = nullptr;
  ChunkDev* d_chunks = nullptr;
};

extern "C" {

int tn_permute_dense_f64(const double* src, double* dst, int rank, const int64_t* dims,
                         const int* perm, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(rank == 0 || (dims && perm), TN_EINVAL, "tn_permute_dense_f64: NULL dims/perm");
  Plan P;
  int rc = make_plan(rank, dims, perm, &P);
  if (rc != TN_OK) return rc;
  if (rank == 0) P.kind = Plan::kCopy, P.total = 1;
  if (P.kind != Plan::kEmpty) TN_REQUIRE(src && dst, TN_EINVAL, "tn_permute_dense_f64: NULL buffer");
  TN_REQUIRE(src != dst || P.kind == Plan::kEmpty, TN_EINVAL,
             "tn_permute_dense_f64: in-place permute is not supported");
  return run_plan(P, src, dst, tn::stream_of(stream));
}

int tn_reshape_elem_f64(const double* old_elem, int bond_num, size_t elem_num,
                        const size_t* offset, double* new_elem, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(bond_num >= 1 && bond_num <= kMaxRank && offset, TN_EINVAL,
             "tn_reshape_elem_f64: bad bond count or NULL offset table");
  // offset[0..r): source strides -> extents; offset[r+i]: destination stride
  // of source axis i -> the output order is the source axes sorted by
  // decreasing destination stride.
  const int r = bond_num;
  std::vector<int64_t> dims(r);
  for (int i = 0; i < r; ++i) {
    size_t outer = (i == 0) ? elem_num : offset[i - 1];
    TN_REQUIRE(offset[i] > 0 && outer % offset[i] == 0, TN_EINVAL,
               "tn_reshape_elem_f64: offset table is not a row-major stride list");
    dims[i] = (int64_t)(outer / offset[i]);
  }
  std::vector<int> perm(r);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
    if (offset[r + a] != offset[r + b]) return offset[r + a] > offset[r + b];
    return dims[a] > dims[b];  // unit axes share strides; order is immaterial
  });
  return tn_permute_dense_f64(old_elem, new_elem, r, dims.data(), perm.data(), stream);
}

int tn_transpose_f64(const double* A, int64_t M, int64_t N, double* AT, void* stream) {
  int64_t dims[2] = {M, N};
  int perm[2] = {1, 0};
  return tn_permute_dense_f64(A, AT, 2, dims, perm, stream);
}

int tn_blockperm_plan_create(tn_blockperm_plan* plan, int rank, int64_t nmoves,
                             const int64_t* meta, const int64_t* sub_dims,
                             const int64_t* bot_acc, const double* sign, int64_t dst_elems) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(plan != nullptr, TN_EINVAL, "tn_blockperm_plan_create: NULL out pointer");
  TN_REQUIRE(rank >= 1 && rank <= kMaxRank, TN_EUNSUPP, "block permute rank %d unsupported", rank);
  TN_REQUIRE(nmoves >= 0 && (nmoves == 0 || (meta && sub_dims && bot_acc)), TN_EINVAL,
             "tn_blockperm_plan_create: NULL move arrays");
  std::vector<MoveDev> mv(nmoves);
  std::vector<ChunkDev> ch;
  int64_t covered = 0;
  const int64_t kChunk = 4096;
  for (int64_t m = 0; m < nmoves; ++m) {
    const int64_t* mm = meta + 6 * m;
    const int64_t* sd = sub_dims + rank * m;
    const int64_t* acc = bot_acc + rank * m;
    MoveDev& D = mv[m];
    std::memset(&D, 0, sizeof(D));
    D.src_off = mm[0];
    D.bin_c = mm[1];
    D.sbin_c = mm[2];
    D.dst_off = mm[3];
    D.bot_c = mm[4];
    D.sbot_c = mm[5];
    D.sign = sign ? sign[m] : 1.0;
    int64_t total = 1;
    for (int b = 0; b < rank; ++b) {
      TN_REQUIRE(sd[b] >= 1, TN_EINVAL, "move %lld: sub-block extent < 1", (long long)m);
      total *= sd[b];
    }
    D.total = total;
    TN_REQUIRE(D.sbin_c >= 1 && D.sbot_c >= 1 && total % D.sbin_c == 0 && total % D.sbot_c == 0,
               TN_EINVAL, "move %lld: sub-block column counts do not divide its size",
               (long long)m);
    // Output order: sort source bonds by decreasing destination stride
    // (sBot_acc is the row-major stride list of the destination order).
    std::vector<int> ord(rank);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      if (acc[a] != acc[b]) return acc[a] > acc[b];
      return sd[a] > sd[b];
    });
    std::vector<int64_t> sacc(rank);
    sacc[rank - 1] = 1;
    for (int b = rank - 1; b > 0; --b) sacc[b - 1] = sacc[b] * sd[b];
    int64_t expect = 1;
    for (int k = rank - 1; k >= 0; --k) {
      int b = ord[k];
      TN_REQUIRE(sd[b] == 1 || acc[b] == expect, TN_EINVAL,
                 "move %lld: destination strides are not a permutation of the sub-block",
                 (long long)m);
      expect *= sd[b];
      D.odim[k] = sd[b];
      D.osrc[k] = sacc[b];
    }
    for (int64_t s0 = 0; s0 < total; s0 += kChunk)
      ch.push_back(ChunkDev{m, s0, std::min(kChunk, total - s0)});
    covered += total;
  }
  TN_REQUIRE((int64_t)ch.size() < (int64_t(1) << 31), TN_EUNSUPP, "too many chunks");
  auto* P = new tn_blockperm_plan_s();
  P->rank = rank;
  P->nmoves = nmoves;
  P->nchunks = (int64_t)ch.size();
  P->dst_elems = dst_elems;
  P->covered = covered;
  if (nmoves > 0) {
    int rc = tn::pool_alloc((void**)&P->d_moves, sizeof(MoveDev) * mv.size());
    if (rc == TN_OK) rc = tn::pool_alloc((void**)&P->d_chunks, sizeof(ChunkDev) * ch.size());
    if (rc == TN_OK) rc = tn_memcpy_h2d(P->d_moves, mv.data(), sizeof(MoveDev) * mv.size(), nullptr);
    if (rc == TN_OK) rc = tn_memcpy_h2d(P->d_chunks, ch.data(), sizeof(ChunkDev) * ch.size(), nullptr);
    if (rc != TN_OK) {
      tn::pool_free(P->d_moves);
      tn::pool_free(P->d_chunks);
      delete P;
      return rc;
    }
  }
  *plan = P;
  return TN_OK;
}

int tn_blockperm_exec(tn_blockperm_plan P, const double* src, double* dst, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_blockperm_exec: NULL plan");
  hipStream_t s = tn::stream_of(stream);
  if (P->covered < P->dst_elems) {
    int rc = tn_zero(dst, sizeof(double) * P->dst_elems, stream);
    if (rc != TN_OK) return rc;
  }
  if (P->nchunks == 0) return TN_OK;
  TN_REQUIRE(src && dst && src != dst, TN_EINVAL, "tn_blockperm_exec: bad buffers");
  hipLaunchKernelGGL(block_permute_kernel, dim3((unsigned)P->nchunks), dim3(256), 0, s, src, dst,
                     P->d_moves, P->d_chunks, P->rank);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_blockperm_plan_destroy(tn_blockperm_plan P) {
  if (!P) return TN_OK;
  tn::pool_free(P->d_moves);
  tn::pool_free(P->d_chunks);
  delete P;
  return TN_OK;
}


struct tn_rectscale_plan_s {
  int64_t nrects = 0, nchunks = 0;
  RectDev* d_rects = nullptr;
  ChunkDev* d_chunks = nullptr;
};

int tn_rectscale_plan_create(tn_rectscale_plan* plan, int64_t nrects, const int64_t* rects) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(plan != nullptr && nrects >= 0 && (nrects == 0 || rects != nullptr), TN_EINVAL,
             "tn_rectscale_plan_create: bad arguments");
  const int64_t kChunk = 4096;
  std::vector<RectDev> rv;
  std::vector<ChunkDev> ch;
  for (int64_t i = 0; i < nrects; ++i) {
    RectDev r{rects[4 * i], rects[4 * i + 1], rects[4 * i + 2], rects[4 * i + 3]};
    TN_REQUIRE(r.off >= 0 && r.rows >= 0 && r.cols >= 0 && (r.rows <= 1 || r.pitch >= r.cols), TN_EINVAL,
               "tn_rectscale_plan_create: rectangle %lld is malformed", (long long)i);
    const int64_t total = r.rows * r.cols;
    if (total == 0) continue;
    const int64_t idx = (int64_t)rv.size();
    rv.push_back(r);
    for (int64_t s0 = 0; s0 < total; s0 += kChunk) ch.push_back(ChunkDev{idx, s0, std::min(kChunk, total - s0)});
  }
  TN_REQUIRE((int64_t)ch.size() < (int64_t(1) << 31), TN_EUNSUPP, "too many chunks");
  auto* P = new tn_rectscale_plan_s();
  P->nrects = (int64_t)rv.size();
  P->nchunks = (int64_t)ch.size();
  if (!rv.empty()) {
    int rc = tn::pool_alloc((void**)&P->d_rects, sizeof(RectDev) * rv.size());
    if (rc == TN_OK) rc = tn::pool_alloc((void**)&P->d_chunks, sizeof(ChunkDev) * ch.size());
    if (rc == TN_OK) rc = tn_memcpy_h2d(P->d_rects, rv.data(), sizeof(RectDev) * rv.size(), nullptr);
    if (rc == TN_OK) rc = tn_memcpy_h2d(P->d_chunks, ch.data(), sizeof(ChunkDev) * ch.size(), nullptr);
    if (rc != TN_OK) {
      tn::pool_free(P->d_rects);
      tn::pool_free(P->d_chunks);
      delete P;
      return rc;
    }
  }
  *plan = P;
  return TN_OK;
}

int tn_rectscale_exec(tn_rectscale_plan P, double* base, double a, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_rectscale_exec: NULL plan");
  if (P->nchunks == 0) return TN_OK;
  TN_REQUIRE(base != nullptr, TN_EINVAL, "tn_rectscale_exec: NULL buffer");
  hipLaunchKernelGGL(rect_scale_kernel, dim3((unsigned)P->nchunks), dim3(256), 0, tn::stream_of(stream), base,
                     P->d_rects, P->d_chunks, a);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_rectscale_plan_destroy(tn_rectscale_plan P) {
  if (!P) return TN_OK;
  tn::pool_free(P->d_rects);
  tn::pool_free(P->d_chunks);
  delete P;
  return TN_OK;
}

}  // extern "C"
