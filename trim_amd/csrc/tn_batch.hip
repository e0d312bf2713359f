// Batched vector operations for many independent problems in one launch.
//
// The per-chain steps of a batched TEBD update between its GEMMs and SVDs
// are all short vector operations: the lambda row scaling of theta
// (diagRowMul, uni10_lapack_cpu.cpp:46-57), truncation and renormalisation
// of the singular values (iTEBD.cpp:157-161), the copy of the kept VT rows
// into the new Gamma (iTEBD.cpp:162-166) and the 1/norm scaling
// (vectorScal, uni10_lapack_cpu.cpp:75-89).  A plan holds one item per
// (chain, operation) with offsets into a single arena; executing it is one
// kernel launch over all items, with no allocation and no host round trip,
// so it can be captured into a graph.
//
// Items in one plan must be independent (an item never reads what another
// item of the same plan writes).  Each item is cut into chunks of
// kChunk elements (one workgroup each); a truncation item is one chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "tn_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int64_t kChunk = 4096;

struct ChunkDev {
  int32_t item;
  int32_t mask1;        // 1 + the mask entry that gates the item (0: never skipped)
  int64_t begin, end;   // element range of the item
};

template <int T>
__device__ inline double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int o = T / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kThreads) vecops_kernel(double* __restrict__ base,
                                                          const tn_vecop* __restrict__ items,
                                                          const ChunkDev* __restrict__ chunks,
                                                          const int* __restrict__ mask) {
  __shared__ double red[kThreads];
  const ChunkDev ck = chunks[blockIdx.x];
  if (mask != nullptr && ck.mask1 > 0 && mask[ck.mask1 - 1] == 0) return;
  const tn_vecop it = items[ck.item];
  double* __restrict__ dst = base + it.dst;
  const double* __restrict__ src = base + it.src;
  switch (it.op) {
    case TN_VECOP_ROWSCALE: {   // dst[i] = src[i] * aux[i / rep]
      const double* __restrict__ v = base + it.aux;
      for (int64_t i = ck.begin + threadIdx.x; i < ck.end; i += kThreads) dst[i] = src[i] * v[i / it.rep];
      break;
    }
    case TN_VECOP_COPY:
      for (int64_t i = ck.begin + threadIdx.x; i < ck.end; i += kThreads) dst[i] = src[i];
      break;
    case TN_VECOP_SCALE_INV: {   // dst[i] = src[i] / aux[0]
      const double inv = 1.0 / base[it.aux];
      for (int64_t i = ck.begin + threadIdx.x; i < ck.end; i += kThreads) dst[i] = src[i] * inv;
      break;
    }
    case TN_VECOP_TRUNCATE: {
      // src = S (n values, descending), keep rep values: dst = S[:rep] / ||S[:rep]||,
      // aux[0] = ||S[:rep]||, aux[1] = sum of the dropped S_i^2
      double a = 0, b = 0;
      for (int64_t i = threadIdx.x; i < it.n; i += kThreads) {
        const double s = src[i];
        if (i < it.rep)
          a += s * s;
        else
          b += s * s;
      }
      const double keep = block_sum<kThreads>(a, red);
      const double drop = block_sum<kThreads>(b, red);
      const double nrm = sqrt(keep);
      for (int64_t i = threadIdx.x; i < it.rep; i += kThreads) dst[i] = src[i] / nrm;
      if (threadIdx.x == 0) {
        base[it.aux] = nrm;
        base[it.aux + 1] = drop;
      }
      break;
    }
  }
}

}  // namespace

struct tn_vecops_plan_s {
  tn_vecop* d_items = nullptr;
  ChunkDev* d_chunks = nullptr;
  int64_t nchunks = 0;
  std::vector<ChunkDev> chunks;   // host copy (for set_mask)
  const int* mask = nullptr;
};

extern "C" {

int tn_vecops_plan_create(tn_vecops_plan* plan, int64_t nitems, const tn_vecop* items) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(plan != nullptr, TN_EINVAL, "tn_vecops_plan_create: NULL out pointer");
  TN_REQUIRE(nitems >= 0 && (nitems == 0 || items), TN_EINVAL, "tn_vecops_plan_create: NULL items");
  std::vector<ChunkDev> chunks;
  for (int64_t i = 0; i < nitems; ++i) {
    const tn_vecop& it = items[i];
    TN_REQUIRE(it.op >= TN_VECOP_ROWSCALE && it.op <= TN_VECOP_TRUNCATE, TN_EINVAL, "vecop %lld: unknown op %d",
               (long long)i, it.op);
    TN_REQUIRE(it.n >= 0 && it.src >= 0 && it.dst >= 0, TN_EINVAL, "vecop %lld: negative size or offset",
               (long long)i);
    TN_REQUIRE(it.op == TN_VECOP_COPY || it.aux >= 0, TN_EINVAL, "vecop %lld: negative aux offset", (long long)i);
    TN_REQUIRE(it.op != TN_VECOP_ROWSCALE || it.rep >= 1, TN_EINVAL, "vecop %lld: row length < 1", (long long)i);
    TN_REQUIRE(it.op != TN_VECOP_TRUNCATE || (it.rep >= 1 && it.rep <= it.n), TN_EINVAL,
               "vecop %lld: kept count outside [1, n]", (long long)i);
    TN_REQUIRE(i < (int64_t(1) << 31), TN_EUNSUPP, "too many vecop items");
    if (it.op == TN_VECOP_TRUNCATE) {
      chunks.push_back(ChunkDev{(int32_t)i, 0, 0, it.n});
      continue;
    }
    for (int64_t b = 0; b < it.n; b += kChunk) chunks.push_back(ChunkDev{(int32_t)i, 0, b, std::min(it.n, b + kChunk)});
  }
  auto* P = new tn_vecops_plan_s();
  P->nchunks = (int64_t)chunks.size();
  P->chunks = chunks;
  int rc = TN_OK;
  if (nitems > 0) rc = tn::pool_alloc((void**)&P->d_items, sizeof(tn_vecop) * nitems);
  if (rc == TN_OK && nitems > 0) rc = tn_memcpy_h2d(P->d_items, items, sizeof(tn_vecop) * nitems, nullptr);
  if (rc == TN_OK && P->nchunks > 0) rc = tn::pool_alloc((void**)&P->d_chunks, sizeof(ChunkDev) * P->nchunks);
  if (rc == TN_OK && P->nchunks > 0)
    rc = tn_memcpy_h2d(P->d_chunks, chunks.data(), sizeof(ChunkDev) * P->nchunks, nullptr);
  if (rc != TN_OK) {
    tn::pool_free(P->d_items);
    tn::pool_free(P->d_chunks);
    delete P;
    return rc;
  }
  *plan = P;
  return TN_OK;
}

int tn_vecops_plan_exec(tn_vecops_plan P, double* base, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_vecops_plan_exec: NULL plan");
  if (P->nchunks == 0) return TN_OK;
  TN_REQUIRE(base != nullptr, TN_EINVAL, "tn_vecops_plan_exec: NULL base");
  TN_REQUIRE(P->nchunks < (int64_t(1) << 31), TN_EUNSUPP, "too many vecop chunks");
  hipLaunchKernelGGL(vecops_kernel, dim3((unsigned)P->nchunks), dim3(kThreads), 0, tn::stream_of(stream), base,
                     P->d_items, P->d_chunks, P->mask);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_vecops_plan_set_mask(tn_vecops_plan P, const int* mask, const int32_t* item_idx) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(P != nullptr, TN_EINVAL, "tn_vecops_plan_set_mask: NULL plan");
  TN_REQUIRE(mask == nullptr || item_idx != nullptr || P->chunks.empty(), TN_EINVAL,
             "tn_vecops_plan_set_mask: NULL item indices");
  for (ChunkDev& c : P->chunks) c.mask1 = (mask && item_idx[c.item] >= 0) ? item_idx[c.item] + 1 : 0;
  int rc = tn_device_sync();   // launches queued with the old chunks finish first
  if (rc == TN_OK && P->nchunks > 0)
    rc = tn_memcpy_h2d(P->d_chunks, P->chunks.data(), sizeof(ChunkDev) * P->nchunks, nullptr);
  if (rc == TN_OK) P->mask = mask;
  return rc;
}

int tn_vecops_plan_destroy(tn_vecops_plan P) {
  if (!P) return TN_OK;
  const hipError_t e = hipDeviceSynchronize();   // launches may still read the plan
  tn::pool_free(P->d_items);
  tn::pool_free(P->d_chunks);
  delete P;
  TN_REQUIRE(e == hipSuccess, TN_EHIP, "hipDeviceSynchronize: %s", hipGetErrorString(e));
  return TN_OK;
}

}  // extern "C"
