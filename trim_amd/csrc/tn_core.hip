// Device selection, error reporting, the caching device-memory pool,
// copies/fills, streams and events of libtn_hip.so.
//
// Replaces the reference's global GPU state and allocators
// (uni10/src/uni10/tools/lib/uni10_tools_gpu.cu:7-163): instead of
// cudaMallocManaged buffers that host loops touch through page migration,
// every buffer is explicit device memory handed out by a size-binned cache,
// with a byte budget used for admission control by the request loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "tn_internal.h"

namespace tn {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
void clear_error() { g_err[0] = 0; }

struct DeviceState {
  std::mutex mu;
  bool ready = false;
  int device = -1;
  int cus = 0;
  size_t hbm = 0;
  char arch[64] = {0};
  hipStream_t stream = nullptr;  // library default stream (default priority)
  // pool.  A freed block keeps the stream it was freed on (the freeing
  // thread's pool stream): an allocation on that stream takes it at once
  // (stream order protects the earlier uses); one on another stream first
  // makes its stream wait for an event recorded on the freeing stream.
  // A null tag means no work can be pending on the block.
  struct Freed {
    void* p;
    hipStream_t s;
  };
  std::multimap<size_t, Freed> free_blocks;       // size -> block
  hipEvent_t reuse_ev = nullptr;                  // cross-stream reuse fence
  std::unordered_map<void*, size_t> live;         // block -> size
  size_t in_use = 0, cached = 0, peak = 0, budget = 0;
  // Memory of captured graphs.  A block allocated or freed while a capture
  // is in progress belongs to that graph: its replays read and write it, so
  // freeing it only marks it released, and it returns to the pool when the
  // graph is destroyed.  Within the capture, a block released earlier is
  // handed to later allocations of the same capture: the graph runs in
  // stream order, as the captured calls did.
  struct GraphMem {
    std::vector<void*> owned;
    std::multimap<size_t, void*> local_free;      // released in this capture
  };
  GraphMem* capturing = nullptr;                  // the capture in progress
  std::unordered_map<void*, GraphMem*> graph_mem; // graph exec -> its memory
  std::unordered_map<void*, bool> owned;          // graph-owned block -> released by its holder
};

static DeviceState& st() {
  static DeviceState s;
  return s;
}

static int init_locked(DeviceState& s, int device) {
  if (s.ready) {
    TN_REQUIRE(device < 0 || device == s.device, TN_EINVAL,
               "tn_init: already initialised on device %d (asked %d)", s.device, device);
    return TN_OK;
  }
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  TN_REQUIRE(e == hipSuccess && n > 0, TN_ENODEV, "no HIP device available (%s)",
             e == hipSuccess ? "count=0" : hipGetErrorString(e));
  if (device < 0) device = 0;
  TN_REQUIRE(device < n, TN_EINVAL, "device %d out of range (%d devices)", device, n);
  TN_HIP_CHECK(hipSetDevice(device));
  hipDeviceProp_t p;
  TN_HIP_CHECK(hipGetDeviceProperties(&p, device));
  TN_REQUIRE(std::strncmp(p.gcnArchName, "gfx950", 6) == 0, TN_ENODEV,
             "device %d is %s; this library is built for gfx950 only", device, p.gcnArchName);
  s.cus = p.multiProcessorCount;
  s.hbm = p.totalGlobalMem;
  std::snprintf(s.arch, sizeof(s.arch), "%s", p.gcnArchName);
  TN_HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
  s.device = device;
  s.ready = true;
  return TN_OK;
}

int ensure_init() {
  DeviceState& s = st();
  if (s.ready) return TN_OK;
  std::lock_guard<std::mutex> g(s.mu);
  return init_locked(s, -1);
}

int cu_count() { return st().cus > 0 ? st().cus : 256; }

hipStream_t stream_of(void* h) { return h ? (hipStream_t)h : st().stream; }

// The stream this thread's allocations and frees are ordered on.
static thread_local void* t_pool_stream = nullptr;

static size_t bin_size(size_t bytes) {
  if (bytes == 0) bytes = 1;
  const size_t small = size_t(1) << 20;
  if (bytes <= small) return (bytes + 511) & ~size_t(511);
  const size_t g = size_t(2) << 20;
  return (bytes + g - 1) / g * g;
}

static int release_cached_locked(DeviceState& s) {
  for (auto& kv : s.free_blocks) {
    hipError_t e = hipFree(kv.second.p);
    if (e != hipSuccess) {
      set_error("hipFree failed: %s", hipGetErrorString(e));
      return TN_EHIP;
    }
  }
  s.free_blocks.clear();
  s.cached = 0;
  return TN_OK;
}

int pool_alloc(void** p, size_t bytes) {
  TN_REQUIRE(p != nullptr, TN_EINVAL, "tn_malloc: NULL out pointer");
  TN_INIT_OR_RETURN();
  DeviceState& s = st();
  std::lock_guard<std::mutex> g(s.mu);
  size_t sz = bin_size(bytes);
  if (s.capturing) {
    auto& lf = s.capturing->local_free;
    auto it = lf.lower_bound(sz);
    if (it != lf.end() && it->first <= sz + sz / 4) {
      void* q = it->second;
      lf.erase(it);
      s.owned[q] = false;    // held again (already counted in use)
      *p = q;
      return TN_OK;
    }
  }
  TN_REQUIRE(s.budget == 0 || s.in_use + sz <= s.budget, TN_ENOMEM,
             "device memory budget exceeded: in use %zu + request %zu > budget %zu",
             s.in_use, sz, s.budget);
  // a block of this stream (or idle) first; else one freed on another stream,
  // fenced by an event (not inside a capture: the graph could not wait on it)
  const hipStream_t mine = stream_of(t_pool_stream);
  auto pick = s.free_blocks.end();
  int seen = 0;
  for (auto it = s.free_blocks.lower_bound(sz); it != s.free_blocks.end() && it->first <= sz + sz / 4 && seen < 16;
       ++it, ++seen) {
    if (it->second.s == mine || it->second.s == nullptr) {
      pick = it;
      break;
    }
    if (pick == s.free_blocks.end() && !s.capturing) pick = it;
  }
  if (pick != s.free_blocks.end()) {
    const DeviceState::Freed f = pick->second;
    if (f.s != nullptr && f.s != mine) {
      if (!s.reuse_ev) TN_HIP_CHECK(hipEventCreateWithFlags(&s.reuse_ev, hipEventDisableTiming));
      TN_HIP_CHECK(hipEventRecord(s.reuse_ev, f.s));
      TN_HIP_CHECK(hipStreamWaitEvent(mine, s.reuse_ev, 0));
    }
    const size_t qs = pick->first;
    s.free_blocks.erase(pick);
    s.cached -= qs;
    s.live[f.p] = qs;
    s.in_use += qs;
    s.peak = std::max(s.peak, s.in_use);
    *p = f.p;
    return TN_OK;
  }
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    int rc = release_cached_locked(s);
    if (rc != TN_OK) return rc;
    e = hipMalloc(&q, sz);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error("hipMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
    return TN_ENOMEM;
  }
  s.live[q] = sz;
  s.in_use += sz;
  s.peak = std::max(s.peak, s.in_use);
  if (s.capturing) {
    s.capturing->owned.push_back(q);
    s.owned[q] = false;
  }
  *p = q;
  return TN_OK;
}

// Returns a live block to its size bin.
static void free_locked(DeviceState& s, void* p) {
  auto it = s.live.find(p);
  if (it == s.live.end()) return;
  const size_t sz = it->second;
  s.live.erase(it);
  s.in_use -= sz;
  s.free_blocks.emplace(sz, DeviceState::Freed{p, stream_of(t_pool_stream)});
  s.cached += sz;
}

// The graph's blocks go back to the pool (those its holders released) or
// to their holders (the rest, freed normally later).
static void release_graph_mem_locked(DeviceState& s, DeviceState::GraphMem* m) {
  for (void* p : m->owned) {
    auto it = s.owned.find(p);
    if (it == s.owned.end()) continue;
    const bool released = it->second;
    s.owned.erase(it);
    if (released) free_locked(s, p);
  }
  delete m;
}

int pool_free(void* p) {
  if (p == nullptr) return TN_OK;
  DeviceState& s = st();
  std::lock_guard<std::mutex> g(s.mu);
  auto it = s.live.find(p);
  TN_REQUIRE(it != s.live.end(), TN_EINVAL, "tn_free: %p was not allocated by tn_malloc", p);
  auto ow = s.owned.find(p);
  if (s.capturing) {
    DeviceState::GraphMem* m = s.capturing;
    const bool mine = ow != s.owned.end() &&
                      std::find(m->owned.begin(), m->owned.end(), p) != m->owned.end();
    if (ow == s.owned.end() || mine) {
      // the captured calls may still read it: it joins this graph and is
      // reusable by the capture's later allocations only
      if (ow == s.owned.end()) m->owned.push_back(p);
      s.owned[p] = true;
      m->local_free.emplace(it->second, p);
      return TN_OK;
    }
  }
  if (ow != s.owned.end()) {
    ow->second = true;       // a graph still uses it: returned when that graph is destroyed
    return TN_OK;
  }
  free_locked(s, p);
  return TN_OK;
}

typedef double dbl2 __attribute__((ext_vector_type(2)));

// One 16-byte element per thread: many small workgroups keep the most
// requests in flight (measured 6.2 TB/s plain, 6.7 TB/s streaming, against
// 5.3 TB/s for the runtime's device-to-device copy, 1 GiB moved).
template <bool NT>
__global__ void __launch_bounds__(256) copy16_kernel(const dbl2* __restrict__ src, dbl2* __restrict__ dst,
                                                     int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if constexpr (NT) {
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
  } else {
    dst[i] = src[i];
  }
}

int copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return TN_OK;
  const int64_t n = (int64_t)(bytes / 16);
  if ((((uintptr_t)dst | (uintptr_t)src | bytes) & 15) != 0 || (n + 255) / 256 >= (int64_t(1) << 31)) {
    TN_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    return TN_OK;
  }
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (bytes >= kStreamingBytes)
    hipLaunchKernelGGL(copy16_kernel<true>, dim3(grid), dim3(256), 0, s, (const dbl2*)src, (dbl2*)dst, n);
  else
    hipLaunchKernelGGL(copy16_kernel<false>, dim3(grid), dim3(256), 0, s, (const dbl2*)src, (dbl2*)dst, n);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

}  // namespace tn

using namespace tn;

// ---------------------------------------------------------------------------
__global__ void tn_fill_zero_kernel(uint64_t* __restrict__ p, size_t n_words,
                                    unsigned char* __restrict__ tail, size_t n_tail) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n_words; i += stride) p[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x < n_tail) tail[threadIdx.x] = 0;
}

extern "C" {

const char* tn_last_error(void) { return g_err; }

int tn_init(int device) {
  DeviceState& s = st();
  std::lock_guard<std::mutex> g(s.mu);
  return init_locked(s, device);
}

int tn_device_info(int* cus, size_t* hbm, char* arch, int arch_len) {
  TN_INIT_OR_RETURN();
  DeviceState& s = st();
  if (cus) *cus = s.cus;
  if (hbm) *hbm = s.hbm;
  if (arch && arch_len > 0) std::snprintf(arch, arch_len, "%s", s.arch);
  return TN_OK;
}

int tn_set_budget(size_t bytes) {
  DeviceState& s = st();
  std::lock_guard<std::mutex> g(s.mu);
  s.budget = bytes;
  return TN_OK;
}

int tn_mem_stats(size_t* in_use, size_t* cached, size_t* peak, size_t* budget) {
  DeviceState& s = st();
  std::lock_guard<std::mutex> g(s.mu);
  if (in_use) *in_use = s.in_use;
  if (cached) *cached = s.cached;
  if (peak) *peak = s.peak;
  if (budget) *budget = s.budget;
  return TN_OK;
}

int tn_malloc(void** dptr, size_t bytes) { return pool_alloc(dptr, bytes); }
int tn_free(void* dptr) { return pool_free(dptr); }

int tn_pool_trim(void) {
  TN_INIT_OR_RETURN();
  TN_HIP_CHECK(hipDeviceSynchronize());
  DeviceState& s = st();
  std::lock_guard<std::mutex> g(s.mu);
  return release_cached_locked(s);
}

int tn_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  TN_INIT_OR_RETURN();
  if (bytes == 0) return TN_OK;
  TN_REQUIRE(dst && src, TN_EINVAL, "tn_memcpy_h2d: NULL pointer");
  hipStream_t s = stream_of(stream);
  TN_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  TN_HIP_CHECK(hipStreamSynchronize(s));  // source may be pageable/transient
  return TN_OK;
}

int tn_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  TN_INIT_OR_RETURN();
  if (bytes == 0) return TN_OK;
  TN_REQUIRE(dst && src, TN_EINVAL, "tn_memcpy_d2h: NULL pointer");
  hipStream_t s = stream_of(stream);
  TN_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  TN_HIP_CHECK(hipStreamSynchronize(s));
  return TN_OK;
}

int tn_host_alloc(void** p, size_t bytes) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(p != nullptr, TN_EINVAL, "tn_host_alloc: NULL out pointer");
  *p = nullptr;
  TN_HIP_CHECK(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault));
  return TN_OK;
}

int tn_host_free(void* p) {
  if (!p) return TN_OK;
  TN_HIP_CHECK(hipHostFree(p));
  return TN_OK;
}

int tn_memcpy_d2h_async(void* dst_pinned, const void* src, size_t bytes, void* stream) {
  TN_INIT_OR_RETURN();
  if (bytes == 0) return TN_OK;
  TN_REQUIRE(dst_pinned && src, TN_EINVAL, "tn_memcpy_d2h_async: NULL pointer");
  TN_HIP_CHECK(hipMemcpyAsync(dst_pinned, src, bytes, hipMemcpyDeviceToHost, stream_of(stream)));
  return TN_OK;
}

int tn_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  TN_INIT_OR_RETURN();
  if (bytes == 0) return TN_OK;
  TN_REQUIRE(dst && src, TN_EINVAL, "tn_memcpy_d2d: NULL pointer");
  return copy_d2d(dst, src, bytes, stream_of(stream));
}

int tn_zero(void* dst, size_t bytes, void* stream) {
  TN_INIT_OR_RETURN();
  if (bytes == 0) return TN_OK;
  TN_REQUIRE(dst != nullptr, TN_EINVAL, "tn_zero: NULL pointer");
  TN_REQUIRE(((uintptr_t)dst & 7) == 0, TN_EINVAL, "tn_zero: pointer not 8-byte aligned");
  size_t words = bytes / 8, tail = bytes % 8;
  size_t blocks = std::min<size_t>((words + 255) / 256, 4096);
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(tn_fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     stream_of(stream), (uint64_t*)dst, words,
                     (unsigned char*)dst + words * 8, tail);
  TN_LAUNCH_CHECK();
  return TN_OK;
}

int tn_stream_create(void** stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(stream != nullptr, TN_EINVAL, "tn_stream_create: NULL out pointer");
  hipStream_t s;
  TN_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return TN_OK;
}

int tn_stream_destroy(void* stream) {
  if (!stream) return TN_OK;
  // blocks freed on it are idle once it drains: untag them before the
  // handle goes away (a later fence would record on a dead stream)
  TN_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  {
    DeviceState& s = st();
    std::lock_guard<std::mutex> g(s.mu);
    for (auto& kv : s.free_blocks)
      if (kv.second.s == (hipStream_t)stream) kv.second.s = nullptr;
  }
  if (t_pool_stream == stream) t_pool_stream = nullptr;
  TN_HIP_CHECK(hipStreamDestroy((hipStream_t)stream));
  return TN_OK;
}

int tn_pool_set_stream(void* stream) {
  t_pool_stream = stream;
  return TN_OK;
}

int tn_stream_sync(void* stream) {
  TN_INIT_OR_RETURN();
  TN_HIP_CHECK(hipStreamSynchronize(stream_of(stream)));
  return TN_OK;
}

int tn_stream_query(void* stream, int* done) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(done != nullptr, TN_EINVAL, "tn_stream_query: NULL out pointer");
  const hipError_t e = hipStreamQuery(stream_of(stream));
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();  // not an error: clear it so later checks do not see it
    *done = 0;
    return TN_OK;
  }
  TN_HIP_CHECK(e);
  *done = 1;
  return TN_OK;
}

int tn_device_sync(void) {
  TN_INIT_OR_RETURN();
  TN_HIP_CHECK(hipDeviceSynchronize());
  return TN_OK;
}

int tn_event_create(void** ev) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(ev != nullptr, TN_EINVAL, "tn_event_create: NULL out pointer");
  hipEvent_t e;
  TN_HIP_CHECK(hipEventCreate(&e));
  *ev = (void*)e;
  return TN_OK;
}

int tn_event_destroy(void* ev) {
  if (!ev) return TN_OK;
  TN_HIP_CHECK(hipEventDestroy((hipEvent_t)ev));
  return TN_OK;
}

int tn_event_record(void* ev, void* stream) {
  TN_INIT_OR_RETURN();
  TN_HIP_CHECK(hipEventRecord((hipEvent_t)ev, stream_of(stream)));
  return TN_OK;
}

int tn_event_query(void* ev, int* done) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(ev && done, TN_EINVAL, "tn_event_query: NULL pointer");
  const hipError_t e = hipEventQuery((hipEvent_t)ev);
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    *done = 0;
    return TN_OK;
  }
  TN_REQUIRE(e == hipSuccess, TN_EHIP, "hipEventQuery: %s", hipGetErrorString(e));
  *done = 1;
  return TN_OK;
}

int tn_graph_begin(void* stream) {
  TN_INIT_OR_RETURN();
  DeviceState& s = st();
  {
    std::lock_guard<std::mutex> g(s.mu);
    TN_REQUIRE(s.capturing == nullptr, TN_EINVAL, "tn_graph_begin: another capture is in progress");
    s.capturing = new DeviceState::GraphMem();
  }
  // relaxed: a pool miss during the capture may call hipMalloc
  const hipError_t e = hipStreamBeginCapture(stream_of(stream), hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    std::lock_guard<std::mutex> g(s.mu);
    release_graph_mem_locked(s, s.capturing);
    s.capturing = nullptr;
    set_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
    return TN_EHIP;
  }
  return TN_OK;
}

int tn_graph_end(void* stream, void** graph_exec) {
  TN_INIT_OR_RETURN();
  DeviceState& s = st();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(stream_of(stream), &g);
  hipGraphExec_t ge = nullptr;
  if (e == hipSuccess && graph_exec) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  std::lock_guard<std::mutex> lk(s.mu);
  DeviceState::GraphMem* m = s.capturing;
  s.capturing = nullptr;
  if (e != hipSuccess || !graph_exec) {
    if (m) release_graph_mem_locked(s, m);   // nothing captured will ever run
    if (!graph_exec) {
      set_error("tn_graph_end: NULL out pointer");
      return TN_EINVAL;
    }
    set_error("graph capture: %s", hipGetErrorString(e));
    return TN_EHIP;
  }
  if (m) s.graph_mem[(void*)ge] = m;
  *graph_exec = (void*)ge;
  return TN_OK;
}

int tn_graph_launch(void* graph_exec, void* stream) {
  TN_INIT_OR_RETURN();
  TN_REQUIRE(graph_exec != nullptr, TN_EINVAL, "tn_graph_launch: NULL graph");
  TN_HIP_CHECK(hipGraphLaunch((hipGraphExec_t)graph_exec, stream_of(stream)));
  return TN_OK;
}

int tn_graph_destroy(void* graph_exec) {
  if (!graph_exec) return TN_OK;
  // a replay may still be running: its memory is reused only after it ends
  const hipError_t e = hipDeviceSynchronize();
  TN_HIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
  DeviceState& s = st();
  {
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.graph_mem.find(graph_exec);
    if (it != s.graph_mem.end()) {
      release_graph_mem_locked(s, it->second);
      s.graph_mem.erase(it);
    }
  }
  TN_REQUIRE(e == hipSuccess, TN_EHIP, "hipDeviceSynchronize: %s", hipGetErrorString(e));
  return TN_OK;
}

int tn_event_elapsed_ms(void* a, void* b, float* ms) {
  TN_INIT_OR_RETURN();
  TN_HIP_CHECK(hipEventSynchronize((hipEvent_t)b));
  TN_HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return TN_OK;
}

}  // extern "C"
