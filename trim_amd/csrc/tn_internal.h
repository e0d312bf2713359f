// Internal helpers shared by the HIP translation units of libtn_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>
#include <utility>

#include "tn_hip.h"

namespace tn {

// Thread-local last-error message (what tn_last_error() returns).
void set_error(const char* fmt, ...);
void clear_error();

// The stream to use for a caller-supplied stream handle (NULL -> default).
hipStream_t stream_of(void* s);

// Ensure the device is initialised; returns TN_OK or an error code.
int ensure_init();

int cu_count();

// Device allocation through the caching pool (used internally for plans and
// scratch; same pool as tn_malloc).
int pool_alloc(void** p, size_t bytes);
int pool_free(void* p);

// Device-to-device copy on stream s: a copy kernel (16 bytes per thread,
// streaming loads and stores when the copy is larger than the on-chip
// caches) when both pointers and the size are 16-byte multiples, the
// runtime's copy otherwise.
int copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t s);

// Copies at or above this size (bytes read) use non-temporal loads and
// stores: neither side is worth keeping in L2 / the Infinity Cache.
constexpr size_t kStreamingBytes = size_t(64) << 20;

// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>) in turn:
// a loop whose index is a constant in its body, so that it can index
// register arrays (a runtime index into one becomes select chains, LDS round
// trips or branch trees).
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

}  // namespace tn

#define TN_HIP_CHECK(expr)                                                   \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::tn::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                      __FILE__, __LINE__);                                   \
      return TN_EHIP;                                                        \
    }                                                                        \
  } while (0)

#define TN_REQUIRE(cond, code, ...)  \
  do {                               \
    if (!(cond)) {                   \
      ::tn::set_error(__VA_ARGS__);  \
      return (code);                 \
    }                                \
  } while (0)

#define TN_INIT_OR_RETURN()        \
  do {                             \
    int _rc = ::tn::ensure_init(); \
    if (_rc != TN_OK) return _rc;  \
  } while (0)

// Kernel launch error check (launch configuration errors only).
#define TN_LAUNCH_CHECK()                                                   \
  do {                                                                      \
    hipError_t _e = hipGetLastError();                                      \
    if (_e != hipSuccess) {                                                 \
      ::tn::set_error("kernel launch failed: %s (%s:%d)",                   \
                      hipGetErrorString(_e), __FILE__, __LINE__);           \
      return TN_EHIP;                                                       \
    }                                                                       \
  } while (0)
