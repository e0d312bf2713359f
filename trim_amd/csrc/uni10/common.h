// Shared helpers of the host C++ Uni10 layer.
#pragma once
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>

#include "tn_hip.h"
#include "uni10_amd/uni10.hpp"

namespace uni10 {
namespace detail {

[[noreturn]] inline void fail(const std::string& msg) { throw std::runtime_error(msg); }

inline void check(int rc, const char* what) {
  if (rc != TN_OK) {
    std::ostringstream os;
    os << what << ": " << tn_last_error() << " (code " << rc << ")";
    throw std::runtime_error(os.str());
  }
}

// SplitMix64 -> uniform[-1, 1): x_i = mix(seed + i * 0x9E3779B97F4A7C15),
// value = 2 * (x >> 11) * 2^-53 - 1, i = 1, 2, ...
inline double splitmix_uniform(uint64_t seed, uint64_t i) {
  uint64_t z = seed + i * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

void* stream();
int gemm_mode();
void sync();

// The directory of the network files the iTEBD drivers read: `netdir` when
// given, else the `nets` directory installed beside libuni10_amd.so.
std::string netdir_or_default(const std::string& netdir);

}  // namespace detail
}  // namespace uni10
