// Counter-based N(0,1) generator shared by the device input generators:
// SplitMix64 hashes of (key, 2c) and (key, 2c+1) give two 53-bit uniforms,
// Box-Muller turns them into one normal.  The numpy restatement in the oracle
// uses the same integer hash, so host and device agree to rounding of the
// transcendental functions.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

inline uint64_t rng_key(uint64_t seed) { return splitmix64(seed); }

__device__ inline double normal_from_counter(uint64_t key, uint64_t ctr) {
  uint64_t h1 = splitmix64(key + 2 * ctr), h2 = splitmix64(key + 2 * ctr + 1);
  double u1 = (double)((h1 >> 11) + 1) * 0x1.0p-53;  // (0, 1]
  double u2 = (double)(h2 >> 11) * 0x1.0p-53;        // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

}  // namespace rta
