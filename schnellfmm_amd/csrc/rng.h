// Counter-based random streams of the library (SplitMix64 finaliser over
// seed * C1 + (i + 1) * C2), shared by the synthetic panel, the RSVD and
// NMF probe matrices and the NMF cross-validation masks.  Element i of the
// stream with seed s is a pure function of (s, i), so any kernel can draw
// any element; schnellfmm_amd/synth.py produces the same numbers on the host.
#pragma once
#include <cstdint>

namespace slfmm {

__device__ __forceinline__ uint64_t mix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed * 0xD1B54A32D192ED03ull + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// U[0,1) with 53 random bits
__device__ __forceinline__ double u01(uint64_t seed, uint64_t i) {
  return (double)(mix64(seed, i) >> 11) * (1.0 / 9007199254740992.0);
}

// N(0,1) from the Box-Muller pair (2i, 2i+1) of the stream
__device__ inline double normal01(uint64_t seed, uint64_t i) {
  const double u1 = 1.0 - u01(seed, 2 * i), u2 = u01(seed, 2 * i + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

}  // namespace slfmm
