// Counter-based random streams of the library (SplitMix64 finaliser over
// seed * C1 + (i + 1) * C2), shared by the synthetic panel, the RSVD and
// NMF probe matrices and the NMF cross-validation masks.  Element i of the
// stream with seed s is a pure function of (s, i), so any kernel can draw
// any element; schnellfmm_amd/synth.py produces the same numbers on the host.
#pragma once
#include <cmath>
#include <cstdint>

namespace slfmm {

constexpr uint64_t kStreamStep = 0x9E3779B97F4A7C15ull;

// Counter of element i of stream `seed`; element i + 1 has counter + kStreamStep.
__host__ __device__ __forceinline__ uint64_t stream_counter(uint64_t seed, uint64_t i) {
  return seed * 0xD1B54A32D192ED03ull + (i + 1) * kStreamStep;
}

__host__ __device__ __forceinline__ uint64_t mix_counter(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t mix64(uint64_t seed, uint64_t i) { return mix_counter(stream_counter(seed, i)); }

// Integer form of a Bernoulli draw: u01(s, i) < rate  <=>  (mix64(s, i) >> 11) < u01_threshold(rate)
// (u01 is m 2^-53 with m < 2^53 an integer, and rate 2^53 is exact).
inline uint64_t u01_threshold(double rate) {
  if (!(rate > 0.0)) return 0;
  if (rate >= 1.0) return 1ull << 53;
  return (uint64_t)std::ceil(rate * 9007199254740992.0);
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
