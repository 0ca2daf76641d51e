// Counter-based random numbers for the path-tracing core (host and device).
//
// The reference draws every random number from one process-global boost
// mt19937 shared by all render threads (stats/Random.cpp:29-56), so a path's
// stream is not reproducible.  Here every draw is a Philox4x32-10 block
// (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC 2011)
// keyed by (seed, domain) and indexed by a 4-word counter, so a path's numbers
// depend only on (seed, job, pixel, sample, depth, stream): results are
// independent of scheduling, batching and thread count.
#ifndef MPT_RNG_H_
#define MPT_RNG_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpt_rng {

// Key of the path-tracing streams ("MPT1").
constexpr uint32_t kPathKey = 0x4D505431u;
// Depth value of the camera stream.
constexpr uint32_t kCameraDepth = 255;

struct U4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * b) >> 32);
}

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = mulhi32(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = mulhi32(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 24-bit uniform in [0, 1), exact in float and double.
__host__ __device__ inline float unitf(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
// floor(3 * u) for the 24-bit uniform of x (the path's IOR channel).
__host__ __device__ inline uint32_t third(uint32_t x) { return (uint32_t)(((uint64_t)(x >> 8) * 3u) >> 24); }

// Path-tracing stream: counter (pixel, sample, depth << 8 | stream, job).
__host__ __device__ inline U4 path_draw(uint32_t seed, uint32_t job, uint32_t pixel, uint32_t sample,
                                        uint32_t depth, uint32_t stream) {
  return philox4x32_10(U4{pixel, sample, (depth << 8) | stream, job}, seed, kPathKey);
}

// Host helpers for the scene generators.
inline void philox_words(uint32_t seed, uint32_t domain, uint32_t a, uint32_t b, uint32_t out[4]) {
  const U4 r = philox4x32_10(U4{a, b, 0, 0}, seed, domain);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
inline double unit32(uint32_t x) { return x * (1.0 / 4294967296.0); }
inline double unit53(uint32_t hi, uint32_t lo) {
  return (double)(((uint64_t)hi << 21) ^ (lo >> 11)) * (1.0 / 9007199254740992.0);
}

}  // namespace mpt_rng

#endif
