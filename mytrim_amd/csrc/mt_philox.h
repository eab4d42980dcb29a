// Counter-based random streams for the cascade kernels.
//
// Philox4x32-10 (J. K. Salmon et al., "Parallel random numbers: as easy as
// 1, 2, 3", SC'11, section 5).  One stream per (seed, history) — the key —
// addressed by (collision index, 64-bit flight key) — the counter — so a
// flight's draws do not depend on which lane or in which order it runs.
//
// Per collision ic >= 1 one block (w0..w3) gives
//   r2 = u53(w0, w1)  impact parameter     (trim.C:143)
//   hh = u32(w2)      target element        (trim.C:147)
//   az = u32(w3)      azimuth of the recoil (replaces trim.C:322-333)
// The first free path's r1 (trim.C:71) is u53 of a splitmix64 hash of
// (flight key, history, seed) (first_draw), a quarter of a Philox block's
// work: a wave evaluates it at almost every step, for its few lanes that
// start a flight.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mt {

struct P4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ P4 philox10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const uint64_t m0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t m1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(m1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(m0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)m1;
    c3 = (uint32_t)m0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return P4{c0, c1, c2, c3};
}

// (0, 1] from the top 53 bits of (a:b)
__host__ __device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  const uint64_t x = ((((uint64_t)a) << 32) | b) >> 11;
  return (double)(x + 1) * (1.0 / 9007199254740992.0);
}
// (0, 1] from 32 bits
__host__ __device__ __forceinline__ double u32d(uint32_t a) {
  return ((double)a + 1.0) * (1.0 / 4294967296.0);
}
// (0, 1] from the top 24 bits, single precision
__host__ __device__ __forceinline__ float u24f(uint32_t a) {
  return (float)((a >> 8) + 1u) * (1.0f / 16777216.0f);
}

// key of the flight spawned at collision ic of flight `parent` (splitmix64 finaliser)
__host__ __device__ __forceinline__ uint64_t child_key(uint64_t parent, uint32_t ic) {
  uint64_t z = parent ^ ((uint64_t)ic * 0x9E3779B97F4A7C15ull);
  z += 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// 64 random bits for the first draw of a flight (r1): splitmix64 finaliser
// over the flight key mixed with (history, seed)
__host__ __device__ __forceinline__ uint64_t first_draw(uint64_t key, uint32_t seed, uint32_t hist) {
  uint64_t z = key ^ (((((uint64_t)hist) << 32) | seed) * 0xD1B54A32D192ED03ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace mt
