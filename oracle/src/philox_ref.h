// ORACLE — test infrastructure only (never linked into the product).
//
// Philox4x32-10 counter-based generator, written from the published
// description: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11 (2011), section 5.
// Round function: two 32x32->64 multiplications by M0=0xD2511F53 and
// M1=0xCD9E8D57, key schedule bumped by the Weyl constants W0=0x9E3779B9,
// W1=0xBB67AE85.  Pinned by the known-answer vectors in tests/test_oracle.py.
//
// The device path keys one stream per (seed, history) and addresses it by
// (collision index, 64-bit flight key), see flight_key_child().  This file is
// the oracle's own implementation of that scheme; the HIP library has an
// independent one.
#pragma once
#include <cstdint>

namespace orc {

struct Philox4 {
  uint32_t v[4];
};

inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                             uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t p0 = (uint64_t)M0 * c0;
    const uint64_t p1 = (uint64_t)M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n1 = lo1;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    const uint32_t n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  Philox4 r;
  r.v[0] = c0;
  r.v[1] = c1;
  r.v[2] = c2;
  r.v[3] = c3;
  return r;
}

// uniform in (0, 1] from 53 bits of (a:b)
inline double u53(uint32_t a, uint32_t b) {
  const uint64_t x = (((uint64_t)a << 32) | b) >> 11;
  return (double)(x + 1) * (1.0 / 9007199254740992.0);
}

// uniform in (0, 1] from 32 bits
inline double u32d(uint32_t a) { return ((double)a + 1.0) * (1.0 / 4294967296.0); }

// key of the flight spawned at collision `ic` of flight `parent`
// (splitmix64 finaliser over the parent key and the collision index)
inline uint64_t flight_key_child(uint64_t parent, uint32_t ic) {
  uint64_t z = parent ^ ((uint64_t)ic * 0x9E3779B97F4A7C15ull);
  z += 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// 64 random bits for a flight's first draw r1: the splitmix64 finaliser
// (Steele, Lea and Flood, OOPSLA 2014; constants of Vigna's splitmix64.c)
// over the flight key xor (history:seed) times an odd constant, after the
// splitmix64 state increment
inline uint64_t first_draw(uint64_t key, uint32_t seed, uint32_t hist) {
  uint64_t z = key ^ ((((uint64_t)hist << 32) | seed) * 0xD1B54A32D192ED03ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace orc
