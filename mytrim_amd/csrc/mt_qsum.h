// Checksum of a queue entry and its lap, accumulated one field at a time in
// two 32-bit words mixed by additions, rotations and exclusive-ors: every
// step is a full-rate vector instruction (a 64-bit multiply costs three
// quarter-rate ones, and an entry is summed by its writer and its reader).
// A reader recomputes it over the fields it loaded; a mismatch means some
// field is not yet the writer's, and it loads again.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mt {

struct QSum {
  uint32_t a, b;
};
__host__ __device__ __forceinline__ QSum sum_init(uint32_t lap) {
  return QSum{0x9E3779B9u ^ lap, 0x7F4A7C15u + (lap << 16)};
}
__host__ __device__ __forceinline__ QSum sum_add(QSum h, uint64_t w) {
  h.a += (uint32_t)w;
  h.b ^= (uint32_t)(w >> 32);
  h.a = __builtin_rotateleft32(h.a, 7) ^ h.b;
  h.b = __builtin_rotateleft32(h.b, 13) + h.a;
  return h;
}
__host__ __device__ __forceinline__ uint32_t sum_end(QSum h) {
  const uint32_t a = h.a + __builtin_rotateleft32(h.b, 16);
  return a ^ __builtin_rotateleft32(a, 11) ^ h.b;
}

}  // namespace mt
