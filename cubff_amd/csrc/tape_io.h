// Moving interaction tapes between the soup in HBM and a lane's LDS tape,
// and the per-slot mutation applied on the way in.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

namespace cubff {

// Program bytes of one half of the tape, with 16-byte global accesses.
template <class Tape>
__device__ __forceinline__ void load_row(const Tape &t, int half,
                                         const uint8_t *__restrict__ src) {
  const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    uint4 v = s4[q];
    const int w = half * 16 + q * 4;
    t.set_word(w + 0, v.x);
    t.set_word(w + 1, v.y);
    t.set_word(w + 2, v.z);
    t.set_word(w + 3, v.w);
  }
}

template <class Tape>
__device__ __forceinline__ void store_row(const Tape &t, int half,
                                          uint8_t *__restrict__ dst) {
  uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int w = half * 16 + q * 4;
    d4[q] = make_uint4(t.word(w), t.word(w + 1), t.word(w + 2), t.word(w + 3));
  }
}

// Mutation of common_language.h:172-180: byte b of slot k is replaced by
// r & 0xFF when ((r >> 8) & (2^30 - 1)) < P, r = SM((N*S(e) + k)*128 + b).
template <class Tape>
__device__ __forceinline__ void mutate(const Tape &t, uint64_t n,
                                       uint64_t seed_value, uint64_t k,
                                       uint32_t prob) {
  const uint64_t base = (n * seed_value + k) * (uint64_t)kTape2;
  for (int b = 0; b < kTape2; b++) {
    const uint64_t r = splitmix64(base + b);
    if (((uint32_t)(r >> 8) & 0x3fffffffu) < prob) t.set(b, (uint8_t)r);
  }
}

// The same draws for the bytes of one half only.
template <class Tape>
__device__ __forceinline__ void mutate_half(const Tape &t, int half, uint64_t n,
                                            uint64_t seed_value, uint64_t k, uint32_t prob) {
  const uint64_t base = (n * seed_value + k) * (uint64_t)kTape2;
  for (int b = half * kTape; b < (half + 1) * kTape; b++) {
    const uint64_t r = splitmix64(base + b);
    if (((uint32_t)(r >> 8) & 0x3fffffffu) < prob) t.set(b, (uint8_t)r);
  }
}

}  // namespace cubff
