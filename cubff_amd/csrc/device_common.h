// Device-side building blocks shared by every cubff kernel (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubff {

constexpr int kTape = 64;            // bytes per program (common.h:55)
constexpr int kTape2 = 2 * kTape;    // bytes per interaction tape
constexpr uint32_t kEvalSteps = 8192;  // common_language.h:184
constexpr uint32_t kNone = 0xffffffffu;

// SplitMix64 finaliser (common_language.h:131-136).
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Epoch seed S(x) = SM(SM(params.seed) ^ SM(x)) (common_language.h:367-369),
// with SM(params.seed) precomputed by the caller.
__host__ __device__ __forceinline__ uint64_t epoch_seed_pre(uint64_t sm_seed,
                                                            uint64_t x) {
  return splitmix64(sm_seed ^ splitmix64(x));
}

// One tape per lane, stored in LDS with a dword-interleaved layout: dword w of
// lane l sits at LDS dword (w * LANES + l).  Every lane therefore always hits
// bank (l mod #banks), whatever byte it reads, so data-dependent accesses from
// 64 lanes never conflict.  Consecutive tape bytes share a dword.
// Pointers into LDS (address space 3): 32-bit addresses, so per-lane byte
// addresses are one 32-bit multiply-add.
typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

template <int LANES>
struct LdsTape {
  static constexpr int kRegionBytes = kTape2 * LANES;
  lds_u8 *p;  // this lane's dword 0

  __device__ __forceinline__ static LdsTape at(uint8_t *region, int lane) {
    return LdsTape{(lds_u8 *)region + lane * 4};
  }
  __device__ __forceinline__ uint32_t off(uint32_t i) const {
    return (i >> 2) * (LANES * 4) + (i & 3);
  }
  __device__ __forceinline__ uint8_t get(uint32_t i) const { return p[off(i)]; }
  __device__ __forceinline__ void set(uint32_t i, uint8_t v) const {
    p[off(i)] = v;
  }
  __device__ __forceinline__ uint32_t word(uint32_t w) const {
    return *(const lds_u32 *)(p + w * (LANES * 4));
  }
  __device__ __forceinline__ void set_word(uint32_t w, uint32_t v) const {
    *(lds_u32 *)(p + w * (LANES * 4)) = v;
  }
};

// Wavefront-aggregated bump allocation: every lane of the (fully active)
// wavefront asks for c slots of *counter and receives the first of a private
// contiguous range; one atomic per wavefront instead of one per lane.
__device__ __forceinline__ uint32_t wave_alloc_u32(uint32_t *counter, uint32_t c) {
  const int lane = threadIdx.x & 63;
  uint32_t incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  uint32_t base = 0;
  if (lane == 63 && incl) base = atomicAdd(counter, incl);
  base = __shfl(base, 63, 64);
  return base + incl - c;
}

// Sum over the 64 lanes of a wavefront.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Soup of a program / slot index of batched soups: base is a prefix-sum
// array of length num_soups + 1.
__device__ __forceinline__ uint32_t soup_of(const uint64_t *base, uint32_t num_soups,
                                            uint64_t x) {
  uint32_t lo = 0, hi = num_soups;  // base[lo] <= x < base[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (base[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

}  // namespace cubff
