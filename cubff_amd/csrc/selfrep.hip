// CheckSelfRep (common_language.h:203-277) for the languages BffFast
// interprets (the standard jump logic: bff, bff_noheads, bff_noheads_4bit,
// bff_perm), in levels of command budgets.
//
// A self-rep check is 13 noise trials per program, each a chain of 5
// Evaluate calls (program || noise, then 4 more generations of second half
// || noise); a position is stable when enough trials agree on its final
// byte.  Chain lengths are heavy-tailed: on a random soup the mean chain is
// ~30 executed commands but about 1 in 100 runs to thousands, so a
// wavefront that runs 64 chains to the end lasts as long as its longest
// (~17x the mean).  Here:
//
//   * one lane per chain (all 64 lanes of a wavefront busy); the final tape
//     of every chain goes to an HBM buffer, and a separate kernel counts the
//     stable positions of each program from its 13 final tapes;
//   * level 0 runs every chain with a small command budget per Evaluate; a
//     chain whose Evaluate exceeds it stops, parks the input of that
//     generation (the first half; the noise half is regenerated) in its own
//     slot of the final-tape buffer, and is listed for the next level;
//   * levels 1 and 2 rerun the listed chains from the parked generation with
//     larger budgets, and level 3 runs what is left to the end, with the
//     full-state cycle skip.
//
// Every level reruns a parked Evaluate from its unchanged input, so results
// are those of running each chain straight through: the final tapes, and
// so the counts, equal the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bff_fast.h"
#include "device_common.h"
#include "kernels.h"
#include "languages.h"

namespace cubff {

namespace {

constexpr int kSrLanes = 64;        // one wavefront per block
constexpr int kSrIters = 13;        // kNumIters, common_language.h:207
constexpr int kSrGens = 4;          // kNumExtraGens, :208
constexpr int kSrLevels = 4;
// Command budget of one Evaluate at levels 0..2 (level 3 has none).  Chosen
// on the measured chain-length distribution of random bff_noheads soups:
// ~3% of chains pass level 0, ~2% level 1, ~0.1% level 2.
constexpr uint32_t kSrBudget[kSrLevels - 1] = {12, 96, 768};

struct SrArgs {
  const uint8_t *soup;
  uint64_t n;           // soup size (the noise seed uses it)
  uint64_t q0;          // first program of this chunk
  uint64_t total;       // chains to run at this level
  uint64_t seed_value;
  uint8_t *finals;      // [chunk programs x 13][128]
  const uint32_t *list_in;  // chains listed by the previous level (null: all)
  uint32_t *list_out;
  uint32_t *count_out;
  uint32_t budget;
};

// Noise of trial `it` of program q (common_language.h:213-220): byte j is
// SM(local ^ SM((it + 1) * 64 + j)) & 0xFF with local = SM(n * seed + q).
__device__ __forceinline__ void sr_noise(uint64_t n, uint64_t seed_value, uint64_t q,
                                         uint32_t it, uint32_t *noise) {
  const uint64_t local = splitmix64(n * seed_value + q);
#pragma unroll
  for (int w = 0; w < kTape / 4; w++) {
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < 4; b++)
      v |= (uint32_t)(splitmix64(local ^ splitmix64((uint64_t)(it + 1) * kTape + w * 4 + b)) &
                      0xFF)
           << (8 * b);
    noise[w] = v;
  }
}

template <int LANG, bool kLast>
__global__ __launch_bounds__(kSrLanes) void k_sr_run(SrArgs a) {
  using L = typename LangOf<LANG>::T;
  using F = typename FastOf<LANG>::T;
  using Tape = typename L::template Tape<kSrLanes>;
  constexpr int kNx = (Tape::kRegionBytes + 15) / 16 * 16;
  constexpr int kAnchor = kNx + (LinTape<kSrLanes>::kRegionBytes + 15) / 16 * 16;
  __shared__ __align__(16) uint8_t lds[kAnchor + (kLast ? LdsTape<kSrLanes>::kRegionBytes : 0)];
  __shared__ uint8_t lut[256];
  fill_fast_lut<LANG>(lut);
  __syncthreads();
  const int lane = threadIdx.x;
  const uint64_t i = (uint64_t)blockIdx.x * kSrLanes + lane;
  bool deferred = false;
  uint32_t c = 0;
  if (i < a.total) {
    c = a.list_in ? a.list_in[i] : (uint32_t)i;
    const uint64_t q = a.q0 + c / kSrIters;
    const uint32_t it = c % kSrIters;
    uint32_t noise[kTape / 4], first[kTape / 4];
    sr_noise(a.n, a.seed_value, q, it, noise);
    uint4 *row = reinterpret_cast<uint4 *>(a.finals + (uint64_t)c * kTape2);
    uint32_t g = 0;
    // The input first half of generation g: the program itself, or the
    // parked half of a chain listed by the previous level.
    const uint4 *src = a.list_in ? row : reinterpret_cast<const uint4 *>(a.soup + q * kTape);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint4 v = src[k];
      first[4 * k] = v.x;
      first[4 * k + 1] = v.y;
      first[4 * k + 2] = v.z;
      first[4 * k + 3] = v.w;
    }
    if (a.list_in) g = row[4].x;
    const auto t = Tape::at(lds, lane);
    const auto nx = LinTape<kSrLanes>::at(lds + kNx, lane);
    for (; g <= (uint32_t)kSrGens; g++) {
#pragma unroll
      for (int w = 0; w < kTape / 4; w++) {
        t.set_word(w, first[w]);
        t.set_word(kTape / 4 + w, noise[w]);
      }
      uint32_t steps;
      if constexpr (kLast) {
        F::evaluate_full(t, nx, LdsTape<kSrLanes>::at(lds + kAnchor, lane),
                         LdsByteLut{(const lds_u8 *)lut}, kEvalSteps, &steps, nullptr);
      } else {
        F::evaluate(t, nx, LdsByteLut{(const lds_u8 *)lut}, kEvalSteps, &steps, nullptr,
                    a.budget);
        if (steps == kOverBudget) {
          deferred = true;
          break;
        }
      }
      if (g < (uint32_t)kSrGens) {
#pragma unroll
        for (int w = 0; w < kTape / 4; w++) first[w] = t.word(kTape / 4 + w);
      }
    }
    if (deferred) {
#pragma unroll
      for (int k = 0; k < 4; k++)
        row[k] = make_uint4(first[4 * k], first[4 * k + 1], first[4 * k + 2], first[4 * k + 3]);
      row[4] = make_uint4(g, 0, 0, 0);
    } else {
#pragma unroll
      for (int k = 0; k < kTape2 / 16; k++)
        row[k] = make_uint4(t.word(4 * k), t.word(4 * k + 1), t.word(4 * k + 2),
                            t.word(4 * k + 3));
    }
  }
  if constexpr (!kLast) {
    const uint32_t at = wave_alloc_u32(a.count_out, deferred ? 1u : 0u);
    if (deferred) a.list_out[at] = c;
  }
}

// Byte stability (common_language.h:259-276): position i of program q is
// stable when some trial a (matching the original byte, for the first half)
// agrees with at least 3 later trials; the result is the smaller of the two
// halves' stable counts.  32 threads per program, 4 positions per thread
// (one 32-bit load from each trial's final tape); lanes 0-15 of the group
// hold the first half, 16-31 the second.
constexpr int kCountProgs = 8;  // programs per 256-thread block
__global__ __launch_bounds__(256) void k_sr_count(const uint8_t *__restrict__ soup,
                                                  const uint8_t *__restrict__ finals, uint64_t q0,
                                                  uint64_t progs, uint64_t *__restrict__ result) {
  const uint64_t ql = (uint64_t)blockIdx.x * kCountProgs + threadIdx.x / 32;
  const int pos4 = threadIdx.x % 32;
  uint32_t cnt = 0;
  if (ql < progs) {
    const uint32_t *f = reinterpret_cast<const uint32_t *>(finals + ql * kSrIters * kTape2) + pos4;
    uint32_t v[kSrIters];
#pragma unroll
    for (int k = 0; k < kSrIters; k++) v[k] = f[k * (kTape2 / 4)];
    const bool in_first = pos4 < kTape / 4;
    const uint32_t orig =
        in_first ? reinterpret_cast<const uint32_t *>(soup + (q0 + ql) * kTape)[pos4] : 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int sh = 8 * b;
      bool stable = false;
      for (int x = 0; x < kSrIters && !stable; x++) {
        const uint32_t vx = (v[x] >> sh) & 0xFF;
        if (in_first && vx != ((orig >> sh) & 0xFF)) continue;
        int c = 1;
#pragma unroll
        for (int y = 0; y < kSrIters; y++) c += (y > x) && vx == ((v[y] >> sh) & 0xFF);
        stable = c > kSrIters / 4;
      }
      cnt += stable;
    }
  }
  // Sums over each half (16 lanes), then the smaller of the two halves.
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  const uint32_t other = __shfl_xor(cnt, 16, 64);
  if (ql < progs && pos4 == 0) result[q0 + ql] = std::min(cnt, other);
}

template <int LANG>
hipError_t run_levels(hipStream_t st, const uint8_t *soup, uint64_t n, uint64_t count,
                      uint64_t seed_value, uint64_t *result, const SrScratch &w) {
  hipError_t e;
  for (uint64_t q0 = 0; q0 < count; q0 += w.chunk) {
    const uint64_t progs = std::min(w.chunk, count - q0);
    SrArgs a{soup, n, q0, progs * kSrIters, seed_value, w.finals, nullptr,
             w.list[0], w.counts, kSrBudget[0]};
    for (int level = 0; level < kSrLevels && a.total; level++) {
      const bool last = level == kSrLevels - 1;
      if (!last) {
        a.list_out = w.list[level & 1];
        a.count_out = w.counts + (level & 1);
        a.budget = kSrBudget[level];
        if ((e = launch_fill_u32(st, a.count_out, 0, 1)) != hipSuccess) return e;
      }
      const dim3 grid((unsigned)((a.total + kSrLanes - 1) / kSrLanes));
      if (last)
        hipLaunchKernelGGL((k_sr_run<LANG, true>), grid, dim3(kSrLanes), 0, st, a);
      else
        hipLaunchKernelGGL((k_sr_run<LANG, false>), grid, dim3(kSrLanes), 0, st, a);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if (last) break;
      // The next level runs the chains this one listed.
      uint32_t listed = 0;
      if ((e = hipMemcpyAsync(&listed, a.count_out, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
          (e = hipStreamSynchronize(st)) != hipSuccess)
        return e;
      if (listed > a.total) return hipErrorIllegalState;
      a.list_in = a.list_out;
      a.total = listed;
    }
    hipLaunchKernelGGL(k_sr_count, dim3((unsigned)((progs + kCountProgs - 1) / kCountProgs)),
                       dim3(256), 0, st, soup, w.finals, q0, progs, result);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

bool selfrep_levels_supported(int lang) {
  return lang == kBff || lang == kBffNoheads || lang == kBffNoheads4bit || lang == kBffPerm;
}

hipError_t launch_selfrep_levels(hipStream_t st, int lang, const uint8_t *soup, uint64_t n,
                                 uint64_t count, uint64_t seed_value, uint64_t *result,
                                 const SrScratch &w) {
  if (count == 0) return hipSuccess;
  if (!w.finals || !w.list[0] || !w.list[1] || !w.counts || w.chunk == 0 ||
      w.chunk * kSrIters > 0xffffffffULL)
    return hipErrorInvalidValue;
  switch (lang) {
    case kBff: return run_levels<kBff>(st, soup, n, count, seed_value, result, w);
    case kBffNoheads: return run_levels<kBffNoheads>(st, soup, n, count, seed_value, result, w);
    case kBffNoheads4bit:
      return run_levels<kBffNoheads4bit>(st, soup, n, count, seed_value, result, w);
    case kBffPerm: return run_levels<kBffPerm>(st, soup, n, count, seed_value, result, w);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cubff
