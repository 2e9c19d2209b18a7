// Batched soups: many independent soups of uneven size, seed, mutation rate
// and language advanced together, one set of launches per epoch.
//
// Per epoch:
//   k_b_fy_*      Fisher-Yates pairing of every soup (common_language.h:418-424),
//                 the exact parallel form of kernels.hip over the concatenated
//                 program index space
//   k_b_stage     gather + mutation (common_language.h:166-180) into a staged
//                 128-byte tape per pair slot
//   k_b_interp<L> evaluate staged tapes in the order of `order` (identity or
//                 longest-predicted first), scatter the halves back
//                 (common_language.h:184-192), record per-slot ops/steps
//   k_b_end_epoch per-soup counter sums, num_runs, epoch counter
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batch.h"
#include "device_common.h"
#include "languages.h"

namespace cubff {

// A workgroup's copy of a prefix-sum array in LDS when it has at most
// kBaseTable entries (else the array itself): the per-soup lookups of every
// thread then hit LDS.  Every thread of the block must call it.
constexpr uint32_t kBaseTable = 256;
__device__ __forceinline__ const uint64_t *block_bases(const uint64_t *__restrict__ base,
                                                       uint32_t num_soups, uint64_t *lds) {
  if (num_soups + 1 > kBaseTable) return base;
  for (uint32_t i = threadIdx.x; i <= num_soups; i += blockDim.x) lds[i] = base[i];
  __syncthreads();
  return lds;
}

// ---------------------------------------------------------------------------
// Pairing.

__global__ void k_b_fy_draw(const BatchDev d, uint32_t ahead) {
  __shared__ uint64_t pbase[kBaseTable];
  const uint64_t *pb = block_bases(d.prog_base, d.num_soups, pbase);
  const uint64_t epoch = *d.epoch + ahead;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.total_programs) return;
  const uint32_t s = soup_of(pb, d.num_soups, t);
  const uint64_t b = pb[s];
  const uint64_t n = pb[s + 1] - b;
  const uint64_t tl = t - b;
  const uint64_t r = splitmix64(epoch_seed_pre(d.sm_seed[s], epoch * n + tl));
  const uint32_t j = (uint32_t)(b + r % (tl + 1));
  d.fy_j[t] = j;
  atomicAdd(&d.fy_cnt[j], 1u);
}

// Contiguous item ranges per bucket: a block-wide scan of the counts and
// one global atomic per block of 256 (small blocks: the pairing of the next
// epoch runs beside the interpreter and must fit between its waves).
constexpr int kAllocBlock = 256;
__global__ __launch_bounds__(kAllocBlock) void k_b_fy_alloc(const BatchDev d) {
  __shared__ uint32_t wtot[kAllocBlock / 64], bbase;
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = p < d.total_programs;
  const uint32_t c = in ? d.fy_cnt[p] : 0u;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wtot[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int i = 0; i < kAllocBlock / 64; i++) {
      const uint32_t x = wtot[i];
      wtot[i] = run;
      run += x;
    }
    bbase = run ? atomicAdd(&d.fy_cnt[d.total_programs], run) : 0u;
  }
  __syncthreads();
  if (!in) return;
  d.fy_base[p] = bbase + wtot[w] + incl - c;
  d.fy_cursor[p] = 0;
}

__global__ void k_b_fy_place(const BatchDev d) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.total_programs) return;
  const uint32_t j = d.fy_j[t];
  d.fy_items[d.fy_base[j] + atomicAdd(&d.fy_cursor[j], 1u)] = (uint32_t)t;
}

__global__ void k_b_fy_sort(const BatchDev d) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.total_programs) return;
  const uint32_t c = d.fy_cnt[p];
  if (c == 0) {
    d.fy_first[p] = kNone;
    return;
  }
  uint32_t *b = d.fy_items + d.fy_base[p];
  for (uint32_t x = 1; x < c; x++) {
    const uint32_t v = b[x];
    uint32_t y = x;
    for (; y > 0 && b[y - 1] > v; y--) b[y] = b[y - 1];
    b[y] = v;
  }
  d.fy_first[p] = b[0];
  for (uint32_t x = 0; x < c; x++) d.fy_succ[b[x]] = x + 1 < c ? b[x + 1] : kNone;
}

__device__ __forceinline__ uint32_t b_fy_in(const BatchDev &d, uint32_t t) {
  for (;;) {
    const uint32_t a = d.fy_j[t] == t ? d.fy_succ[t] : d.fy_first[t];
    if (a == kNone) return t;
    t = a;
  }
}

// Writes the pairing as local (per-soup) program indices, like the
// reference's shuffle_idx.
__global__ void k_b_fy_final(const BatchDev d, uint32_t ahead) {
  __shared__ uint64_t pbase[kBaseTable];
  const uint64_t *pb = block_bases(d.prog_base, d.num_soups, pbase);
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.total_programs) return;
  const uint32_t j = d.fy_j[t];
  uint32_t v;
  if (j == (uint32_t)t) {
    v = b_fy_in(d, (uint32_t)t);
  } else {
    const uint32_t c = d.fy_succ[t];
    v = c == kNone ? j : b_fy_in(d, c);
  }
  const uint32_t s = soup_of(pb, d.num_soups, t);
  const uint64_t e = *d.epoch + ahead;
  d.shuf[e % kPairBufs][t] = (uint32_t)(v - pb[s]);
  d.inv[e % kPairBufs][v] = (uint32_t)t;
}

__global__ void k_b_fill(uint32_t *__restrict__ p, uint32_t v, uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// ---------------------------------------------------------------------------
// Stage: one thread per 16 bytes of a slot's tape.  Thread q of slot g copies
// bytes [16q, 16q + 16) of program s[2k] (q < 4) or s[2k+1] (q >= 4) and
// applies the mutation draws of those positions.
__global__ void k_b_stage(const BatchDev d) {
  __shared__ uint64_t pbase[kBaseTable], sbase[kBaseTable];
  const uint64_t *prog_base = block_bases(d.prog_base, d.num_soups, pbase);
  const uint64_t *slot_base = block_bases(d.slot_base, d.num_soups, sbase);
  const uint64_t epoch = *d.epoch;
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= d.total_slots * 8) return;
  const uint64_t g = u >> 3;
  const uint32_t q = (uint32_t)(u & 7);
  const uint32_t s = soup_of(slot_base, d.num_soups, g);
  const uint64_t k = g - slot_base[s];
  const uint64_t pb = prog_base[s];
  const uint64_t n = prog_base[s + 1] - pb;
  const uint32_t p = d.shuf[epoch % kPairBufs][pb + 2 * k + (q >> 2)];
  uint4 v = reinterpret_cast<const uint4 *>(d.soup + (pb + p) * kTape)[q & 3];
  uint32_t w[4] = {v.x, v.y, v.z, v.w};
  const uint64_t seed_value = epoch_seed_pre(d.sm_seed[s], epoch);
  const uint32_t prob = d.prob[s];
  const uint64_t base = (n * seed_value + k) * (uint64_t)kTape2 + q * 16;
#pragma unroll
  for (int b = 0; b < 16; b++) {
    const uint64_t r = splitmix64(base + b);
    if (((uint32_t)(r >> 8) & 0x3fffffffu) < prob) {
      const int sh = 8 * (b & 3);
      w[b >> 2] = (w[b >> 2] & ~(0xffu << sh)) | ((uint32_t)(r & 0xff) << sh);
    }
  }
  reinterpret_cast<uint4 *>(d.stage + g * kTape2)[q] = make_uint4(w[0], w[1], w[2], w[3]);
  if ((q & 3) == 0) {  // q=0: s[2k], q=4: s[2k+1]
    const uint64_t h = 2 * g + (q >> 2);  // = pb + 2k + (q >> 2)
    d.tgt[h] = (uint32_t)(pb + p);
    // The program at the same position of the next pairing is staged now
    // in half inv_e[program] (the next pairing is drawn two epochs ahead).
    if (d.halves_next)
      d.halves_next[h] = d.inv[epoch % kPairBufs][pb + d.shuf[(epoch + 1) % kPairBufs][h]];
  }
}

__global__ void k_b_copy(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src,
                         uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Interpreter: lane `lane_g` evaluates one list entry in each of the ranges
// [rL, rL + L) (L = lanes in the grid), T ranges at most: rL + lane_g in
// even rounds, rL + L - 1 - lane_g in odd ones.  Under the learned order
// (longest predicted first) the lanes that ran the longest pairs of one
// round then run the shortest of the next.
template <int LANG, int LANES>
__global__ __launch_bounds__(LANES) void k_b_interp(const BatchDev d,
                                                    const uint32_t *__restrict__ list,
                                                    uint64_t count, uint32_t per_lane) {
  using L = typename LangOf<LANG>::T;
  using Tape = typename L::template Tape<LANES>;
  using Aux = typename L::template Aux<LANES>;
  constexpr int kAuxBytes = lang_needs_aux(LANG) ? Aux::kRegionBytes : 0;
  __shared__ __align__(16) uint8_t lds[Tape::kRegionBytes + kAuxBytes];
  const int lane = threadIdx.x;
  const auto t = Tape::at(lds, lane);
  const auto aux = Aux::at(lds + Tape::kRegionBytes, lane);
  const uint64_t lanes_total = (uint64_t)gridDim.x * LANES;
  const uint64_t lane_g = (uint64_t)blockIdx.x * LANES + lane;
  for (uint32_t r = 0; r < per_lane; r++) {
    const uint64_t e = r * lanes_total + ((r & 1) ? lanes_total - 1 - lane_g : lane_g);
    if (e >= count) break;
    const uint32_t g = list[e];
    const uint4 *src = reinterpret_cast<const uint4 *>(d.stage + (uint64_t)g * kTape2);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const uint4 v = src[q];
      t.set_word(q * 4 + 0, v.x);
      t.set_word(q * 4 + 1, v.y);
      t.set_word(q * 4 + 2, v.z);
      t.set_word(q * 4 + 3, v.w);
    }
    const uint32_t s = soup_of(d.slot_base, d.num_soups, g);
    const uint64_t k = g - d.slot_base[s];
    uint32_t ops = 0, steps = 0, work = 0;
    if (k < d.num_indices[s]) ops = L::evaluate(t, aux, kEvalSteps, &steps, &work);
    const uint32_t p1 = d.tgt[2 * (uint64_t)g], p2 = d.tgt[2 * (uint64_t)g + 1];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      uint4 *dst = reinterpret_cast<uint4 *>(d.soup + (uint64_t)(h ? p2 : p1) * kTape);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int w = h * 16 + q * 4;
        dst[q] = make_uint4(t.word(w), t.word(w + 1), t.word(w + 2), t.word(w + 3));
      }
    }
    d.pair_ops[g] = ops;
    d.pair_steps[g] = steps;
    d.pair_work[g] = work;
  }
}

// Per-soup interval counters (common_language.h:192, :494): kEndSplit
// blocks per soup each sum a share of its slots' ops and steps and add them
// atomically; the first also adds num_runs += num_indices, and block 0
// advances the epoch.
constexpr int kEndSplit = 16;
__global__ __launch_bounds__(256) void k_b_end_epoch(const BatchDev d) {
  __shared__ unsigned long long part[2][4];
  const uint32_t s = blockIdx.x / kEndSplit, piece = blockIdx.x % kEndSplit;
  const uint64_t lo = d.slot_base[s], n = d.slot_base[s + 1] - lo;
  const uint64_t b0 = lo + n * piece / kEndSplit, b1 = lo + n * (piece + 1) / kEndSplit;
  unsigned long long ops = 0, steps = 0;
  for (uint64_t g = b0 + threadIdx.x; g < b1; g += blockDim.x) {
    ops += d.pair_ops[g];
    steps += d.pair_steps[g];
  }
  ops = wave_sum_u64(ops);
  steps = wave_sum_u64(steps);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][w] = ops;
    part[1][w] = steps;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long *c = d.counters + kNumBatchCounters * (uint64_t)s;
    const unsigned long long o = part[0][0] + part[0][1] + part[0][2] + part[0][3];
    const unsigned long long t = part[1][0] + part[1][1] + part[1][2] + part[1][3];
    if (o) atomicAdd(&c[0], o);
    if (t) atomicAdd(&c[2], t);
    if (piece == 0) atomicAdd(&c[1], (unsigned long long)d.num_indices[s]);
    if (blockIdx.x == 0) *d.epoch += 1;
  }
}

// ---------------------------------------------------------------------------
// Host launchers.

namespace {
inline unsigned nblk(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
}  // namespace

hipError_t batch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, uint64_t count) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_b_fill, dim3(nblk(count, 256)), dim3(256), 0, st, p, v, count);
  return hipGetLastError();
}

hipError_t batch_pair_shuffle(hipStream_t st, const BatchDev &d, uint32_t ahead) {
  const uint64_t n = d.total_programs;
  const unsigned g = nblk(n, 256);
  hipLaunchKernelGGL(k_b_fill, dim3(nblk(n + 1, 256)), dim3(256), 0, st, d.fy_cnt, 0u, n + 1);
  hipLaunchKernelGGL(k_b_fy_draw, dim3(g), dim3(256), 0, st, d, ahead);
  hipLaunchKernelGGL(k_b_fy_alloc, dim3(nblk(n, kAllocBlock)), dim3(kAllocBlock), 0, st, d);
  hipLaunchKernelGGL(k_b_fy_place, dim3(g), dim3(256), 0, st, d);
  hipLaunchKernelGGL(k_b_fy_sort, dim3(g), dim3(256), 0, st, d);
  hipLaunchKernelGGL(k_b_fy_final, dim3(g), dim3(256), 0, st, d, ahead);
  return hipGetLastError();
}

hipError_t batch_stage(hipStream_t st, const BatchDev &d) {
  hipLaunchKernelGGL(k_b_stage, dim3(nblk(d.total_slots * 8, 256)), dim3(256), 0, st, d);
  return hipGetLastError();
}

hipError_t batch_copy_u32(hipStream_t st, uint32_t *dst, const uint32_t *src, uint64_t count) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_b_copy, dim3(nblk(count, 256)), dim3(256), 0, st, dst, src, count);
  return hipGetLastError();
}

hipError_t batch_end_epoch(hipStream_t st, const BatchDev &d) {
  hipLaunchKernelGGL(k_b_end_epoch, dim3(d.num_soups * kEndSplit), dim3(256), 0, st, d);
  return hipGetLastError();
}

template <int LANES>
static hipError_t interp_lanes(hipStream_t st, int lang, const BatchDev &d,
                               const uint32_t *list, uint64_t count, uint32_t per_lane) {
  const uint64_t lanes = (count + per_lane - 1) / per_lane;
  const dim3 grid(nblk(lanes, LANES)), block(LANES);
  switch (lang) {
#define CUBFF_B_CASE(LG) \
  case LG: hipLaunchKernelGGL((k_b_interp<LG, LANES>), grid, block, 0, st, d, list, count, per_lane); break;
    CUBFF_B_CASE(kBff)
    CUBFF_B_CASE(kBffNoheads)
    CUBFF_B_CASE(kBff8)
    CUBFF_B_CASE(kBff8Noheads)
    CUBFF_B_CASE(kBffNoheads4bit)
    CUBFF_B_CASE(kBffPerm)
    CUBFF_B_CASE(kBffSelfmove)
    CUBFF_B_CASE(kForth)
    CUBFF_B_CASE(kForthcopy)
    CUBFF_B_CASE(kForthtrivial)
    CUBFF_B_CASE(kForthtrivialReset)
    CUBFF_B_CASE(kSubleq)
    CUBFF_B_CASE(kRsubleq4)
#undef CUBFF_B_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t batch_interp(hipStream_t st, int lang, const BatchDev &d, const uint32_t *list,
                        uint64_t count, int wg, uint32_t per_lane) {
  if (count == 0) return hipSuccess;
  if (per_lane == 0) per_lane = 1;
  switch (wg) {
    case 64: return interp_lanes<64>(st, lang, d, list, count, per_lane);
    case 128: return interp_lanes<128>(st, lang, d, list, count, per_lane);
    case 256: return interp_lanes<256>(st, lang, d, list, count, per_lane);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cubff
