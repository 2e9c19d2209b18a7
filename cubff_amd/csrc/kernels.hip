// HIP kernels of the cubff per-epoch hot path (gfx950).
//
//   k_init            InitPrograms            common_language.h:138-155
//   k_epoch<L>        MutateAndRunPrograms    common_language.h:157-193
//   k_selfrep<L>      CheckSelfRep            common_language.h:203-277
//   k_eval_tapes<L>   RunOneProgram (batched) common_language.h:195-201
//   pairing kernels   host pairing block      common_language.h:418-486
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "kernels.h"
#include "languages.h"
#include "tape_io.h"

namespace cubff {

// ---------------------------------------------------------------------------
// InitPrograms: byte j of program i = SM(64*N*seed + 64*i + j) & 0xFF.  One
// thread per 16 bytes, written with one 16-byte store.
__global__ void k_init(uint8_t *__restrict__ soup, uint64_t n,
                       uint64_t seed_value, uint64_t count, int zero_init) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // 16-byte unit
  if (q >= count * (kTape / 16)) return;
  uint32_t w[4];
  if (zero_init) {
    w[0] = w[1] = w[2] = w[3] = 0;
  } else {
    const uint64_t base = (uint64_t)kTape * n * seed_value + q * 16;
#pragma unroll
    for (int d = 0; d < 4; d++) {
      uint32_t v = 0;
#pragma unroll
      for (int b = 0; b < 4; b++)
        v |= (uint32_t)(splitmix64(base + d * 4 + b) & 0xFF) << (8 * b);
      w[d] = v;
    }
  }
  reinterpret_cast<uint4 *>(soup)[q] = make_uint4(w[0], w[1], w[2], w[3]);
}

// ---------------------------------------------------------------------------
// Epoch kernel.  One lane per pair slot k: gather programs s[2k], s[2k+1]
// into the lane's LDS tape with 16-byte loads, apply the slot's 128 mutation
// draws, evaluate if k < num_indices, scatter with 16-byte stores, record
// per-slot ops/steps and add the block's ops to the interval counter.

template <int LANG>
__global__ __launch_bounds__(kEpochLanes) void k_epoch(
    uint8_t *__restrict__ soup, const uint32_t *__restrict__ shuf,
    uint64_t n, const uint64_t *__restrict__ num_indices_p,
    uint64_t seed_value, uint32_t prob, uint32_t *__restrict__ pair_ops,
    uint32_t *__restrict__ pair_steps, unsigned long long *__restrict__ counters) {
  using L = typename LangOf<LANG>::T;
  using Tape = typename L::template Tape<kEpochLanes>;
  using Aux = typename L::template Aux<kEpochLanes>;
  constexpr int kAuxBytes = lang_needs_aux(LANG) ? Aux::kRegionBytes : 0;
  __shared__ __align__(16) uint8_t lds[Tape::kRegionBytes + kAuxBytes];
  __shared__ unsigned long long block_ops, block_steps;
  const int lane = threadIdx.x;
  if (lane == 0) {
    block_ops = 0;
    block_steps = 0;
  }
  const uint64_t k = (uint64_t)blockIdx.x * kEpochLanes + lane;
  const uint64_t slots = n / 2;
  const auto t = Tape::at(lds, lane);
  const auto aux = Aux::at(lds + Tape::kRegionBytes, lane);
  uint32_t ops = 0, steps = 0;
  __syncthreads();
  if (k < slots) {
    const uint32_t p1 = shuf[2 * k], p2 = shuf[2 * k + 1];
    load_row(t, 0, soup + (uint64_t)p1 * kTape);
    load_row(t, 1, soup + (uint64_t)p2 * kTape);
    mutate(t, n, seed_value, k, prob);
    if (k < *num_indices_p) ops = L::evaluate(t, aux, kEvalSteps, &steps);
    store_row(t, 0, soup + (uint64_t)p1 * kTape);
    store_row(t, 1, soup + (uint64_t)p2 * kTape);  // p2's half wins if p1 == p2
    pair_ops[k] = ops;
    pair_steps[k] = steps;
  }
  const uint64_t wsum = wave_sum_u64(ops);
  const uint64_t wsteps = wave_sum_u64(steps);
  if ((lane & 63) == 0 && wsum) atomicAdd(&block_ops, (unsigned long long)wsum);
  if ((lane & 63) == 0 && wsteps) atomicAdd(&block_steps, (unsigned long long)wsteps);
  __syncthreads();
  if (lane == 0 && block_ops) atomicAdd(&counters[0], block_ops);
  if (lane == 0 && block_steps) atomicAdd(&counters[2], block_steps);
  if (blockIdx.x == 0 && lane == 0) atomicAdd(&counters[1], (unsigned long long)*num_indices_p);
}

// ---------------------------------------------------------------------------
// CheckSelfRep.  A block scores kSelfrepProgs programs; lane (q, it) runs
// noise trial `it` of program q: Evaluate(prog || noise), then 4 more
// generations of (second half || noise).  Final tapes stay in LDS for the
// cross-trial byte-stability count.
constexpr int kIters = 13;   // kNumIters, common_language.h:207
constexpr int kGens = 4;     // kNumExtraGens, :208

template <int LANG>
__global__ __launch_bounds__(kSelfrepLanes) void k_selfrep(
    const uint8_t *__restrict__ soup, uint64_t n, uint64_t count,
    uint64_t seed_value, uint64_t *__restrict__ result, uint8_t *__restrict__ trace) {
  using L = typename LangOf<LANG>::T;
  constexpr int kLanes = kSelfrepLanes;  // kSelfrepProgs * 13 (+ padding)
  using Tape = typename L::template Tape<kLanes>;
  using Aux = typename L::template Aux<kLanes>;
  constexpr int kAuxBytes = lang_needs_aux(LANG) ? Aux::kRegionBytes : 0;
  __shared__ __align__(16) uint8_t lds[Tape::kRegionBytes + kAuxBytes];
  __shared__ uint32_t res[kSelfrepProgs][2];
  const int lane = threadIdx.x;
  const int q = lane / kIters, it = lane % kIters;
  const uint64_t idx = (uint64_t)blockIdx.x * kSelfrepProgs + q;
  const auto t = Tape::at(lds, lane);
  const auto aux = Aux::at(lds + Tape::kRegionBytes, lane);
  if (lane < kSelfrepProgs * 2) res[lane / 2][lane % 2] = 0;
  const bool active = q < kSelfrepProgs && idx < count;
  if (active) {
    const uint64_t local = splitmix64(n * seed_value + idx);
    uint32_t noise[kTape / 4];
#pragma unroll
    for (int w = 0; w < kTape / 4; w++) {
      uint32_t v = 0;
#pragma unroll
      for (int b = 0; b < 4; b++)
        v |= (uint32_t)(splitmix64(local ^ splitmix64((uint64_t)(it + 1) * kTape +
                                                      w * 4 + b)) & 0xFF)
             << (8 * b);
      noise[w] = v;
    }
    const uint32_t *prog = reinterpret_cast<const uint32_t *>(soup + idx * kTape);
#pragma unroll
    for (int w = 0; w < kTape / 4; w++) {
      t.set_word(w, prog[w]);
      t.set_word(16 + w, noise[w]);
    }
    // Optional record of program 0's tapes before and after every
    // generation, [trial][generation][before, after][128] (the states the
    // reference prints under `debug`, common_language.h:225-255).
    uint32_t *rec = trace && idx == 0
                        ? reinterpret_cast<uint32_t *>(trace + (uint64_t)it * (kGens + 1) * 2 * kTape2)
                        : nullptr;
    auto keep = [&](int slot) {
      if (rec)
        for (int w = 0; w < kTape2 / 4; w++) rec[slot * (kTape2 / 4) + w] = t.word(w);
    };
    uint32_t steps;
    keep(0);
    L::evaluate(t, aux, kEvalSteps, &steps);
    keep(1);
    for (int g = 0; g < kGens; g++) {
#pragma unroll
      for (int w = 0; w < kTape / 4; w++) {
        t.set_word(w, t.word(16 + w));
        t.set_word(16 + w, noise[w]);
      }
      keep(2 * g + 2);
      L::evaluate(t, aux, kEvalSteps, &steps);
      keep(2 * g + 3);
    }
  }
  __syncthreads();
  // Byte-stability count (common_language.h:259-276): position i of program
  // q is stable when some trial a (matching the original byte, for the first
  // half) agrees with at least 3 later trials.
  for (int item = lane; item < kSelfrepProgs * kTape2; item += kLanes) {
    const int qq = item / kTape2, i = item % kTape2;
    const uint64_t pidx = (uint64_t)blockIdx.x * kSelfrepProgs + qq;
    if (pidx >= count) continue;
    const uint8_t orig = i < kTape ? soup[pidx * kTape + i] : 0;
    uint8_t v[kIters];
#pragma unroll
    for (int a = 0; a < kIters; a++)
      v[a] = Tape::at(lds, qq * kIters + a).get(i);
    for (int a = 0; a < kIters; a++) {
      if (i < kTape && v[a] != orig) continue;
      int c = 1;
      for (int b = a + 1; b < kIters; b++) c += v[a] == v[b];
      if (c > kIters / 4) {
        atomicAdd(&res[qq][i / kTape], 1u);
        break;
      }
    }
  }
  __syncthreads();
  if (lane < kSelfrepProgs) {
    const uint64_t pidx = (uint64_t)blockIdx.x * kSelfrepProgs + lane;
    if (pidx < count) result[pidx] = min(res[lane][0], res[lane][1]);
  }
}

// ---------------------------------------------------------------------------
// Batched single-tape evaluation: lane per tape, tapes packed 128 B apart.
template <int LANG>
__global__ __launch_bounds__(kEpochLanes) void k_eval_tapes(
    uint8_t *__restrict__ tapes, uint64_t count, uint32_t stepcount,
    uint32_t *__restrict__ ops_out, uint32_t *__restrict__ steps_out) {
  using L = typename LangOf<LANG>::T;
  using Tape = typename L::template Tape<kEpochLanes>;
  using Aux = typename L::template Aux<kEpochLanes>;
  constexpr int kAuxBytes = lang_needs_aux(LANG) ? Aux::kRegionBytes : 0;
  __shared__ __align__(16) uint8_t lds[Tape::kRegionBytes + kAuxBytes];
  const int lane = threadIdx.x;
  const uint64_t k = (uint64_t)blockIdx.x * kEpochLanes + lane;
  if (k >= count) return;
  const auto t = Tape::at(lds, lane);
  const auto aux = Aux::at(lds + Tape::kRegionBytes, lane);
  load_row(t, 0, tapes + k * kTape2);
  load_row(t, 1, tapes + k * kTape2 + kTape);
  uint32_t steps = 0;
  const uint32_t ops = L::evaluate(t, aux, stepcount, &steps);
  store_row(t, 0, tapes + k * kTape2);
  store_row(t, 1, tapes + k * kTape2 + kTape);
  ops_out[k] = ops;
  steps_out[k] = steps;
}

// ---------------------------------------------------------------------------
// Pairing.

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, uint64_t count) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

__global__ void k_set_u64(uint64_t *__restrict__ p, uint64_t v) { *p = v; }

// --fixed_shuffle (common_language.h:468-474).
__global__ void k_pair_fixed(uint32_t *__restrict__ s, uint64_t n,
                             uint64_t offset, uint32_t flip) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = (uint32_t)(((i * offset) % n) ^ flip);
}

// --nopermute_programs (common_language.h:478-486).
__global__ void k_pair_cyclic(uint32_t *__restrict__ s, uint64_t n,
                              uint32_t odd) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = odd ? (uint32_t)i : (uint32_t)(i == 0 ? n - 1 : i - 1);
}

// Fisher-Yates over the identity (common_language.h:418-424): for t = N-1
// down to 0, swap(s[t], s[j_t]) with j_t = SM(S(e*N + t)) % (t + 1).
//
// Parallel exact form.  Let in(t) be the value at position t just before
// step t.  Position t is last written before step t by a(t), the smallest
// t' > t with j_{t'} = t, which moved in(a(t)) there; so in(t) = in(a(t)),
// or t itself when no such step exists.  Step t then fixes
//   s[t] = in(t)                         if j_t == t,
//        = in(c(t))                      if c(t) exists,
//        = j_t                           otherwise,
// where c(t) is the smallest t' > t with j_{t'} = j_t (the last step before
// t that wrote position j_t).  Steps are bucketed by j_t; each bucket sorted
// ascending gives c() as the in-bucket successor and a(t) as the first
// bucket element above t.  The a()-chains strictly increase, so chasing them
// terminates.

// Every kernel below draws gridDim.y epochs at once: blockIdx.y = y selects
// epoch `epoch + y`, whose arrays start y * n entries in (cnt: y * (n + 1),
// its last entry the bucket allocator's counter).
__global__ void k_fy_draw(uint32_t *__restrict__ j, uint32_t *__restrict__ cnt,
                          uint64_t n, uint64_t sm_seed, uint64_t epoch) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t y = blockIdx.y;
  j += y * n;
  cnt += y * (n + 1);
  const uint64_t r = splitmix64(epoch_seed_pre(sm_seed, (epoch + y) * n + t));
  const uint32_t jt = (uint32_t)(r % (t + 1));
  j[t] = jt;
  atomicAdd(&cnt[jt], 1u);
}

__global__ void k_fy_alloc(uint32_t *__restrict__ cnt,
                           uint32_t *__restrict__ base,
                           uint32_t *__restrict__ cursor, uint64_t n) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t y = blockIdx.y;
  const bool in = p < n;
  const uint32_t c = in ? cnt[y * (n + 1) + p] : 0u;
  // all lanes take part
  const uint32_t b = wave_alloc_u32(cnt + y * (n + 1) + n, c);
  if (!in) return;
  base[y * n + p] = b;
  cursor[y * n + p] = 0;
}

__global__ void k_fy_place(const uint32_t *__restrict__ j,
                           const uint32_t *__restrict__ base,
                           uint32_t *__restrict__ cursor,
                           uint32_t *__restrict__ items, uint64_t n) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t o = (uint64_t)blockIdx.y * n;
  const uint32_t jt = j[o + t];
  items[o + base[o + jt] + atomicAdd(&cursor[o + jt], 1u)] = (uint32_t)t;
}

__global__ void k_fy_sort(const uint32_t *__restrict__ cnt,
                          const uint32_t *__restrict__ base,
                          uint32_t *__restrict__ items,
                          uint32_t *__restrict__ first,
                          uint32_t *__restrict__ succ, uint64_t n) {
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint64_t o = (uint64_t)blockIdx.y * n;
  const uint32_t c = cnt[blockIdx.y * (n + 1) + p];
  if (c == 0) {
    first[o + p] = kNone;
    return;
  }
  uint32_t *b = items + o + base[o + p];
  for (uint32_t x = 1; x < c; x++) {  // insertion sort; buckets are tiny
    const uint32_t v = b[x];
    uint32_t y = x;
    for (; y > 0 && b[y - 1] > v; y--) b[y] = b[y - 1];
    b[y] = v;
  }
  first[o + p] = b[0];
  for (uint32_t x = 0; x < c; x++) succ[o + b[x]] = x + 1 < c ? b[x + 1] : kNone;
}

__device__ __forceinline__ uint32_t fy_in(uint32_t t,
                                          const uint32_t *__restrict__ j,
                                          const uint32_t *__restrict__ first,
                                          const uint32_t *__restrict__ succ) {
  for (;;) {
    const uint32_t a = j[t] == t ? succ[t] : first[t];
    if (a == kNone) return t;
    t = a;
  }
}

__global__ void k_fy_final(const uint32_t *__restrict__ j,
                           const uint32_t *__restrict__ first,
                           const uint32_t *__restrict__ succ,
                           uint32_t *__restrict__ s, uint64_t n) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t o = (uint64_t)blockIdx.y * n;
  j += o;
  first += o;
  succ += o;
  const uint32_t jt = j[t];
  uint32_t v;
  if (jt == (uint32_t)t) {
    v = fy_in((uint32_t)t, j, first, succ);
  } else {
    const uint32_t c = succ[t];
    v = c == kNone ? jt : fy_in(c, j, first, succ);
  }
  s[o + t] = v;
}

// 2-D greedy neighbour pairing (common_language.h:435-463).  Visiting
// programs in the Fisher-Yates order `order`, program i proposes edge
// (i, nbr_i) with nbr_i = nbr[row_off[i] + S(S(e) ^ S(i)) % deg(i)]; the
// sequential loop accepts an edge when neither end is used yet.  Exact
// parallel form by rounds: an undecided edge whose position in the order is
// the minimum over undecided edges at both its ends is accepted (every
// earlier edge at its ends is decided and none of them was accepted); an
// undecided edge with a used end is rejected (the edge that used it came
// earlier).  Each round decides at least the globally first undecided edge.

__global__ void k_g2_propose(const uint32_t *__restrict__ order,
                             const uint64_t *__restrict__ row_off,
                             const uint32_t *__restrict__ nbr,
                             uint32_t *__restrict__ eu, uint32_t *__restrict__ ew,
                             uint8_t *__restrict__ state, uint64_t n,
                             uint64_t sm_seed, uint64_t s_epoch) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = order[k];
  const uint64_t deg = row_off[i + 1] - row_off[i];
  if (deg == 0) {
    state[k] = 2;  // no edge: skipped (:445-447)
    return;
  }
  const uint64_t idx =
      epoch_seed_pre(sm_seed, s_epoch ^ epoch_seed_pre(sm_seed, i)) % deg;
  eu[k] = i;
  ew[k] = nbr[row_off[i] + idx];
  state[k] = 0;  // undecided
}

__global__ void k_g2_minpri(const uint32_t *__restrict__ eu,
                            const uint32_t *__restrict__ ew,
                            const uint8_t *__restrict__ state,
                            uint32_t *__restrict__ minpri, uint64_t n) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || state[k] != 0) return;
  atomicMin(&minpri[eu[k]], (uint32_t)k);
  atomicMin(&minpri[ew[k]], (uint32_t)k);
}

__global__ void k_g2_accept(const uint32_t *__restrict__ eu,
                            const uint32_t *__restrict__ ew,
                            uint8_t *__restrict__ state,
                            const uint32_t *__restrict__ minpri,
                            uint8_t *__restrict__ used, uint64_t n) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || state[k] != 0) return;
  const uint32_t u = eu[k], w = ew[k];
  if (minpri[u] == (uint32_t)k && minpri[w] == (uint32_t)k) {
    state[k] = 1;
    used[u] = 1;
    used[w] = 1;
  }
}

__global__ void k_g2_reject(const uint32_t *__restrict__ eu,
                            const uint32_t *__restrict__ ew,
                            uint8_t *__restrict__ state,
                            const uint8_t *__restrict__ used,
                            uint32_t *__restrict__ minpri,
                            unsigned int *__restrict__ undecided, uint64_t n) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  minpri[k] = kNone;  // reset for the next round (one entry per program)
  if (state[k] != 0) return;
  if (used[eu[k]] || used[ew[k]]) state[k] = 2;
  else atomicAdd(undecided, 1u);
}

// Single-block exclusive scan of flags (1 byte each) into u32 ranks.
// Writes the total to *total.  Used for the 2-D pairing's compactions.
__global__ __launch_bounds__(1024) void k_scan_flags(
    const uint8_t *__restrict__ flags, uint8_t want, uint32_t *__restrict__ rank,
    uint64_t n, uint64_t *__restrict__ total) {
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x;
  const uint64_t per = (n + 1023) / 1024;
  const uint64_t lo = tid * per, hi = min(n, lo + per);
  uint32_t c = 0;
  for (uint64_t i = lo; i < hi; i++) c += flags[i] == want;
  part[tid] = c;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    uint32_t v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = part[tid] - c;
  for (uint64_t i = lo; i < hi; i++) {
    rank[i] = run;
    run += flags[i] == want;
  }
  if (tid == 1023) *total = part[1023];
}

__global__ void k_g2_emit_pairs(const uint32_t *__restrict__ eu,
                                const uint32_t *__restrict__ ew,
                                const uint8_t *__restrict__ state,
                                const uint32_t *__restrict__ rank,
                                uint32_t *__restrict__ s, uint64_t n) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || state[k] != 1) return;
  const uint32_t m = rank[k];
  s[2 * (uint64_t)m] = eu[k];
  s[2 * (uint64_t)m + 1] = ew[k];
}

// Leftovers in ascending program order after the accepted pairs (:458-463).
__global__ void k_g2_emit_rest(const uint8_t *__restrict__ used,
                               const uint32_t *__restrict__ rank,
                               const uint64_t *__restrict__ num_pairs,
                               uint32_t *__restrict__ s, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || used[i]) return;
  const uint64_t pos = 2 * *num_pairs + rank[i];
  if (pos < n) s[pos] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// Byte histogram of the soup (the h0 and frequent-byte metrics,
// common_language.h:519-540): per-block LDS counters, then 256 global adds.
__global__ __launch_bounds__(256) void k_byte_hist(const uint4 *__restrict__ soup, uint64_t units,
                                                   unsigned long long *__restrict__ counts) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < units;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 v = soup[u];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int b = 0; b < 16; b++) atomicAdd(&h[(w[b >> 2] >> (8 * (b & 3))) & 255u], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&counts[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Host-side launch helpers (declared in kernels.h).

namespace {
inline unsigned blocks_for(uint64_t n, unsigned b) {
  return (unsigned)((n + b - 1) / b);
}

// Instantiates KERNEL<L> for every language and launches the one selected.
#define CUBFF_LAUNCH_LANG(lang, KERNEL, grid, block, st, ...)             \
  switch (lang) {                                                         \
    case kBff: hipLaunchKernelGGL(KERNEL<kBff>, grid, block, 0, st, __VA_ARGS__); break; \
    case kBffNoheads: hipLaunchKernelGGL(KERNEL<kBffNoheads>, grid, block, 0, st, __VA_ARGS__); break; \
    case kBff8: hipLaunchKernelGGL(KERNEL<kBff8>, grid, block, 0, st, __VA_ARGS__); break; \
    case kBff8Noheads: hipLaunchKernelGGL(KERNEL<kBff8Noheads>, grid, block, 0, st, __VA_ARGS__); break; \
    case kBffNoheads4bit: hipLaunchKernelGGL(KERNEL<kBffNoheads4bit>, grid, block, 0, st, __VA_ARGS__); break; \
    case kBffPerm: hipLaunchKernelGGL(KERNEL<kBffPerm>, grid, block, 0, st, __VA_ARGS__); break; \
    case kBffSelfmove: hipLaunchKernelGGL(KERNEL<kBffSelfmove>, grid, block, 0, st, __VA_ARGS__); break; \
    case kForth: hipLaunchKernelGGL(KERNEL<kForth>, grid, block, 0, st, __VA_ARGS__); break; \
    case kForthcopy: hipLaunchKernelGGL(KERNEL<kForthcopy>, grid, block, 0, st, __VA_ARGS__); break; \
    case kForthtrivial: hipLaunchKernelGGL(KERNEL<kForthtrivial>, grid, block, 0, st, __VA_ARGS__); break; \
    case kForthtrivialReset: hipLaunchKernelGGL(KERNEL<kForthtrivialReset>, grid, block, 0, st, __VA_ARGS__); break; \
    case kSubleq: hipLaunchKernelGGL(KERNEL<kSubleq>, grid, block, 0, st, __VA_ARGS__); break; \
    case kRsubleq4: hipLaunchKernelGGL(KERNEL<kRsubleq4>, grid, block, 0, st, __VA_ARGS__); break; \
    default: return hipErrorInvalidValue;                                 \
  }                                                                       \
  return hipGetLastError();
}  // namespace

hipError_t launch_byte_hist(hipStream_t st, const uint8_t *soup, uint64_t bytes,
                            unsigned long long *counts) {
  const uint64_t units = bytes / 16;  // the soup is 64 N bytes
  if (units == 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<uint64_t>(blocks_for(units, 256), 2048);
  hipLaunchKernelGGL(k_byte_hist, dim3(grid), dim3(256), 0, st,
                     reinterpret_cast<const uint4 *>(soup), units, counts);
  return hipGetLastError();
}

hipError_t launch_init(hipStream_t st, uint8_t *soup, uint64_t n,
                       uint64_t seed_value, uint64_t count, int zero_init) {
  const uint64_t units = count * (kTape / 16);
  if (units == 0) return hipSuccess;
  hipLaunchKernelGGL(k_init, dim3(blocks_for(units, 256)), dim3(256), 0, st,
                     soup, n, seed_value, count, zero_init);
  return hipGetLastError();
}

hipError_t launch_epoch(hipStream_t st, int lang, uint8_t *soup,
                        const uint32_t *shuf, uint64_t n,
                        const uint64_t *num_indices, uint64_t seed_value,
                        uint32_t prob, uint32_t *pair_ops, uint32_t *pair_steps,
                        unsigned long long *counters) {
  const uint64_t slots = n / 2;
  if (slots == 0) return hipSuccess;
  CUBFF_LAUNCH_LANG(lang, k_epoch, dim3(blocks_for(slots, kEpochLanes)),
                    dim3(kEpochLanes), st, soup, shuf, n, num_indices,
                    seed_value, prob, pair_ops, pair_steps, counters)
}

hipError_t launch_selfrep(hipStream_t st, int lang, const uint8_t *soup,
                          uint64_t n, uint64_t count, uint64_t seed_value,
                          uint64_t *result, uint8_t *trace) {
  if (count == 0) return hipSuccess;
  CUBFF_LAUNCH_LANG(lang, k_selfrep, dim3(blocks_for(count, kSelfrepProgs)),
                    dim3(kSelfrepLanes), st, soup, n, count, seed_value, result, trace)
}

hipError_t launch_eval_tapes(hipStream_t st, int lang, uint8_t *tapes,
                             uint64_t count, uint32_t steps, uint32_t *ops,
                             uint32_t *steps_out) {
  if (count == 0) return hipSuccess;
  CUBFF_LAUNCH_LANG(lang, k_eval_tapes, dim3(blocks_for(count, kEpochLanes)),
                    dim3(kEpochLanes), st, tapes, count, steps, ops, steps_out)
}

hipError_t launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v,
                           uint64_t count) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(count, 256)), dim3(256), 0,
                     st, p, v, count);
  return hipGetLastError();
}

hipError_t launch_set_u64(hipStream_t st, uint64_t *p, uint64_t v) {
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, p, v);
  return hipGetLastError();
}

hipError_t launch_pair_fixed(hipStream_t st, uint32_t *s, uint64_t n,
                             uint64_t epoch) {
  // (1 << (epoch % max_pow2 + 1)) - 1 in int arithmetic; max_pow2 < 32.
  const uint64_t max_pow2 = 31 - __builtin_clz((unsigned)n);
  const int sh = (int)(epoch % max_pow2 + 1);
  const uint64_t offset = sh >= 31 ? 0x7fffffffULL : ((1ULL << sh) - 1);
  hipLaunchKernelGGL(k_pair_fixed, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                     s, n, offset, (uint32_t)(epoch & 1));
  return hipGetLastError();
}

hipError_t launch_pair_cyclic(hipStream_t st, uint32_t *s, uint64_t n,
                              uint64_t epoch) {
  hipLaunchKernelGGL(k_pair_cyclic, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                     s, n, (uint32_t)(epoch % 2 == 1));
  return hipGetLastError();
}

hipError_t launch_pair_shuffle_many(hipStream_t st, const FyScratch &w, uint32_t *s,
                                    uint64_t n, uint64_t params_seed, uint64_t epoch,
                                    uint32_t epochs) {
  const uint64_t sm_seed = splitmix64(params_seed);
  const dim3 g(blocks_for(n, 256), epochs);
  hipError_t e;
  if ((e = launch_fill_u32(st, w.cnt, 0, (n + 1) * epochs)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_fy_draw, g, dim3(256), 0, st, w.j, w.cnt, n, sm_seed, epoch);
  hipLaunchKernelGGL(k_fy_alloc, g, dim3(256), 0, st, w.cnt, w.base, w.cursor, n);
  hipLaunchKernelGGL(k_fy_place, g, dim3(256), 0, st, w.j, w.base, w.cursor, w.items, n);
  hipLaunchKernelGGL(k_fy_sort, g, dim3(256), 0, st, w.cnt, w.base, w.items, w.first, w.succ, n);
  hipLaunchKernelGGL(k_fy_final, g, dim3(256), 0, st, w.j, w.first, w.succ, s, n);
  return hipGetLastError();
}

hipError_t launch_pair_shuffle(hipStream_t st, const FyScratch &w, uint32_t *s,
                               uint64_t n, uint64_t params_seed,
                               uint64_t epoch) {
  return launch_pair_shuffle_many(st, w, s, n, params_seed, epoch, 1);
}

hipError_t launch_pair_2d(hipStream_t st, const FyScratch &w,
                          const G2Scratch &g2, const uint64_t *row_off,
                          const uint32_t *nbr, uint32_t *s, uint64_t n,
                          uint64_t params_seed, uint64_t epoch,
                          uint64_t *num_indices) {
  hipError_t e;
  if ((e = launch_pair_shuffle(st, w, g2.order, n, params_seed, epoch)) !=
      hipSuccess)
    return e;
  const uint64_t sm_seed = splitmix64(params_seed);
  const uint64_t s_epoch = epoch_seed_pre(sm_seed, epoch);
  const unsigned g = blocks_for(n, 256);
  hipLaunchKernelGGL(k_g2_propose, dim3(g), dim3(256), 0, st, g2.order,
                     row_off, nbr, g2.eu, g2.ew, g2.state, n, sm_seed, s_epoch);
  if ((e = launch_fill_u32(st, g2.minpri, kNone, n)) != hipSuccess) return e;
  if ((e = launch_fill_u32(st, reinterpret_cast<uint32_t *>(g2.used), 0,
                           (n + 3) / 4)) != hipSuccess)
    return e;
  // Rounds until no edge is undecided.  The count is read back every
  // kCheck rounds; extra rounds past completion are no-ops.
  constexpr int kCheck = 8;
  for (uint64_t round = 0;; round += kCheck) {
    for (int r = 0; r < kCheck; r++) {
      if ((e = launch_fill_u32(st, g2.undecided, 0, 1)) != hipSuccess) return e;
      hipLaunchKernelGGL(k_g2_minpri, dim3(g), dim3(256), 0, st, g2.eu, g2.ew,
                         g2.state, g2.minpri, n);
      hipLaunchKernelGGL(k_g2_accept, dim3(g), dim3(256), 0, st, g2.eu, g2.ew,
                         g2.state, g2.minpri, g2.used, n);
      hipLaunchKernelGGL(k_g2_reject, dim3(g), dim3(256), 0, st, g2.eu, g2.ew,
                         g2.state, g2.used, g2.minpri, g2.undecided, n);
    }
    uint32_t left = 0;
    if ((e = hipMemcpyAsync(&left, g2.undecided, 4, hipMemcpyDeviceToHost,
                            st)) != hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    if (left == 0) break;
    if (round > 4 * n + 64) return hipErrorUnknown;  // cannot happen
  }
  hipLaunchKernelGGL(k_scan_flags, dim3(1), dim3(1024), 0, st, g2.state,
                     (uint8_t)1, g2.rank, n, num_indices);
  hipLaunchKernelGGL(k_g2_emit_pairs, dim3(g), dim3(256), 0, st, g2.eu, g2.ew,
                     g2.state, g2.rank, s, n);
  hipLaunchKernelGGL(k_scan_flags, dim3(1), dim3(1024), 0, st, g2.used,
                     (uint8_t)0, g2.rank, n, g2.rest_total);
  hipLaunchKernelGGL(k_g2_emit_rest, dim3(g), dim3(256), 0, st, g2.used,
                     g2.rank, num_indices, s, n);
  return hipGetLastError();
}

}  // namespace cubff
