// Many epochs of one soup in a single launch, run as a dataflow graph.
//
// An epoch's pairs only wait for the pairs of the previous epoch that last
// wrote their two programs, never for the whole previous epoch.  With the
// pairings of all epochs of a segment known in advance (the Fisher-Yates,
// fixed and cyclic pairings do not depend on the soup), pair (j+1, k) can
// start as soon as the two pairs of epoch j holding its programs are done.
// A single epoch lasts as long as its longest interaction (a few in 10^5
// run thousands of commands); over many epochs the dataflow order only
// waits along the longest dependency chain, which is several times shorter
// than the sum of the per-epoch maxima.  Results are bit-identical to one
// MutateAndRunPrograms (common_language.h:157-193) per epoch in order: every
// program is read by exactly the interaction that follows its last writer.
//
// Work items are "pair k of segment epoch j" (id j * slots + k).  A wave
// runs up to 64 items at a time, one per lane, then takes a release fence
// and decrements the pending count of the epoch-(j+1) pairs holding each
// item's two programs.  An item whose count this wave brought to zero is
// ready: the wave keeps the first 64 such items for its next round (no
// queue traffic on the common path) and appends the rest to a global ready
// queue, from which waves with free lanes claim runs of items.  Epoch 0's
// items are handed to the waves directly, 64 each, and the rest queued.  A
// queued item is run after its pending count reads zero and an acquire
// fence: the counts are read-modify-written after each writer's release
// fence, so reading zero synchronizes with both writers.  Every wait is
// bounded by a wall-clock limit: past it the kernel raises the abort flag,
// every wave leaves at its next claim and the host reports the error.
//
// The gain is largest where single interactions run far longer than the
// rest (the Forth family at 8192 steps): the rest of the soup moves on
// while they run.  Where the longest interaction is short compared with the
// fences, queue operations and global-memory latency of a round (about
// 100-150 us per wave), one launch per epoch stays faster.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "kernels.h"
#include "languages.h"
#include "tape_io.h"

namespace cubff {

namespace {

__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lane 0 of a wave: claims up to `want` queued items [*h, *h + *k).  With
// `wait` false it makes one attempt and may return *k = 0.  With `wait`
// true it waits for work: a full run of `want` is claimed at once, fewer
// only after `patience` ticks without one; *stop is set once every item of
// the segment is done or the run was aborted.  Idle polls back off
// exponentially and read head and tail with one 64-bit load, which keeps
// the traffic on the control block low.
__device__ void flow_claim(FlowCtl *c, uint32_t total, uint64_t ticks, uint64_t patience,
                           uint32_t want, bool wait, uint32_t *h, uint32_t *k,
                           uint32_t *stop) {
  const uint64_t t0 = wall_clock64();
  uint32_t nap = 1, polls = 0;
  for (;;) {
    const uint64_t ht = __hip_atomic_load(reinterpret_cast<const uint64_t *>(&c->head),
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t hh = (uint32_t)ht, tt = (uint32_t)(ht >> 32);
    const uint32_t avail = tt - hh;  // tail never trails head
    if (avail >= want || (avail > 0 && (!wait || wall_clock64() - t0 >= patience))) {
      const uint32_t kk = min(avail, want);
      uint32_t expect = hh;
      if (__hip_atomic_compare_exchange_strong(&c->head, &expect, hh + kk, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_fetch_add(&c->claims, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *h = hh;
        *k = kk;
        return;
      }
      if (wait) continue;  // another wave claimed first; retry at once
    }
    if (!wait) return;
    if ((++polls & 3) == 0) {
      if (ld_agent(&c->abort) || ld_agent(&c->done) >= total) break;
      if (wall_clock64() - t0 > ticks) {
        st_agent(&c->abort, 1);
        break;
      }
    }
    for (uint32_t q = 0; q < nap; q++) __builtin_amdgcn_s_sleep(1);
    nap = min(nap * 2, 128u);
  }
  *stop = 1;
}

// Exclusive prefix sum over the 64 lanes of a wavefront; *sum gets the total.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t *sum) {
  const int lane = threadIdx.x & 63;
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  *sum = __shfl(incl, 63, 64);
  return incl - v;
}

// Waits until *p != kNone (want_zero false) or *p == 0 (want_zero true);
// returns the value, or kNone after raising the abort flag.
__device__ uint32_t flow_wait(FlowCtl *c, const uint32_t *p, bool want_zero,
                              uint64_t ticks) {
  const uint64_t t0 = wall_clock64();
  for (;;) {
    const uint32_t v = ld_agent(p);
    if (want_zero ? v == 0 : v != kNone) return v;
    if (ld_agent(&c->abort)) return kNone;
    if (wall_clock64() - t0 > ticks) {
      st_agent(&c->abort, 1);
      return kNone;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

}  // namespace

template <int LANG>
__global__ __launch_bounds__(kEpochLanes) void k_flow(FlowArgs a) {
  using L = typename LangOf<LANG>::T;
  using Tape = typename L::template Tape<kEpochLanes>;
  using Aux = typename L::template Aux<kEpochLanes>;
  constexpr int kAuxBytes = lang_needs_aux(LANG) ? Aux::kRegionBytes : 0;
  __shared__ __align__(16) uint8_t lds[Tape::kRegionBytes + kAuxBytes];
  __shared__ uint32_t keep[kEpochLanes];  // items a wave made ready and runs next
  const int lane = threadIdx.x, wl = lane & 63, wbase = lane - wl;
  const auto t = Tape::at(lds, lane);
  const auto aux = Aux::at(lds + Tape::kRegionBytes, lane);
  const uint32_t total = a.epochs * a.slots;
  uint64_t my_ops = 0, my_steps = 0;
  // First items: this wave's share of epoch 0's first `fixed` items.
  const uint32_t w0 = (blockIdx.x * kEpochLanes + wbase);  // = 64 * wave index
  uint32_t kept = w0 < a.fixed ? min(a.fixed - w0, 64u) : 0;  // wave-uniform
  if ((uint32_t)wl < kept) keep[lane] = w0 + wl;
  if (blockIdx.x == 0 && lane == 0) atomicAdd(&a.counters[1], (unsigned long long)total);
  for (;;) {
    // This wave's next items: those it made ready itself (their pending
    // counts reached zero in its own decrements), topped up from the queue.
    uint32_t item = (uint32_t)wl < kept ? keep[lane] : kNone;
    bool queued = false;
    if (kept < 64) {
      uint32_t h = 0, k = 0, stop = 0;
      if (wl == 0)
        flow_claim(a.ctl, total, a.spin_ticks, a.patience_ticks, 64 - kept, kept == 0, &h, &k,
                   &stop);
      h = __shfl(h, 0, 64);
      k = __shfl(k, 0, 64);
      if (__shfl(stop, 0, 64)) break;
      if ((uint32_t)wl >= kept && (uint32_t)wl < kept + k) {
        item = flow_wait(a.ctl, a.queue + h + (wl - kept), false, a.spin_ticks);
        queued = true;
      }
    }
    if (item != kNone && item >= total) {  // never queued by this kernel
      st_agent(&a.ctl->abort, 1);
      item = kNone;
    }
    uint32_t j = 0, kk = 0;
    if (item != kNone) {
      j = item / a.slots;
      kk = item - j * a.slots;
      if (queued && j > 0 && flow_wait(a.ctl, a.cnt + item, true, a.spin_ticks) != 0)
        item = kNone;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    uint32_t p1 = 0, p2 = 0;
    if (item != kNone) {
      const uint32_t *s = a.shuf + (uint64_t)j * a.n;
      p1 = s[2 * kk];
      p2 = s[2 * kk + 1];
      load_row(t, 0, a.soup + (uint64_t)p1 * kTape);
      load_row(t, 1, a.soup + (uint64_t)p2 * kTape);
      mutate(t, a.n, epoch_seed_pre(a.sm_seed, a.first_epoch + j), kk, a.prob);
      uint32_t steps = 0;
      const uint32_t ops = L::evaluate(t, aux, kEvalSteps, &steps);
      store_row(t, 0, a.soup + (uint64_t)p1 * kTape);
      store_row(t, 1, a.soup + (uint64_t)p2 * kTape);
      my_ops += ops;
      my_steps += steps;
      if (j + 1 == a.epochs) {
        a.pair_ops[kk] = ops;
        a.pair_steps[kk] = steps;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const uint32_t ran = __builtin_popcountll(__ballot(item != kNone));
    if (wl == 0 && ran) __hip_atomic_fetch_add(&a.ctl->done, ran, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
    uint32_t r0 = kNone, r1 = kNone;
    if (item != kNone && j + 1 < a.epochs) {
      const uint32_t *po = a.pairof + (uint64_t)(j + 1) * a.n;
      const uint32_t base = (j + 1) * a.slots;
      const uint32_t q1 = base + po[p1], q2 = base + po[p2];
      if (__hip_atomic_fetch_add(a.cnt + q1, 0xffffffffu, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) == 1)
        r0 = q1;
      if (__hip_atomic_fetch_add(a.cnt + q2, 0xffffffffu, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) == 1)
        r1 = q2;
    }
    // Ready items in wave order: the first 64 stay with this wave, the rest
    // go to the queue.
    if (r0 == kNone) {
      r0 = r1;
      r1 = kNone;
    }
    const uint32_t nr = (r0 != kNone) + (r1 != kNone);
    uint32_t made = 0;
    const uint32_t pos = wave_excl_scan(nr, &made);
    uint32_t sp0 = kNone, sp1 = kNone;  // items that go to the queue
    if (nr >= 1) {
      if (pos < 64) keep[wbase + pos] = r0;
      else sp0 = r0;
    }
    if (nr == 2) {
      if (pos + 1 < 64) keep[wbase + pos + 1] = r1;
      else if (sp0 == kNone) sp0 = r1;
      else sp1 = r1;
    }
    const uint32_t spill = (sp0 != kNone) + (sp1 != kNone);
    const uint32_t at = wave_alloc_u32(&a.ctl->tail, spill);
    if (spill && at + spill > total) {
      st_agent(&a.ctl->abort, 1);
    } else {
      if (sp0 != kNone) st_agent(a.queue + at, sp0);
      if (sp1 != kNone) st_agent(a.queue + at + 1, sp1);
    }
    kept = min(made, 64u);
    // keep[] is written and read by lanes of this wavefront only.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  const uint64_t wops = wave_sum_u64(my_ops);
  const uint64_t wsteps = wave_sum_u64(my_steps);
  if (wl == 0 && wops) atomicAdd(&a.counters[0], (unsigned long long)wops);
  if (wl == 0 && wsteps) atomicAdd(&a.counters[2], (unsigned long long)wsteps);
}

// pairof[s[i]] = i / 2: the pair of each program in one epoch.
__global__ void k_flow_pairof(const uint32_t *__restrict__ s,
                              uint32_t *__restrict__ pairof, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pairof[s[i]] = (uint32_t)(i / 2);
}

// Pending counts 2; epoch 0's items past the first `fixed` (which the
// waves take directly, 64 each) queued; the rest of the queue empty.
__global__ void k_flow_init(uint32_t *__restrict__ cnt, uint32_t *__restrict__ queue,
                            FlowCtl *ctl, uint32_t slots, uint32_t total, uint32_t fixed) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) {
    cnt[i] = 2;
    queue[i] = i < slots - fixed ? fixed + i : kNone;
  }
  if (i == 0) {
    ctl->head = 0;
    ctl->tail = slots - fixed;
    ctl->abort = 0;
    ctl->claims = 0;
    ctl->done = 0;
  }
}

namespace {
inline unsigned blocks_of(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <int LANG>
hipError_t launch_flow_lang(hipStream_t st, FlowArgs a) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_flow<LANG>, kEpochLanes, 0)) !=
      hipSuccess)
    return e;
  // At most one epoch's pairs are ever in flight (a program is in one running
  // interaction at a time), so twice that many lanes is plenty.
  const unsigned want = 2 * blocks_of(a.slots, kEpochLanes);
  const unsigned grid = std::max(1u, std::min((unsigned)(std::max(per_cu, 1) * cus), want));
  // Epoch 0's first 64 x (waves) items go to the waves directly.
  const uint64_t waves = (uint64_t)grid * (kEpochLanes / 64);
  a.fixed = (uint32_t)std::min<uint64_t>(a.slots, 64 * waves);
  const uint32_t total = a.epochs * a.slots;
  hipLaunchKernelGGL(k_flow_init, dim3(blocks_of(total, 256)), dim3(256), 0, st, a.cnt, a.queue,
                     a.ctl, a.slots, total, a.fixed);
  hipLaunchKernelGGL(k_flow<LANG>, dim3(grid), dim3(kEpochLanes), 0, st, a);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_flow_pairof(hipStream_t st, const uint32_t *shuf, uint32_t *pairof,
                              uint64_t n, uint32_t epochs) {
  for (uint32_t j = 1; j < epochs; j++)
    hipLaunchKernelGGL(k_flow_pairof, dim3(blocks_of(n, 256)), dim3(256), 0, st,
                       shuf + (uint64_t)j * n, pairof + (uint64_t)j * n, n);
  return hipGetLastError();
}

hipError_t launch_flow(hipStream_t st, int lang, const FlowArgs &a) {
  switch (lang) {
    case kBff: return launch_flow_lang<kBff>(st, a);
    case kBffNoheads: return launch_flow_lang<kBffNoheads>(st, a);
    case kBff8: return launch_flow_lang<kBff8>(st, a);
    case kBff8Noheads: return launch_flow_lang<kBff8Noheads>(st, a);
    case kBffNoheads4bit: return launch_flow_lang<kBffNoheads4bit>(st, a);
    case kBffPerm: return launch_flow_lang<kBffPerm>(st, a);
    case kBffSelfmove: return launch_flow_lang<kBffSelfmove>(st, a);
    case kForth: return launch_flow_lang<kForth>(st, a);
    case kForthcopy: return launch_flow_lang<kForthcopy>(st, a);
    case kForthtrivial: return launch_flow_lang<kForthtrivial>(st, a);
    case kForthtrivialReset: return launch_flow_lang<kForthtrivialReset>(st, a);
    case kSubleq: return launch_flow_lang<kSubleq>(st, a);
    case kRsubleq4: return launch_flow_lang<kRsubleq4>(st, a);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cubff
