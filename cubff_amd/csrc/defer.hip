// Epochs that do not wait for their longest interactions.
//
// One launch per epoch is as slow as the epoch's longest interaction: at
// 2^17 programs a typical interaction executes ~23 commands and the longest
// ~1,200-3,100.  Here the epoch kernel (k_epoch_defer) runs every
// interaction with a command budget; one that exceeds it is abandoned (its
// tape is discarded, nothing is written back) and appended to a list of
// deferred interactions, together with every interaction whose program is
// still owed by a deferred one.  A small persistent kernel on a second
// stream (k_resolve) runs the deferred interactions to completion, one per
// wavefront, in list order, each after the deferred interactions it
// depends on, while the epoch kernels go on.  Every program is therefore
// read by exactly the interaction that follows its last writer, as in the
// epoch-by-epoch loop, and the results are bit-identical to it.
//
// owner[p] names p's pending writer: kNone (the last writer was an epoch
// kernel, visible to later kernels of the same stream), a list index (a
// deferred interaction will write p), or kResolved (the resolver wrote p: an
// epoch kernel reading it takes an acquire fence).  A deferred interaction
// records the pending writers of its two programs as dependencies.  The
// resolver publishes a result with a release fence before marking it done
// and handing the programs back (owner: index -> kResolved).  List entries
// carry the run's generation number in their ready/done words, so the list
// is never cleared.  Every wait is bounded by a wall-clock limit; past it
// the abort flag stops the resolver and the host reports the error.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"
#include "bff_fast.h"
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

// Waits until *p == want; false (abort raised or seen) past the limit.
__device__ bool wait_eq(DeferCtl *c, const uint32_t *p, uint32_t want, uint64_t ticks) {
  const uint64_t t0 = wall_clock64();
  uint32_t nap = 1;
  for (;;) {
    if (ld_agent(p) == want) return true;
    if (ld_agent(&c->abort)) return false;
    if (wall_clock64() - t0 > ticks) {
      st_agent(&c->abort, 1);
      return false;
    }
    for (uint32_t q = 0; q < nap; q++) __builtin_amdgcn_s_sleep(1);
    nap = min(nap * 2, 32u);
  }
}

}  // namespace

template <int LANG>
__global__ __launch_bounds__(kEpochLanes) void k_epoch_defer(
    DeferArgs a, const uint32_t *__restrict__ shuf, const uint64_t *__restrict__ num_indices_p,
    uint32_t e_rel) {
  using L = typename LangOf<LANG>::T;
  using Tape = typename L::template Tape<kEpochLanes>;
  using Aux = typename L::template Aux<kEpochLanes>;
  // Budgeted runs of the standard BFF languages go through BffFast, whose
  // next-command table shares the second region with the aux tape of a
  // complete run (the two never overlap in time).
  constexpr bool kFast = FastOf<LANG>::kHas;
  constexpr int kAuxBytes = second_region_bytes<LANG, kEpochLanes>();
  __shared__ __align__(16) uint8_t lds[Tape::kRegionBytes + kAuxBytes];
  __shared__ uint8_t lut[kFast ? 256 : 1];
  __shared__ unsigned long long block_ops, block_steps;
  __shared__ uint32_t block_budget;
  const int lane = threadIdx.x;
  fill_fast_lut<LANG>(lut);
  if (lane == 0) {
    block_ops = 0;
    block_steps = 0;
    // While the resolver is behind, run every interaction to the end here:
    // only interactions whose programs are still owed are deferred, which
    // keeps the list (and the chains hanging off it) bounded.
    // Without a resolver (it left: see k_resolve), nothing is given up.
    const uint32_t backlog = ld_agent(&a.ctl->tail) - ld_agent(&a.ctl->resolved);
    block_budget = backlog > a.max_backlog || ld_agent(&a.ctl->abort) ? 0xffffffffu : a.budget;
  }
  __syncthreads();
  const uint32_t budget = block_budget;
  const uint64_t k = (uint64_t)blockIdx.x * kEpochLanes + lane;
  const auto t = Tape::at(lds, lane);
  const auto aux = Aux::at(lds + Tape::kRegionBytes, lane);
  uint32_t ops = 0, steps = 0, p1 = 0, p2 = 0, o1 = kNone, o2 = kNone;
  const bool eval = k < *num_indices_p;  // common_language.h:185
  if (k < a.slots) {
    p1 = shuf[2 * k];
    p2 = shuf[2 * k + 1];
    o1 = ld_agent(a.owner + p1);
    o2 = ld_agent(a.owner + p2);
    // An owing entry the resolver has finished (it hands the program back
    // right after): the program is resolved.  While the resolver is behind
    // (no budget), wait for owing entries instead of deferring more: the
    // resolver never waits on epoch kernels, so the wait ends, and the
    // list cannot grow without bound.
    const bool wait = budget == 0xffffffffu;
    if (o1 < kResolved &&
        (ld_agent(&a.items[o1].done) == a.gen ||
         (wait && wait_eq(a.ctl, &a.items[o1].done, a.gen, a.spin_ticks))))
      o1 = kResolved;
    if (o2 < kResolved &&
        (ld_agent(&a.items[o2].done) == a.gen ||
         (wait && wait_eq(a.ctl, &a.items[o2].done, a.gen, a.spin_ticks))))
      o2 = kResolved;
  }
  // A program last written by the resolver: see its bytes.
  if (__ballot(o1 == kResolved || o2 == kResolved))
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  bool defer = false;
  if (k < a.slots) {
    defer = o1 < kResolved || o2 < kResolved;  // a program is still owed
    if (!defer) {
      load_row(t, 0, a.soup + (uint64_t)p1 * kTape);
      load_row(t, 1, a.soup + (uint64_t)p2 * kTape);
      mutate(t, a.n, epoch_seed_pre(a.sm_seed, a.first_epoch + e_rel), k, a.prob);
      if (eval) {
        if constexpr (kFast)
          ops = FastOf<LANG>::T::evaluate(t, LinTape<kEpochLanes>::at(lds + Tape::kRegionBytes, lane),
                                          LdsByteLut{(const lds_u8 *)lut}, kEvalSteps, &steps,
                                          nullptr, budget);
        else
          ops = L::template evaluate<true>(t, aux, kEvalSteps, &steps, nullptr, budget);
      }
      if (steps == kOverBudget) {
        defer = true;
        ops = steps = 0;
      } else {
        store_row(t, 0, a.soup + (uint64_t)p1 * kTape);
        store_row(t, 1, a.soup + (uint64_t)p2 * kTape);
        if (o1 != kNone) st_agent(a.owner + p1, kNone);
        if (o2 != kNone) st_agent(a.owner + p2, kNone);
        if (e_rel + 1 == a.epochs) {
          a.pair_ops[k] = ops;
          a.pair_steps[k] = steps;
        }
      }
    }
  }
  const uint32_t at = wave_alloc_u32(&a.ctl->tail, defer ? 1u : 0u);
  // The resolver left before anything was deferred (tail held kDeferGone
  // when this allocation landed): no entry of this run exists, so no
  // program is owed and the interaction was deferred for its budget alone.
  // Run it here to the end, from the soup's unchanged bytes.
  const bool gone = defer && at >= kDeferGone;
  if (gone) {
    load_row(t, 0, a.soup + (uint64_t)p1 * kTape);
    load_row(t, 1, a.soup + (uint64_t)p2 * kTape);
    mutate(t, a.n, epoch_seed_pre(a.sm_seed, a.first_epoch + e_rel), k, a.prob);
    if (eval) ops = L::evaluate(t, aux, kEvalSteps, &steps);
    store_row(t, 0, a.soup + (uint64_t)p1 * kTape);
    store_row(t, 1, a.soup + (uint64_t)p2 * kTape);
    if (e_rel + 1 == a.epochs) {
      a.pair_ops[k] = ops;
      a.pair_steps[k] = steps;
    }
    defer = false;
  }
  const bool put = defer && at < a.capacity;
  if (defer && !put) st_agent(&a.ctl->abort, 1);
  if (put) {
    DeferItem *it = a.items + at;
    it->e = e_rel | (eval ? 0u : kNoEval);
    it->k = (uint32_t)k;
    it->p1 = p1;
    it->p2 = p2;
    it->dep1 = o1 < kResolved ? o1 : kNone;
    it->dep2 = o2 < kResolved ? o2 : kNone;
    st_agent(a.owner + p1, at);
    st_agent(a.owner + p2, at);
  }
  if (__ballot(put)) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (put) st_agent(&a.items[at].ready, a.gen);
  const uint64_t wsum = wave_sum_u64(ops);
  const uint64_t wsteps = wave_sum_u64(steps);
  if ((lane & 63) == 0 && wsum) atomicAdd(&block_ops, (unsigned long long)wsum);
  if ((lane & 63) == 0 && wsteps) atomicAdd(&block_steps, (unsigned long long)wsteps);
  __syncthreads();
  if (lane == 0 && block_ops) atomicAdd(&a.counters[0], block_ops);
  if (lane == 0 && block_steps) atomicAdd(&a.counters[2], block_steps);
  if (blockIdx.x == 0 && lane == 0) atomicAdd(&a.counters[1], (unsigned long long)*num_indices_p);
}

// The resolver: one deferred interaction per wavefront at a time, run by
// lane 0 (the list is short and its entries are the long interactions).
template <int LANG>
__global__ __launch_bounds__(64) void k_resolve(DeferArgs a) {
  using L = typename LangOf<LANG>::T;
  using Tape = typename L::template Tape<1>;
  using Aux = typename L::template Aux<1>;
  constexpr int kAuxBytes = lang_needs_aux(LANG) ? Aux::kRegionBytes : 0;
  __shared__ __align__(16) uint8_t lds[Tape::kRegionBytes + kAuxBytes];
  if (threadIdx.x != 0) return;
  const auto t = Tape::at(lds, 0);
  const auto aux = Aux::at(lds + Tape::kRegionBytes, 0);
  DeferCtl *c = a.ctl;
  for (;;) {
    // Claim the next list entry, or leave once the list is closed and done.
    uint32_t my = kNone;
    {
      // The epoch kernels bump counters[1] once each: while they progress,
      // waiting here for the next entry is not a stall.
      uint64_t t0 = wall_clock64();
      unsigned long long beat = __hip_atomic_load(&a.counters[1], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
      uint32_t nap = 1;
      for (;;) {
        if (ld_agent(&c->abort)) break;
        const uint32_t h = ld_agent(&c->head);
        const uint32_t traw = ld_agent(&c->tail);
        if (traw >= kDeferGone) break;  // another wavefront left with the list empty
        const uint32_t tl = min(traw, a.capacity);
        if (h < tl) {
          uint32_t expect = h;
          if (__hip_atomic_compare_exchange_strong(&c->head, &expect, h + 1, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            my = h;
            break;
          }
          continue;
        }
        if (ld_agent(&c->closed) && h >= ld_agent(&c->final_tail)) break;
        const unsigned long long b2 =
            __hip_atomic_load(&a.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (b2 != beat) {
          beat = b2;
          t0 = wall_clock64();
        }
        // Nothing deferred yet and no epoch kernel finished for a while: the
        // epoch kernels cannot start beside this kernel (dispatches are
        // serialized, as under some profilers).  Leave, marking the list
        // gone in the same atomic step that checks it is still empty; the
        // epoch kernels then run every interaction themselves (their
        // allocations return kDeferGone or more).  If an epoch kernel
        // deferred something meanwhile, stay and resolve it.
        if (tl == 0 && wall_clock64() - t0 > a.idle_ticks) {
          uint32_t expect = 0;
          if (__hip_atomic_compare_exchange_strong(&c->tail, &expect, kDeferGone,
                                                   __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT))
            break;
          continue;
        }
        if (wall_clock64() - t0 > a.spin_ticks) {
          st_agent(&c->abort, 1);
          break;
        }
        for (uint32_t q = 0; q < nap; q++) __builtin_amdgcn_s_sleep(1);
        nap = min(nap * 2, 64u);
      }
    }
    if (my == kNone) break;
    DeferItem *it = a.items + my;
    if (!wait_eq(c, &it->ready, a.gen, a.spin_ticks)) break;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t e = it->e & ~kNoEval, k = it->k, p1 = it->p1, p2 = it->p2;
    const bool eval = !(it->e & kNoEval);
    const uint32_t d1 = it->dep1, d2 = it->dep2;
    if (d1 != kNone && !wait_eq(c, &a.items[d1].done, a.gen, a.spin_ticks)) break;
    if (d2 != kNone && !wait_eq(c, &a.items[d2].done, a.gen, a.spin_ticks)) break;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    load_row(t, 0, a.soup + (uint64_t)p1 * kTape);
    load_row(t, 1, a.soup + (uint64_t)p2 * kTape);
    mutate(t, a.n, epoch_seed_pre(a.sm_seed, a.first_epoch + e), k, a.prob);
    uint32_t steps = 0, ops = 0;
    if (eval) ops = L::evaluate(t, aux, kEvalSteps, &steps);
    store_row(t, 0, a.soup + (uint64_t)p1 * kTape);
    store_row(t, 1, a.soup + (uint64_t)p2 * kTape);
    if (ops) atomicAdd(&a.counters[0], (unsigned long long)ops);
    if (steps) atomicAdd(&a.counters[2], (unsigned long long)steps);
    if (e + 1 == a.epochs) {
      a.pair_ops[k] = ops;
      a.pair_steps[k] = steps;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    st_agent(&it->done, a.gen);
    __hip_atomic_fetch_add(&c->resolved, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t expect = my;
    __hip_atomic_compare_exchange_strong(a.owner + p1, &expect, kResolved, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    expect = my;
    __hip_atomic_compare_exchange_strong(a.owner + p2, &expect, kResolved, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void k_defer_start(DeferCtl *c) {
  c->head = 0;
  c->tail = 0;
  c->closed = 0;
  c->abort = 0;
  c->final_tail = 0;
  c->resolved = 0;
}

// After the segment's last epoch kernel: no entry is appended any more.
__global__ void k_defer_close(DeferCtl *c) {
  st_agent(&c->final_tail, ld_agent(&c->tail));
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  st_agent(&c->closed, 1);
}

namespace {
inline unsigned blocks_of(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
}  // namespace

template <int L>
hipError_t epoch_defer_lang(hipStream_t st, const DeferArgs &a, const uint32_t *shuf,
                            const uint64_t *num_indices, uint32_t e_rel) {
  hipLaunchKernelGGL(k_epoch_defer<L>, dim3(blocks_of(a.slots, kEpochLanes)),
                     dim3(kEpochLanes), 0, st, a, shuf, num_indices, e_rel);
  return hipGetLastError();
}

template <int L>
hipError_t resolve_lang(hipStream_t st, const DeferArgs &a, unsigned waves) {
  hipLaunchKernelGGL(k_resolve<L>, dim3(waves), dim3(64), 0, st, a);
  return hipGetLastError();
}

#define CUBFF_DEFER_LANG(lang, F, ARGS)                                      \
  switch (lang) {                                                           \
    case kBff: return F<kBff> ARGS;                                         \
    case kBffNoheads: return F<kBffNoheads> ARGS;                           \
    case kBff8: return F<kBff8> ARGS;                                       \
    case kBff8Noheads: return F<kBff8Noheads> ARGS;                         \
    case kBffNoheads4bit: return F<kBffNoheads4bit> ARGS;                   \
    case kBffPerm: return F<kBffPerm> ARGS;                                 \
    case kBffSelfmove: return F<kBffSelfmove> ARGS;                         \
    case kForth: return F<kForth> ARGS;                                     \
    case kForthcopy: return F<kForthcopy> ARGS;                             \
    case kForthtrivial: return F<kForthtrivial> ARGS;                       \
    case kForthtrivialReset: return F<kForthtrivialReset> ARGS;             \
    case kSubleq: return F<kSubleq> ARGS;                                   \
    case kRsubleq4: return F<kRsubleq4> ARGS;                               \
    default: return hipErrorInvalidValue;                                   \
  }

hipError_t launch_defer_start(hipStream_t st, DeferCtl *c) {
  hipLaunchKernelGGL(k_defer_start, dim3(1), dim3(1), 0, st, c);
  return hipGetLastError();
}

hipError_t launch_defer_close(hipStream_t st, DeferCtl *c) {
  hipLaunchKernelGGL(k_defer_close, dim3(1), dim3(1), 0, st, c);
  return hipGetLastError();
}

hipError_t launch_epoch_defer(hipStream_t st, int lang, const DeferArgs &a,
                              const uint32_t *shuf, const uint64_t *num_indices, uint32_t e_rel) {
  if (a.slots == 0) return hipSuccess;
  CUBFF_DEFER_LANG(lang, epoch_defer_lang, (st, a, shuf, num_indices, e_rel))
}

hipError_t launch_resolve(hipStream_t st, int lang, const DeferArgs &a, unsigned waves) {
  CUBFF_DEFER_LANG(lang, resolve_lang, (st, a, waves))
}

}  // namespace cubff
