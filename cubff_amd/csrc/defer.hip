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
// The resolver runs an entry only once the entries it depends on are done:
// each entry counts its unfinished dependencies and links itself to each
// of them (one successor per program of the dependency); whoever finishes
// the last dependency -- the resolver wavefront that ran it, or the
// creator if it was already done when linking -- puts the entry on a
// ready queue, from which resolver wavefronts take work.  No wavefront
// ever holds an entry while it waits, so however long the chains behind a
// long interaction grow, every wavefront stays free for runnable work.
//
// owner[p] names p's pending writer: kNone (the last writer was an epoch
// kernel, visible to later kernels of the same stream), a list index (a
// deferred interaction will write p), or kResolved (the resolver wrote p: an
// epoch kernel reading it takes an acquire fence).  A deferred interaction
// records the pending writers of its two programs as dependencies.  The
// resolver publishes a result before marking it done and handing the
// programs back (owner: index -> kResolved).  List entries carry the run's
// generation number in their ready/done words, so the list is never
// cleared.  Every wait is bounded by a wall-clock limit; past it the abort
// flag stops the resolver and the host reports the error.
//
// No cache-maintaining fence is used between the two kernels.  On gfx950 an
// agent-scope acquire or release writes back or invalidates the whole L2 of
// the wavefront's XCD, which slowed every memory-bound kernel running
// beside them (the pairing kernels of later epochs, 20x).  Instead, every
// value that crosses between the concurrently running kernels -- list
// entries, the soup rows the resolver reads and writes, the rows of
// programs the resolver wrote -- is read and written with agent-scope
// (device-coherent) atomic accesses, and a writer waits until its stores
// have completed (s_waitcnt vmcnt(0)) before it raises the flag that
// publishes them; a reader issues its data loads only after it has read
// the flag.  Rows that only ever pass from one epoch kernel to a later one
// on the same stream use plain accesses (the end of a kernel writes back
// its L2).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "kernels.h"
#include "bff_fast.h"
#include "bff_lean.h"
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

// Every store this wavefront has issued has completed: s_waitcnt vmcnt(0)
// (expcnt and lgkmcnt left at their maxima).  The workgroup-scope fences
// around it emit no instruction; they keep the compiler from moving memory
// accesses across this point.
__device__ __forceinline__ void publish_point() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0x0F70);
}
// After a flag was read: the compiler keeps later loads after it (the
// hardware issues them after the flag's value decided the branch).
__device__ __forceinline__ void after_flag() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }

// A soup row moved between the soup and a lane's tape with 8-byte
// device-coherent accesses.
template <class Tape>
__device__ __forceinline__ void load_row_coherent(const Tape &t, int half, const uint8_t *src) {
  uint64_t *s = reinterpret_cast<uint64_t *>(const_cast<uint8_t *>(src));
  uint64_t v[kTape / 8];
#pragma unroll
  for (int q = 0; q < kTape / 8; q++)
    v[q] = __hip_atomic_load(s + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int q = 0; q < kTape / 8; q++) {
    t.set_word(half * 16 + 2 * q, (uint32_t)v[q]);
    t.set_word(half * 16 + 2 * q + 1, (uint32_t)(v[q] >> 32));
  }
}
template <class Tape>
__device__ __forceinline__ void store_row_coherent(const Tape &t, int half, uint8_t *dst) {
  uint64_t *d = reinterpret_cast<uint64_t *>(dst);
#pragma unroll
  for (int q = 0; q < kTape / 8; q++) {
    const uint64_t v = (uint64_t)t.word(half * 16 + 2 * q) |
                       ((uint64_t)t.word(half * 16 + 2 * q + 1) << 32);
    __hip_atomic_store(d + q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Puts list entry x on the ready queue.
__device__ __forceinline__ void push_ready(const DeferArgs &a, uint32_t x) {
  const uint32_t q = __hip_atomic_fetch_add(&a.ctl->qtail, 1u, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
  if (q < a.capacity) st_agent(a.queue + q, x);
}

// Waits until *p == want, or until the wall clock passes `deadline`;
// false then (nothing is raised: the caller goes on without).
__device__ bool wait_eq_until(const DeferCtl *c, const uint32_t *p, uint32_t want,
                              uint64_t deadline) {
  uint32_t nap = 1;
  for (;;) {
    if (ld_agent(p) == want) return true;
    if (ld_agent(&c->abort) || wall_clock64() > deadline) return false;
    for (uint32_t q = 0; q < nap; q++) __builtin_amdgcn_s_sleep(1);
    nap = min(nap * 2, 16u);
  }
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

// Progress of the epoch kernels, which the resolver's stall detection
// watches: counters[1] (num_runs), bumped once per epoch kernel, for one
// soup; the control block's heartbeat word for batched soups.
template <bool kBatch>
__device__ __forceinline__ uint64_t heartbeat(const DeferArgs &a) {
  if constexpr (kBatch)
    return __hip_atomic_load(&a.ctl->pad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    return __hip_atomic_load(&a.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
    // A program owed by an entry the resolver has not finished is waited
    // for, at most owe_ticks: a long interaction deferred at the end of one
    // epoch is mostly resolved early in the next, and every interaction
    // deferred behind it would make two more programs owed.  While the
    // resolver is behind (no budget), wait for owing entries for as long as
    // it takes instead of deferring more: the resolver never waits on epoch
    // kernels, so the wait ends, and the list cannot grow without bound.
    const bool wait = budget == 0xffffffffu;
    const uint64_t deadline = wall_clock64() + a.owe_ticks;
    if (o1 < kResolved &&
        (ld_agent(&a.items[o1].done) == a.gen ||
         (wait ? wait_eq(a.ctl, &a.items[o1].done, a.gen, a.spin_ticks)
               : wait_eq_until(a.ctl, &a.items[o1].done, a.gen, deadline))))
      o1 = kResolved;
    if (o2 < kResolved &&
        (ld_agent(&a.items[o2].done) == a.gen ||
         (wait ? wait_eq(a.ctl, &a.items[o2].done, a.gen, a.spin_ticks)
               : wait_eq_until(a.ctl, &a.items[o2].done, a.gen, deadline))))
      o2 = kResolved;
  }
  // Programs last written by the resolver are read with coherent loads,
  // issued after the owner word or done flag that said so.
  after_flag();
  bool defer = false;
  if (k < a.slots) {
    defer = o1 < kResolved || o2 < kResolved;  // a program is still owed
    if (!defer) {
      if (o1 == kResolved) load_row_coherent(t, 0, a.soup + (uint64_t)p1 * kTape);
      else load_row(t, 0, a.soup + (uint64_t)p1 * kTape);
      if (o2 == kResolved) load_row_coherent(t, 1, a.soup + (uint64_t)p2 * kTape);
      else load_row(t, 1, a.soup + (uint64_t)p2 * kTape);
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
  // A slot pairing a program with itself (a 2-D pattern may, main.cc:289-296)
  // has one program and so at most one dependency.
  const uint32_t d1 = o1 < kResolved ? o1 : kNone;
  const uint32_t d2 = (o2 < kResolved && p2 != p1) ? o2 : kNone;
  DeferItem *it = a.items + (put ? at : 0u);
  if (put) {
    st_agent(&it->e, e_rel | (eval ? 0u : kNoEval));
    st_agent(&it->k, (uint32_t)k);
    st_agent(&it->p1, p1);
    st_agent(&it->p2, p2);
    st_agent(&it->dep1, d1);
    st_agent(&it->dep2, d2);
    st_agent(&it->succ1, kNone);
    st_agent(&it->succ2, kNone);
    st_agent(&it->pending, (uint32_t)(d1 != kNone) + (uint32_t)(d2 != kNone) + 1u);
  }
  if (__ballot(put)) publish_point();
  if (put) {
    st_agent(&it->ready, a.gen);
    st_agent(a.owner + p1, at);
    st_agent(a.owner + p2, at);
    // Link to each dependency (the successor slot of the program this
    // entry takes over); one already done counts as satisfied.  Then drop
    // the creation count with the satisfied ones: at zero, the entry is
    // runnable.
    uint32_t sat = 0;
    if (d1 != kNone) {
      DeferItem *d = a.items + d1;
      uint32_t *slot = ld_agent(&d->p1) == p1 ? &d->succ1 : &d->succ2;
      sat += __hip_atomic_exchange(slot, at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             kLinkDone;
    }
    if (d2 != kNone) {
      DeferItem *d = a.items + d2;
      uint32_t *slot = ld_agent(&d->p1) == p2 ? &d->succ1 : &d->succ2;
      sat += __hip_atomic_exchange(slot, at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             kLinkDone;
    }
    if (__hip_atomic_fetch_sub(&it->pending, sat + 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) == sat + 1u)
      push_ready(a, at);
  }
  const uint64_t wsum = wave_sum_u64(ops);
  const uint64_t wsteps = wave_sum_u64(steps);
  if ((lane & 63) == 0 && wsum) atomicAdd(&block_ops, (unsigned long long)wsum);
  if ((lane & 63) == 0 && wsteps) atomicAdd(&block_steps, (unsigned long long)wsteps);
  __syncthreads();
  if (lane == 0 && block_ops) atomicAdd(&a.counters[0], block_ops);
  if (lane == 0 && block_steps) atomicAdd(&a.counters[2], block_steps);
  if (blockIdx.x == 0 && lane == 0) atomicAdd(&a.counters[1], (unsigned long long)*num_indices_p);
}

// k_epoch_defer for the languages BffLean runs: the two programs are
// gathered and mutated in registers, the word table is built from them in
// LDS, and the budgeted run and the write-back go through it.  The
// deferral protocol is k_epoch_defer's.
template <int LANG>
__global__ __launch_bounds__(kEpochLanes) void k_epoch_defer_lean(
    DeferArgs a, const uint32_t *__restrict__ shuf, const uint64_t *__restrict__ num_indices_p,
    uint32_t e_rel) {
  using F = typename LeanOf<LANG>::T;
  using W = LeanLds<kEpochLanes>;
  __shared__ __align__(16) uint8_t lds[W::kRegionBytes];
  __shared__ uint8_t lut[256];
  __shared__ unsigned long long block_ops, block_steps;
  __shared__ uint32_t block_budget, block_wait;
  const int lane = threadIdx.x;
  fill_lean_lut<LANG>(lut);
  if (lane == 0) {
    block_ops = 0;
    block_steps = 0;
    // While the resolver is behind, wait for owing entries for as long as
    // it takes instead of deferring behind them: no chain grows, and only
    // interactions past the budget, which depend on nothing listed, are
    // added (the resolver runs those side by side).  Complete runs here
    // would go without the cycle skip, which the resolver's have.  Without
    // a resolver (abort raised) nothing is given up.
    const uint32_t backlog = ld_agent(&a.ctl->tail) - ld_agent(&a.ctl->resolved);
    const bool abort = ld_agent(&a.ctl->abort) != 0;
    block_budget = abort ? 0xffffffffu : a.budget;
    block_wait = abort || backlog > a.max_backlog;
  }
  __syncthreads();
  const uint32_t budget = block_budget;
  const uint64_t k = (uint64_t)blockIdx.x * kEpochLanes + lane;
  const W w = W::at(lds, lane);
  const LdsByteLut blut{(const lds_u8 *)lut};
  uint32_t ops = 0, steps = 0, p1 = 0, p2 = 0, o1 = kNone, o2 = kNone;
  const bool eval = k < *num_indices_p;  // common_language.h:185
  const uint64_t t_start = wall_clock64();
  const uint64_t deadline = t_start + a.owe_ticks;
  // Runs stop at the budget or this long after the kernel's start,
  // whichever comes first (a stopped run loses nothing: the resolver goes on
  // from it).
  const uint64_t run_deadline =
      a.run_ticks && budget != 0xffffffffu ? t_start + a.run_ticks : ~0ull;
  if (k < a.slots) {
    p1 = shuf[2 * k];
    p2 = shuf[2 * k + 1];
    o1 = ld_agent(a.owner + p1);
    o2 = ld_agent(a.owner + p2);
    if (o1 < kResolved && ld_agent(&a.items[o1].done) == a.gen) o1 = kResolved;
    if (o2 < kResolved && ld_agent(&a.items[o2].done) == a.gen) o2 = kResolved;
  }
  after_flag();
  bool defer = false;
  const uint64_t seed_value = epoch_seed_pre(a.sm_seed, a.first_epoch + e_rel);
  // Two passes: interactions whose programs are both free run in the first;
  // those that owe a program wait for it after that (so the wait overlaps
  // the other lanes' runs: a wavefront's time is the longer of the two, not
  // their sum) and run in the second, or are deferred if it is still owed
  // at the deadline.  Wavefronts with nothing owed skip the second pass.
  bool todo = k < a.slots;
  // A run stopped by its budget is handed to the resolver as it stands:
  // its tape goes to the programs' rows (which it owns) and `rs` to the
  // list entry, and the resolver goes on from there.
  bool resumable = false;
  typename F::Resume rs{};
#pragma unroll 1
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      if (!F::any_lane(todo)) break;
      if (todo) {
        const bool wait = block_wait != 0;
        if (o1 < kResolved && (wait ? wait_eq(a.ctl, &a.items[o1].done, a.gen, a.spin_ticks)
                                    : wait_eq_until(a.ctl, &a.items[o1].done, a.gen, deadline)))
          o1 = kResolved;
        if (o2 < kResolved && (wait ? wait_eq(a.ctl, &a.items[o2].done, a.gen, a.spin_ticks)
                                    : wait_eq_until(a.ctl, &a.items[o2].done, a.gen, deadline)))
          o2 = kResolved;
        defer = o1 < kResolved || o2 < kResolved;  // a program is still owed
        todo = !defer;
      }
      after_flag();
    }
    if (todo && o1 >= kResolved && o2 >= kResolved) {
      todo = false;
      if (o1 == kResolved) load_raw_row_coherent(w, 0, a.soup + (uint64_t)p1 * kTape);
      else load_raw_row(w, 0, a.soup + (uint64_t)p1 * kTape);
      if (o2 == kResolved) load_raw_row_coherent(w, 1, a.soup + (uint64_t)p2 * kTape);
      else load_raw_row(w, 1, a.soup + (uint64_t)p2 * kTape);
      mutate_raw(w, a.n, seed_value, k, a.prob);
      if (eval) ops = F::evaluate(w, blut, kEvalSteps, &steps, nullptr, budget, &rs, run_deadline);
      else raw_as_final(w);  // not evaluated: the mutated tape goes back as it is
      if (steps == kOverBudget) {
        ops = steps = 0;
        defer = true;
        // Two halves of one row cannot both be kept: such a run starts over.
        resumable = p1 != p2;
      } else {
        store_lean_row(w, 0, a.soup + (uint64_t)p1 * kTape);
        store_lean_row(w, 1, a.soup + (uint64_t)p2 * kTape);
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
  // The resolver left before anything was deferred: run the interaction
  // here to the end, from the soup's unchanged bytes.
  const bool gone = defer && at >= kDeferGone;
  if (gone) {
    load_raw_row(w, 0, a.soup + (uint64_t)p1 * kTape);
    load_raw_row(w, 1, a.soup + (uint64_t)p2 * kTape);
    mutate_raw(w, a.n, seed_value, k, a.prob);
    if (eval) ops = F::evaluate(w, blut, kEvalSteps, &steps, nullptr, 0xffffffffu);
    else raw_as_final(w);
    store_lean_row(w, 0, a.soup + (uint64_t)p1 * kTape);
    store_lean_row(w, 1, a.soup + (uint64_t)p2 * kTape);
    if (e_rel + 1 == a.epochs) {
      a.pair_ops[k] = ops;
      a.pair_steps[k] = steps;
    }
    defer = false;
  }
  const bool put = defer && at < a.capacity;
  if (defer && !put) st_agent(&a.ctl->abort, 1);
  const uint32_t d1 = o1 < kResolved ? o1 : kNone;
  const uint32_t d2 = (o2 < kResolved && p2 != p1) ? o2 : kNone;
  DeferItem *it = a.items + (put ? at : 0u);
  if (put) {
    st_agent(&it->e, e_rel | (eval ? 0u : kNoEval) | (resumable ? kResume : 0u));
    if (resumable) {
      store_lean_row_coherent(w, 0, a.soup + (uint64_t)p1 * kTape);
      store_lean_row_coherent(w, 1, a.soup + (uint64_t)p2 * kTape);
      st_agent(&it->ops, rs.pos);
      st_agent(&it->steps, rs.i);
      st_agent(&it->pad[0], rs.iters);
      st_agent(&it->pad[1], rs.extra);
    }
    st_agent(&it->k, (uint32_t)k);
    st_agent(&it->p1, p1);
    st_agent(&it->p2, p2);
    st_agent(&it->dep1, d1);
    st_agent(&it->dep2, d2);
    st_agent(&it->succ1, kNone);
    st_agent(&it->succ2, kNone);
    st_agent(&it->pending, (uint32_t)(d1 != kNone) + (uint32_t)(d2 != kNone) + 1u);
  }
  if (__ballot(put)) publish_point();
  if (put) {
    st_agent(&it->ready, a.gen);
    st_agent(a.owner + p1, at);
    st_agent(a.owner + p2, at);
    uint32_t sat = 0;
    if (d1 != kNone) {
      DeferItem *d = a.items + d1;
      uint32_t *slot = ld_agent(&d->p1) == p1 ? &d->succ1 : &d->succ2;
      sat += __hip_atomic_exchange(slot, at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             kLinkDone;
    }
    if (d2 != kNone) {
      DeferItem *d = a.items + d2;
      uint32_t *slot = ld_agent(&d->p1) == p2 ? &d->succ1 : &d->succ2;
      sat += __hip_atomic_exchange(slot, at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             kLinkDone;
    }
    if (__hip_atomic_fetch_sub(&it->pending, sat + 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) == sat + 1u)
      push_ready(a, at);
  }
  const uint64_t wsum = wave_sum_u64(ops);
  const uint64_t wsteps = wave_sum_u64(steps);
  if ((lane & 63) == 0 && wsum) atomicAdd(&block_ops, (unsigned long long)wsum);
  if ((lane & 63) == 0 && wsteps) atomicAdd(&block_steps, (unsigned long long)wsteps);
  __syncthreads();
  if (lane == 0 && block_ops) atomicAdd(&a.counters[0], block_ops);
  if (lane == 0 && block_steps) atomicAdd(&a.counters[2], block_steps);
  if (blockIdx.x == 0 && lane == 0) atomicAdd(&a.counters[1], (unsigned long long)*num_indices_p);
}

// Mutation parameters of slot g of batched soups in epoch `epoch`
// (common_language.h:166-180 per soup): the soup's size and epoch seed, the
// slot's index within its soup, the soup's threshold, whether it is
// evaluated (common_language.h:185).
struct SlotSoup {
  uint32_t s;
  uint64_t n, seed_value, kl;
  uint32_t prob;
  bool eval;
};
__device__ __forceinline__ SlotSoup slot_soup(const DeferArgs &a, uint64_t g, uint64_t epoch,
                                              const uint64_t *slot_base) {
  SlotSoup r;
  r.s = soup_of(slot_base, a.num_soups, g);
  r.n = a.prog_base[r.s + 1] - a.prog_base[r.s];
  r.seed_value = epoch_seed_pre(a.sm_seeds[r.s], epoch);
  r.kl = g - a.slot_base[r.s];
  r.prob = a.probs[r.s];
  r.eval = r.kl < a.num_idx[r.s];
  return r;
}

// The deferring epoch kernel of batched soups (languages BffFast runs): a
// persistent kernel whose lanes take slots of a.list[0, a.slots) as they
// free up (an atomic on *a.claim per wavefront refill), each slot starting
// from the tape the stage kernel gathered and mutated.  A lane moves
// through the states
//   free -> wait   (claimed; its programs' pending writers are checked
//                   without blocking the wavefront: a program owed by an
//                   unfinished list entry is waited for until owe_ticks
//                   after the claim, then the slot is deferred)
//        -> run    (a resumable BffFast run advanced kRunSlice commands per
//                   turn; past the command budget the slot is deferred)
//        -> done / defer -> free,
// and wavefronts write back finished slots and create the list entries of
// deferred ones together, once enough lanes are waiting for that (or
// nothing else is left).  A half whose program the resolver wrote is read
// again, coherently, and mutated again (the stage may have read it while it
// was being written).  Entries, links and the ready queue follow
// k_epoch_defer's protocol.
constexpr uint32_t kRunSlice = 8;
constexpr uint32_t kServiceMin = 16;
constexpr uint32_t kSlotBaseLds = 256;  // soups + 1 whose slot bases fit in LDS
enum : uint32_t { kLaneFree = 0, kLaneWait, kLaneRun, kLaneDone, kLaneDefer };

template <int LANG, int LANES>
__global__ __launch_bounds__(LANES) void k_b_refill_defer(DeferArgs a, uint32_t e_rel) {
  using L = typename LangOf<LANG>::T;
  using F = typename FastOf<LANG>::T;
  using Tape = typename L::template Tape<LANES>;
  constexpr int kNx = (Tape::kRegionBytes + 15) / 16 * 16;
  constexpr int kAnchor = kNx + (LinTape<LANES>::kRegionBytes + 15) / 16 * 16;
  __shared__ __align__(16) uint8_t lds[kAnchor + LdsTape<LANES>::kRegionBytes];
  __shared__ uint8_t lut[256];
  __shared__ uint64_t sbase_lds[kSlotBaseLds];
  __shared__ uint32_t block_budget;
  const int lane = threadIdx.x, wl = lane & 63;
  fill_fast_lut<LANG>(lut);
  // The soups' slot bases in LDS for the per-slot soup lookups.
  const bool sb_lds = a.num_soups + 1 <= kSlotBaseLds;
  if (sb_lds)
    for (uint32_t x = lane; x <= a.num_soups; x += LANES) sbase_lds[x] = a.slot_base[x];
  const uint64_t *slot_base = sb_lds ? sbase_lds : a.slot_base;
  if (lane == 0) {
    // While the resolver is behind, run every slot to the end here and wait
    // for owed programs as long as it takes (see k_epoch_defer).
    const uint32_t backlog = ld_agent(&a.ctl->tail) - ld_agent(&a.ctl->resolved);
    block_budget = backlog > a.max_backlog || ld_agent(&a.ctl->abort) ? 0xffffffffu : a.budget;
    if (blockIdx.x == 0)
      __hip_atomic_fetch_add(&a.ctl->pad, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const uint32_t budget0 = block_budget;
  const bool wait_all = budget0 == 0xffffffffu;
  const auto t = Tape::at(lds, lane);
  const auto nx = LinTape<LANES>::at(lds + kNx, lane);
  const auto anchor = LdsTape<LANES>::at(lds + kAnchor, lane);
  const LdsByteLut blut{(const lds_u8 *)lut};
  const uint64_t epoch = a.first_epoch + e_rel;
  typename F::Run r;
  r.done = true;
  r.iters = 0;
  uint32_t st = kLaneFree, g = 0, p1 = 0, p2 = 0, o1 = kNone, o2 = kNone, budget = budget0;
  // The lane's run was stopped by the budget (not deferred while waiting):
  // the resolver goes on from its tape and state.
  bool stopped = false;
  uint64_t deadline = 0;
  SlotSoup q{};
  bool exhausted = false;
  // The slot's input tape: the staged one, with the halves the resolver
  // wrote read again; then the start of its run.
  const auto start_slot = [&]() {
    after_flag();
    const uint4 *src = reinterpret_cast<const uint4 *>(a.stage + (uint64_t)g * kTape2);
    uint32_t words[kTape2 / 4];
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const uint4 v = src[w];
      words[4 * w] = v.x;
      words[4 * w + 1] = v.y;
      words[4 * w + 2] = v.z;
      words[4 * w + 3] = v.w;
    }
#pragma unroll
    for (int w = 0; w < kTape2 / 4; w++) t.set_word(w, words[w]);
    const bool reread = o1 == kResolved || o2 == kResolved;
    if (o1 == kResolved) {
      load_row_coherent(t, 0, a.soup + (uint64_t)p1 * kTape);
      mutate_half(t, 0, q.n, q.seed_value, q.kl, q.prob);
    }
    if (o2 == kResolved) {
      load_row_coherent(t, 1, a.soup + (uint64_t)p2 * kTape);
      mutate_half(t, 1, q.n, q.seed_value, q.kl, q.prob);
    }
    if (!q.eval) {
      r.done = true;
      r.iters = 0;
    } else if (reread) {
      F::start(r, t, nx, blut, kEvalSteps);
    } else {
      F::start_words(r, words, nx, blut, kEvalSteps);
    }
    st = kLaneRun;
  };
  for (;;) {
    const uint64_t need = __ballot(st == kLaneFree || st == kLaneDone || st == kLaneDefer);
    const uint32_t nneed = (uint32_t)__popcll(need);
    const bool busy = __any(st == kLaneRun || st == kLaneWait);
    if (nneed >= kServiceMin || (nneed > 0 && !busy)) {
      if (st == kLaneDone) {  // write back (common_language.h:186-192)
        uint32_t steps = 0, work = 0, ops = 0;
        if (q.eval) ops = F::finish(r, kEvalSteps, &steps, &work);
        store_row(t, 0, a.soup + (uint64_t)p1 * kTape);
        store_row(t, 1, a.soup + (uint64_t)p2 * kTape);
        if (o1 != kNone) st_agent(a.owner + p1, kNone);
        if (o2 != kNone) st_agent(a.owner + p2, kNone);
        a.pair_ops[g] = ops;
        a.pair_steps[g] = steps;
        a.pair_work[g] = work;
        st = kLaneFree;
      }
      const bool defer = st == kLaneDefer;
      const uint32_t at = wave_alloc_u32(&a.ctl->tail, defer ? 1u : 0u);
      // The resolver left before anything was deferred: run the slot here
      // to the end.
      if (defer && at >= kDeferGone) {
        budget = 0xffffffffu;
        start_slot();
      }
      const bool put = defer && at < a.capacity;
      if (defer && at >= a.capacity && at < kDeferGone) st_agent(&a.ctl->abort, 1);
      const uint32_t d1 = o1 < kResolved ? o1 : kNone;
      const uint32_t d2 = (o2 < kResolved && p2 != p1) ? o2 : kNone;
      DeferItem *it = a.items + (put ? at : 0u);
      // A stopped run's tape goes to its programs' rows (the entry owns
      // them from here on; a slot pairing a program with itself starts over)
      // and its state to the entry.
      const bool handed = put && stopped && p1 != p2;
      if (put) {
        st_agent(&it->e, e_rel | (q.eval ? 0u : kNoEval) | (handed ? kResume : 0u));
        if (handed) {
          store_row_coherent(t, 0, a.soup + (uint64_t)p1 * kTape);
          store_row_coherent(t, 1, a.soup + (uint64_t)p2 * kTape);
          st_agent(&it->ops, (uint32_t)r.pc | ((uint32_t)r.h0 << 8) | ((uint32_t)r.h1 << 16));
          st_agent(&it->steps, r.i);
          st_agent(&it->pad[0], r.iters);
          st_agent(&it->pad[1], r.extra);
        }
        st_agent(&it->k, g);
        st_agent(&it->p1, p1);
        st_agent(&it->p2, p2);
        st_agent(&it->dep1, d1);
        st_agent(&it->dep2, d2);
        st_agent(&it->succ1, kNone);
        st_agent(&it->succ2, kNone);
        st_agent(&it->pending, (uint32_t)(d1 != kNone) + (uint32_t)(d2 != kNone) + 1u);
      }
      if (__ballot(put)) publish_point();
      if (put) {
        st_agent(&it->ready, a.gen);
        st_agent(a.owner + p1, at);
        st_agent(a.owner + p2, at);
        uint32_t sat = 0;
        if (d1 != kNone) {
          DeferItem *d = a.items + d1;
          uint32_t *slot = ld_agent(&d->p1) == p1 ? &d->succ1 : &d->succ2;
          sat += __hip_atomic_exchange(slot, at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                 kLinkDone;
        }
        if (d2 != kNone) {
          DeferItem *d = a.items + d2;
          uint32_t *slot = ld_agent(&d->p1) == p2 ? &d->succ1 : &d->succ2;
          sat += __hip_atomic_exchange(slot, at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                 kLinkDone;
        }
        if (__hip_atomic_fetch_sub(&it->pending, sat + 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) == sat + 1u)
          push_ready(a, at);
        a.pair_ops[g] = 0;
        a.pair_steps[g] = 0;
        a.pair_work[g] = budget;
      }
      if (defer && !(at >= kDeferGone)) st = kLaneFree;
      if (!exhausted) {
        const uint64_t fm = __ballot(st == kLaneFree);
        const uint32_t nf = (uint32_t)__popcll(fm);
        uint32_t base = 0;
        if (wl == 0 && nf) base = atomicAdd(a.claim, nf);
        base = __shfl(base, 0, 64);
        exhausted = (uint64_t)base + nf >= a.slots;
        const uint64_t e = (uint64_t)base + __popcll(fm & ((1ull << wl) - 1));
        if (st == kLaneFree && e < a.slots) {
          stopped = false;
          g = a.list[e];
          q = slot_soup(a, g, epoch, slot_base);
          p1 = a.tgt[2 * (uint64_t)g];
          p2 = a.tgt[2 * (uint64_t)g + 1];
          o1 = ld_agent(a.owner + p1);
          o2 = ld_agent(a.owner + p2);
          budget = budget0;
          deadline = wall_clock64() + (wait_all ? a.spin_ticks : a.owe_ticks);
          st = kLaneWait;
        }
      }
    }
    if (!__any(st != kLaneFree)) {
      if (exhausted) break;
      continue;
    }
    if (st == kLaneWait) {
      if (o1 < kResolved && ld_agent(&a.items[o1].done) == a.gen) o1 = kResolved;
      if (o2 < kResolved && ld_agent(&a.items[o2].done) == a.gen) o2 = kResolved;
      if (o1 >= kResolved && o2 >= kResolved) {
        start_slot();
      } else if (ld_agent(&a.ctl->abort) || wall_clock64() > deadline) {
        if (wait_all && !ld_agent(&a.ctl->abort)) st_agent(&a.ctl->abort, 1);
        st = kLaneDefer;
      }
    }
    if (st == kLaneRun) {
      if (!r.done) F::template advance<true>(r, t, nx, anchor, blut, kEvalSteps, kRunSlice);
      if (r.done) {
        st = kLaneDone;
      } else if (r.iters >= budget) {
        st = kLaneDefer;
        stopped = true;
      }
    }
    // Only lanes waiting on owed programs: give the memory system a rest.
    if (!__any(st == kLaneRun) && __any(st == kLaneWait)) __builtin_amdgcn_s_sleep(2);
  }
}

// After a batched segment: the last epoch's per-slot counts of the
// entries the resolver ran.
__global__ void k_b_defer_fixup(DeferArgs a, uint32_t listed) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= listed) return;
  const DeferItem &it = a.items[x];
  if ((it.e & ~(kNoEval | kResume)) + 1 == a.epochs) {
    a.pair_ops[it.k] = it.ops;
    a.pair_steps[it.k] = it.steps;
  }
}

// The resolver: one deferred interaction per wavefront at a time, run by
// lane 0 (the list is short and its entries are the long interactions).
template <int LANG, bool kBatch>
__global__ __launch_bounds__(64) void k_resolve(DeferArgs a) {
  using L = typename LangOf<LANG>::T;
  using Tape = typename L::template Tape<1>;
  using Aux = typename L::template Aux<1>;
  constexpr bool kFast = FastOf<LANG>::kHas;
  // Tape, second region (aux, or nx for BffFast), and BffFast's cycle
  // anchor, each 16-byte aligned.
  constexpr int kSecond = (Tape::kRegionBytes + 15) / 16 * 16;
  constexpr int kThird = kSecond + (second_region_bytes<LANG, 1>() + 15) / 16 * 16;
  // The languages BffLean runs use its word table instead.
  constexpr bool kLean = LeanOf<LANG>::kHas;
  constexpr int kLeanAnchor = (LeanLds<1>::kRegionBytes + 15) / 16 * 16;
  constexpr int kBytes = kLean ? kLeanAnchor + LeanAnchorLds<1>::kRegionBytes
                               : kThird + (kFast ? LdsTape<1>::kRegionBytes : 0);
  __shared__ __align__(16) uint8_t lds[kBytes];
  __shared__ uint8_t lut[kFast ? 256 : 1];
  if constexpr (kLean) fill_lean_lut<LANG>(lut);
  else fill_fast_lut<LANG>(lut);
  __syncthreads();
  if (threadIdx.x != 0) return;
  // The resolver's one-lane wavefronts share SIMDs with the epoch kernel's
  // busy ones; issuing first, they finish the long interactions (whose
  // delay grows the chains of deferred interactions behind them) sooner.
  __builtin_amdgcn_s_setprio(3);
  const auto t = Tape::at(lds, 0);
  const auto aux = Aux::at(lds + kSecond, 0);
  DeferCtl *c = a.ctl;
  for (;;) {
    // Claim the next ready-queue slot by fetch-and-add, whether or not an
    // entry has been put there yet, then wait for it on that slot (the
    // control words all wavefronts share are read only every 8th poll, so
    // the waiting wavefronts do not crowd one memory channel).  Leave once
    // the list is closed with fewer entries than this slot, or past its
    // capacity.
    const uint32_t qi = __hip_atomic_fetch_add(&c->head, 1u, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
    if (qi >= a.capacity) break;
    uint32_t my = kNone;
    {
      // The epoch kernels bump counters[1] once each: while they progress,
      // waiting here is not a stall.
      uint64_t t0 = wall_clock64();
      uint64_t beat = heartbeat<kBatch>(a);
      uint32_t nap = 1;
      for (uint32_t poll = 0;; poll++) {
        const uint32_t v = ld_agent(a.queue + qi);
        if (v != kNone) {
          my = v;
          break;
        }
        if ((poll & 7) == 0) {
          if (ld_agent(&c->abort)) break;
          const uint32_t traw = ld_agent(&c->tail);
          if (traw >= kDeferGone) break;  // another wavefront left with the list empty
          if (ld_agent(&c->closed) && qi >= ld_agent(&c->final_tail)) break;
          const uint64_t b2 = heartbeat<kBatch>(a);
          if (b2 != beat) {
            beat = b2;
            t0 = wall_clock64();
          }
          // Nothing deferred yet and no epoch kernel finished for a while:
          // the epoch kernels cannot start beside this kernel (dispatches
          // are serialized, as under some profilers).  Leave, marking the
          // list gone in the same atomic step that checks it is still
          // empty; the epoch kernels then run every interaction themselves
          // (their allocations return kDeferGone or more).  If an epoch
          // kernel deferred something meanwhile, stay and resolve it.
          if (traw == 0 && wall_clock64() - t0 > a.idle_ticks) {
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
        }
        for (uint32_t q = 0; q < nap; q++) __builtin_amdgcn_s_sleep(1);
        nap = min(nap * 2, 64u);
      }
    }
    if (my == kNone) break;
    DeferItem *it = a.items + my;
    // Its fields were published before it could be queued, and every entry
    // it depends on was done before it was.
    after_flag();
    const uint32_t ew = ld_agent(&it->e);
    const uint32_t e = ew & ~(kNoEval | kResume), k = ld_agent(&it->k), p1 = ld_agent(&it->p1),
                   p2 = ld_agent(&it->p2);
    const bool eval = !(ew & kNoEval);
    uint32_t s = 0, steps = 0, ops = 0;
    if constexpr (kLean) {
      using F = typename LeanOf<LANG>::T;
      const auto w = LeanLds<1>::at(lds, 0);
      const auto anchor = LeanAnchorLds<1>::at(lds + kLeanAnchor, 0);
      const LdsByteLut blut{(const lds_u8 *)lut};
      load_raw_row_coherent(w, 0, a.soup + (uint64_t)p1 * kTape);
      load_raw_row_coherent(w, 1, a.soup + (uint64_t)p2 * kTape);
      // An entry an epoch kernel's budget stopped goes on from its tape,
      // which is in the rows, and the state in the entry; others start
      // from the mutated programs.
      const bool resume = (ew & kResume) != 0;
      if constexpr (kBatch) {
        const SlotSoup q = slot_soup(a, k, a.first_epoch + e, a.slot_base);
        s = q.s;
        if (!resume) mutate_raw(w, q.n, q.seed_value, q.kl, q.prob);
      } else {
        if (!resume) mutate_raw(w, a.n, epoch_seed_pre(a.sm_seed, a.first_epoch + e), k, a.prob);
      }
      if (eval) {
        typename F::Run r;
        F::start_raw(r, w, blut, kEvalSteps);
        if (resume)
          F::resume_from(r,
                         typename F::Resume{ld_agent(&it->ops), ld_agent(&it->steps),
                                            ld_agent(&it->pad[0]), ld_agent(&it->pad[1])},
                         kEvalSteps);
        F::template advance<false, true>(r, w, blut, kEvalSteps, 0xffffffffu, anchor);
        ops = F::finish(r, kEvalSteps, &steps, nullptr);
      } else {
        raw_as_final(w);
      }
      store_lean_row_coherent(w, 0, a.soup + (uint64_t)p1 * kTape);
      store_lean_row_coherent(w, 1, a.soup + (uint64_t)p2 * kTape);
    } else {
      load_row_coherent(t, 0, a.soup + (uint64_t)p1 * kTape);
      load_row_coherent(t, 1, a.soup + (uint64_t)p2 * kTape);
      if constexpr (kBatch) {
        const SlotSoup q = slot_soup(a, k, a.first_epoch + e, a.slot_base);
        s = q.s;
        mutate(t, q.n, q.seed_value, q.kl, q.prob);
      } else {
        mutate(t, a.n, epoch_seed_pre(a.sm_seed, a.first_epoch + e), k, a.prob);
      }
      if (eval) {
        if constexpr (kFast)
          ops = FastOf<LANG>::T::evaluate_full(t, LinTape<1>::at(lds + kSecond, 0),
                                               LdsTape<1>::at(lds + kThird, 0),
                                               LdsByteLut{(const lds_u8 *)lut}, kEvalSteps,
                                               &steps, nullptr);
        else
          ops = L::evaluate(t, aux, kEvalSteps, &steps);
      }
      store_row_coherent(t, 0, a.soup + (uint64_t)p1 * kTape);
      store_row_coherent(t, 1, a.soup + (uint64_t)p2 * kTape);
    }
    unsigned long long *ctr = a.counters + (kBatch ? 4 * (uint64_t)s : 0);
    if (ops) atomicAdd(&ctr[0], (unsigned long long)ops);
    if (steps) atomicAdd(&ctr[2], (unsigned long long)steps);
    if constexpr (kBatch) {
      // The slot's per-epoch counts are written after the segment, from here.
      st_agent(&it->ops, ops);
      st_agent(&it->steps, steps);
    } else if (e + 1 == a.epochs) {
      a.pair_ops[k] = ops;
      a.pair_steps[k] = steps;
    }
    publish_point();
    st_agent(&it->done, a.gen);
    __hip_atomic_fetch_add(&c->resolved, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Release the entries waiting on this one: the last dependency done
    // makes an entry runnable.
    for (int h = 0; h < 2; h++) {
      const uint32_t z = __hip_atomic_exchange(h ? &it->succ2 : &it->succ1, kLinkDone,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (z != kNone && __hip_atomic_fetch_sub(&a.items[z].pending, 1u, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT) == 1u)
        push_ready(a, z);
    }
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
  c->qtail = 0;
}

// After the segment's last epoch kernel: no entry is appended any more.
__global__ void k_defer_close(DeferCtl *c) {
  st_agent(&c->final_tail, ld_agent(&c->tail));
  publish_point();
  st_agent(&c->closed, 1);
}

namespace {
inline unsigned blocks_of(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
}  // namespace

template <int L>
hipError_t epoch_defer_lang(hipStream_t st, const DeferArgs &a, const uint32_t *shuf,
                            const uint64_t *num_indices, uint32_t e_rel) {
  if constexpr (LeanOf<L>::kHas)
    hipLaunchKernelGGL(k_epoch_defer_lean<L>, dim3(blocks_of(a.slots, kEpochLanes)),
                       dim3(kEpochLanes), 0, st, a, shuf, num_indices, e_rel);
  else
    hipLaunchKernelGGL(k_epoch_defer<L>, dim3(blocks_of(a.slots, kEpochLanes)),
                       dim3(kEpochLanes), 0, st, a, shuf, num_indices, e_rel);
  return hipGetLastError();
}

template <int L>
hipError_t resolve_lang(hipStream_t st, const DeferArgs &a, unsigned waves) {
  if (a.num_soups)
    hipLaunchKernelGGL((k_resolve<L, true>), dim3(waves), dim3(64), 0, st, a);
  else
    hipLaunchKernelGGL((k_resolve<L, false>), dim3(waves), dim3(64), 0, st, a);
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

bool batch_defers(int lang) {
  return lang == kBff || lang == kBffNoheads || lang == kBffNoheads4bit || lang == kBffPerm;
}

template <int LANES>
hipError_t b_refill_defer_lanes(hipStream_t st, int lang, const DeferArgs &a, uint32_t e_rel,
                                uint32_t waves_per_simd) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
      return hipErrorInvalidDevice;
    cus = v;
  }
  const uint64_t want = (uint64_t)cus * 4 * waves_per_simd * 64 / LANES;
  const dim3 grid((unsigned)std::min<uint64_t>(want, blocks_of(a.slots, LANES))), block(LANES);
  switch (lang) {
#define CUBFF_RD_CASE(LG) \
  case LG: hipLaunchKernelGGL((k_b_refill_defer<LG, LANES>), grid, block, 0, st, a, e_rel); break;
    CUBFF_RD_CASE(kBff)
    CUBFF_RD_CASE(kBffNoheads)
    CUBFF_RD_CASE(kBffNoheads4bit)
    CUBFF_RD_CASE(kBffPerm)
#undef CUBFF_RD_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_b_epoch_defer(hipStream_t st, int lang, const DeferArgs &a, uint32_t e_rel,
                                int wg, uint32_t waves_per_simd) {
  if (a.slots == 0) return hipSuccess;
  if (!a.num_soups || !a.list || !a.stage || !a.tgt || !a.claim) return hipErrorInvalidValue;
  if (waves_per_simd == 0) waves_per_simd = 1;
  switch (wg) {
    case 64: return b_refill_defer_lanes<64>(st, lang, a, e_rel, waves_per_simd);
    case 128: return b_refill_defer_lanes<128>(st, lang, a, e_rel, waves_per_simd);
    case 256: return b_refill_defer_lanes<256>(st, lang, a, e_rel, waves_per_simd);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_b_defer_fixup(hipStream_t st, const DeferArgs &a, uint32_t listed) {
  if (listed == 0) return hipSuccess;
  hipLaunchKernelGGL(k_b_defer_fixup, dim3(blocks_of(listed, 256)), dim3(256), 0, st, a, listed);
  return hipGetLastError();
}

}  // namespace cubff
