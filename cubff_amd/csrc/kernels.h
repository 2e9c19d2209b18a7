// Launch helpers for the cubff HIP kernels (host side of kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace cubff {

constexpr int kEpochLanes = 256;      // lanes (= pair slots) per epoch block
constexpr int kSelfrepProgs = 16;     // programs per self-rep block
constexpr int kSelfrepLanes = 256;    // >= kSelfrepProgs * 13, multiple of 64
constexpr int kNumCounters = 4;       // ops, num_runs, steps, unused

// Device scratch of the Fisher-Yates pairing; every array has n entries
// except cnt, which has n + 1 (the last entry is the bucket allocator).
struct FyScratch {
  uint32_t *j, *cnt, *base, *cursor, *items, *first, *succ;
};

// Device scratch of the 2-D greedy pairing (n entries each; used is padded
// to a multiple of 4 bytes).
struct G2Scratch {
  uint32_t *order, *eu, *ew, *minpri, *rank, *undecided;
  uint8_t *state, *used;
  uint64_t *rest_total;
};

// Control block of the multi-epoch dataflow kernel: queue head (next entry
// to claim) and tail (next free entry), adjacent and 8-byte aligned so they
// can be read together; the abort flag; claims made; items done.
struct alignas(16) FlowCtl {
  uint32_t head, tail, abort, claims, done, pad[3];
};

// Arguments of the multi-epoch dataflow kernel (k_flow in flow.hip) for a
// segment of `epochs` epochs starting at first_epoch: shuf and pairof hold
// epochs x n entries (pairing of segment epoch j, and the pair of each
// program in it), cnt and queue epochs x n/2.
struct FlowArgs {
  uint8_t *soup;
  const uint32_t *shuf;
  const uint32_t *pairof;
  uint32_t *cnt;
  uint32_t *queue;
  FlowCtl *ctl;
  uint64_t n;
  uint32_t slots, epochs;
  uint64_t sm_seed, first_epoch;
  uint32_t prob;
  uint32_t *pair_ops, *pair_steps;  // written for the segment's last epoch
  unsigned long long *counters;
  uint64_t spin_ticks;      // wall-clock ticks any single wait may take
  uint64_t patience_ticks;  // wait for a full run of 64 items this long
  uint32_t fixed;           // epoch-0 items handed to the waves directly (set at launch)
};

// Deferred-interaction path (defer.hip).  Control block: ready-queue
// slots claimed by the resolver (head), list entries appended (tail), the
// closed flag and the list length at closing, the abort flag, entries
// resolved, ready-queue slots filled (qtail).
struct alignas(16) DeferCtl {
  uint32_t head, tail, closed, final_tail, abort, resolved, qtail, pad;
};

// One deferred interaction: segment epoch (kNoEval set for a slot past the
// epoch's evaluated pairs: mutated and written back only), pair slot, its
// programs, the list entries that owe them (kNone if none), and the run
// generation once the entry is complete (ready) and once its result is
// written (done); the entries that wait on this one for its first and its
// second program (kNone until one links, kLinkDone once this entry is
// done), and the count of its unfinished dependencies (plus one while its
// creator links it); the interaction's ops and steps once run (batches).
struct DeferItem {
  uint32_t e, k, p1, p2, dep1, dep2, ready, done;
  uint32_t succ1, succ2, pending, ops, steps, pad[3];
};
constexpr uint32_t kLinkDone = 0xfffffffdu;

constexpr uint32_t kResolved = 0xfffffffeu;  // owner[p]: last written by the resolver
constexpr uint32_t kNoEval = 0x80000000u;    // DeferItem::e flag
// DeferItem::e flag of an entry stopped by its command budget: its tape is
// in its programs' rows, and ops, steps, pad[0], pad[1] hold the stopped
// run's position and heads, steps, commands and accelerated ones (a batched
// entry's ops and steps then receive its results).
constexpr uint32_t kResume = 0x40000000u;
// DeferCtl::tail once the resolver has left without anything deferred: it
// sets tail from 0 to this value in one compare-and-swap, so an allocation
// either lands before (the resolver stays) or returns at least kDeferGone
// (the epoch kernel runs that interaction itself).  Above any capacity.
constexpr uint32_t kDeferGone = 0x80000000u;

struct DeferArgs {
  uint8_t *soup;
  uint32_t *owner;  // n entries
  DeferItem *items;
  uint32_t *queue;  // capacity slots: entries whose dependencies are done, kNone when empty
  DeferCtl *ctl;
  uint64_t n;
  uint32_t slots, epochs, capacity, budget, gen, prob;
  uint32_t max_backlog;  // entries not yet resolved past which nothing more is given up
  uint64_t sm_seed, first_epoch;
  uint32_t *pair_ops, *pair_steps;  // written for the segment's last epoch
  unsigned long long *counters;
  uint64_t spin_ticks;
  uint64_t idle_ticks;  // resolver: leave if nothing arrives and no epoch ends this long
  uint64_t owe_ticks;   // epoch kernel: wait this long for an owing entry before deferring
  uint64_t run_ticks;   // epoch kernel: a run still going this long after the kernel's start is deferred (0: no limit)
  // Batched soups (num_soups > 0; batch.h): soup s owns programs
  // [prog_base[s], prog_base[s+1]) and pair slots [slot_base[s],
  // slot_base[s+1]); `k` of a list entry is a slot over all soups.  An
  // epoch launch runs the slots list[0, slots); each slot's gathered and
  // mutated tape is stage[slot] and its programs tgt[2 slot], tgt[2 slot+1].
  // Counters are per soup (counters + 4 s); per-slot ops and steps of every
  // epoch go to pair_ops / pair_steps (0 for a deferred slot, whose counts
  // the resolver adds to its soup's counters).
  uint32_t num_soups;
  const uint64_t *prog_base, *slot_base, *sm_seeds, *num_idx;
  const uint32_t *probs, *list, *tgt;
  const uint8_t *stage;
  uint32_t *pair_work;  // per slot: commands executed (the budget for a deferred slot)
  uint32_t *claim;      // next unclaimed position of list (0 at the launch)
};

hipError_t launch_init(hipStream_t st, uint8_t *soup, uint64_t n,
                       uint64_t seed_value, uint64_t count, int zero_init);
hipError_t launch_epoch(hipStream_t st, int lang, uint8_t *soup,
                        const uint32_t *shuf, uint64_t n,
                        const uint64_t *num_indices, uint64_t seed_value,
                        uint32_t prob, uint32_t *pair_ops, uint32_t *pair_steps,
                        unsigned long long *counters);
// trace (nullable): program 0's tapes before and after each generation of
// each trial, 13 * 5 * 2 * 128 bytes.
// Scratch of the levelled self-rep check (selfrep.hip) for `chunk` programs
// at a time: a 128-byte final tape per chain (13 per program), two chain
// lists of chunk * 13 entries, and their two counters.
struct SrScratch {
  uint8_t *finals;
  uint32_t *list[2];
  uint32_t *counts;
  uint64_t chunk;
};
// Programs per chunk: 2^22 (7 GiB of final tapes).  The last levels are
// bound by the latency of the few longest chains, paid once per chunk.
constexpr uint64_t kSelfrepChunk = 1ull << 22;

// True for the languages the levelled check runs (BffFast's).
bool selfrep_levels_supported(int lang);
// CheckSelfRep of programs [0, count) of the soup into result[0, count),
// chunk by chunk; synchronises `st` between levels.
hipError_t launch_selfrep_levels(hipStream_t st, int lang, const uint8_t *soup, uint64_t n,
                                 uint64_t count, uint64_t seed_value, uint64_t *result,
                                 const SrScratch &w);

hipError_t launch_selfrep(hipStream_t st, int lang, const uint8_t *soup,
                          uint64_t n, uint64_t count, uint64_t seed_value,
                          uint64_t *result, uint8_t *trace = nullptr);
// Records a failure for cubff_last_error() (capi.hip); returns code.
int report_error(int code, const std::string &msg);
hipError_t launch_eval_tapes(hipStream_t st, int lang, uint8_t *tapes,
                             uint64_t count, uint32_t steps, uint32_t *ops,
                             uint32_t *steps_out);
hipError_t launch_byte_hist(hipStream_t st, const uint8_t *soup, uint64_t bytes,
                            unsigned long long *counts);
hipError_t launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v,
                           uint64_t count);
hipError_t launch_set_u64(hipStream_t st, uint64_t *p, uint64_t v);
hipError_t launch_pair_fixed(hipStream_t st, uint32_t *s, uint64_t n,
                             uint64_t epoch);
hipError_t launch_pair_cyclic(hipStream_t st, uint32_t *s, uint64_t n,
                              uint64_t epoch);
// The Fisher-Yates pairings of epochs epoch .. epoch + epochs - 1 into
// s[y * n ..] (y = 0 .. epochs - 1); w's arrays hold `epochs` times their
// one-epoch sizes.
hipError_t launch_pair_shuffle_many(hipStream_t st, const FyScratch &w, uint32_t *s,
                                    uint64_t n, uint64_t params_seed, uint64_t epoch,
                                    uint32_t epochs);
hipError_t launch_pair_shuffle(hipStream_t st, const FyScratch &w, uint32_t *s,
                               uint64_t n, uint64_t params_seed, uint64_t epoch);
hipError_t launch_pair_2d(hipStream_t st, const FyScratch &w,
                          const G2Scratch &g2, const uint64_t *row_off,
                          const uint32_t *nbr, uint32_t *s, uint64_t n,
                          uint64_t params_seed, uint64_t epoch,
                          uint64_t *num_indices);

hipError_t launch_flow_pairof(hipStream_t st, const uint32_t *shuf, uint32_t *pairof,
                              uint64_t n, uint32_t epochs);
hipError_t launch_flow(hipStream_t st, int lang, const FlowArgs &a);
// Step records of any language (trace.hip): kStepRecordBytes per step.
constexpr int kStepRecordBytes = 16 + 2 * kTape2;
hipError_t launch_trace_steps(hipStream_t st, int lang, uint8_t *tape, uint32_t steps,
                              uint8_t *out, uint64_t cap, uint64_t *meta);
hipError_t launch_trace_bff(hipStream_t st, int lang, const uint8_t *tape, uint32_t steps,
                            uint8_t *out, uint64_t cap, uint64_t *meta);
hipError_t launch_defer_start(hipStream_t st, DeferCtl *c);
hipError_t launch_defer_close(hipStream_t st, DeferCtl *c);
hipError_t launch_epoch_defer(hipStream_t st, int lang, const DeferArgs &a,
                              const uint32_t *shuf, const uint64_t *num_indices, uint32_t e_rel);
hipError_t launch_resolve(hipStream_t st, int lang, const DeferArgs &a, unsigned waves);
// Batched soups: the deferring epoch kernel over a.list[0, a.slots) for a
// language BffFast runs (batch_defers()), persistent with workgroups of wg
// lanes, waves_per_simd wavefronts per SIMD; and after the resolver is
// done, the last epoch's per-slot counts of the first `listed` entries.
bool batch_defers(int lang);
hipError_t launch_b_epoch_defer(hipStream_t st, int lang, const DeferArgs &a, uint32_t e_rel,
                                int wg, uint32_t waves_per_simd);
hipError_t launch_b_defer_fixup(hipStream_t st, const DeferArgs &a, uint32_t listed);

}  // namespace cubff
