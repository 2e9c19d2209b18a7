// Device view and launchers of the batched-soup path (batch.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubff {

constexpr int kNumBatchCounters = 4;  // ops, num_runs, steps, unused
// Pairing buffers: epoch e's pairing is in shuf[e % kPairBufs]; the pairing
// of epoch e + 2 is drawn while epoch e runs.
constexpr uint32_t kPairBufs = 3;

// Everything a batched launch needs, passed by value to every kernel.
// Program indices are global (soup s owns [prog_base[s], prog_base[s+1]));
// slot indices likewise over slot_base.
struct BatchDev {
  uint32_t num_soups;
  uint64_t total_programs, total_slots;
  const uint64_t *prog_base;   // num_soups + 1
  const uint64_t *slot_base;   // num_soups + 1
  const uint64_t *sm_seed;     // SplitMix64(params seed) per soup
  const uint32_t *prob;        // mutation threshold per soup (of 2^30)
  const uint64_t *num_indices; // evaluated slots per soup
  uint8_t *soup;               // total_programs * 64
  uint32_t *shuf[kPairBufs];   // total_programs each, local indices per soup:
                               // the pairing of epoch e is in shuf[e % 3]
  uint8_t *stage;              // total_slots * 128
  uint32_t *tgt;               // total_slots * 2, global program indices
  uint32_t *inv[kPairBufs];    // inverse pairings: inv[e % 3][program] = its
                               // position in shuf[e % 3] (global indices),
                               // which is also the staged half (2 * slot + h)
                               // that holds it in epoch e
  uint32_t *halves_next;       // nullable, total_slots * 2: written by the
                               // stage of epoch e with the staged halves
                               // holding the programs of epoch e + 1's slots
  uint32_t *pair_ops, *pair_steps;     // total_slots
  uint32_t *pair_work;                 // total_slots: commands actually interpreted
  unsigned long long *counters;        // num_soups * kNumBatchCounters
  uint64_t *epoch;             // device scalar: the epoch being run
  // Fisher-Yates scratch (total_programs entries; fy_cnt has one more).
  uint32_t *fy_j, *fy_cnt, *fy_base, *fy_cursor, *fy_items, *fy_first, *fy_succ;
};

// The epoch is read from d.epoch; batch_end_epoch adds num_runs and
// advances it, so one captured epoch can be replayed.  batch_pair_shuffle
// draws the pairing of epoch *d.epoch + ahead (ahead 0, 1 or 2).
hipError_t batch_pair_shuffle(hipStream_t st, const BatchDev &d, uint32_t ahead);
hipError_t batch_stage(hipStream_t st, const BatchDev &d);
hipError_t batch_copy_u32(hipStream_t st, uint32_t *dst, const uint32_t *src, uint64_t count);
hipError_t batch_end_epoch(hipStream_t st, const BatchDev &d);
hipError_t batch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, uint64_t count);
// Evaluates list[0..count) (staged slot indices) with language `lang`,
// workgroup size wg in {64, 128, 256} and up to per_lane tapes per lane.
hipError_t batch_interp(hipStream_t st, int lang, const BatchDev &d,
                        const uint32_t *list, uint64_t count, int wg,
                        uint32_t per_lane);

}  // namespace cubff
