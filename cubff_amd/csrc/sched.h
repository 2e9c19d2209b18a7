// Learned scheduler of the batched path: an MLP cost model predicts each
// staged pair's interaction length; pairs are binned into 8 length buckets,
// longest first, before the interpreter launch.  Scheduling only reorders
// work: results are bit-identical with and without it.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "batch.h"

namespace cubff {

struct SchedState {
  bool ready = false;
  uint64_t capacity = 0;  // slots the buffers hold
  // MLP parameters (fp32 master copies + bf16 copies used by the GEMMs).
  float *params = nullptr;        // all fp32 parameters, one allocation
  float *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  uint16_t *params_bf16 = nullptr;
  // Per-slot activations / outputs.
  uint16_t *feat = nullptr;       // [capacity, 256] bf16 byte histograms
  uint16_t *act1 = nullptr, *act2 = nullptr;  // [capacity, 256] bf16
  float *pred = nullptr;          // [capacity] predicted log2(steps + 1)
  uint32_t *bucket_cnt = nullptr; // per list x 8 buckets (+ cursors)
  uint64_t trained_epochs = 0;
};

// Identifies a list of slots of one language inside plain/order lists.
struct SchedList {
  int lang;
  uint64_t begin, count;
};

int sched_ensure(SchedState &s, const BatchDev &d, size_t num_lists);
void sched_free(SchedState &s);
hipError_t sched_order_list(hipStream_t st, SchedState &s, const BatchDev &d,
                            const uint32_t *plain, uint32_t *order, uint64_t count,
                            size_t list_index);
hipError_t sched_record(hipStream_t st, SchedState &s, const BatchDev &d);

template <class LL>
hipError_t sched_order(hipStream_t st, SchedState &s, const BatchDev &d,
                       const uint32_t *plain, uint32_t *order, const LL *lists,
                       size_t num_lists) {
  hipError_t e = hipSuccess;
  for (size_t i = 0; i < num_lists && e == hipSuccess; i++)
    e = sched_order_list(st, s, d, plain + lists[i].begin, order + lists[i].begin,
                         lists[i].count, i);
  return e;
}

}  // namespace cubff
