// Learned scheduler of the batched path: the MLP cost model (mlp.h) predicts
// each staged pair's interaction length; pairs are binned into 8 length
// buckets (quantiles of the epoch's predictions), longest bucket first,
// before the interpreter launch.  Scheduling
// only reorders work: results are bit-identical with and without it.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "batch.h"
#include "mlp.h"

namespace cubff {

constexpr int kBuckets = 8;

struct SchedState {
  bool ready = false;
  Mlp mlp;
  uint32_t *bucket = nullptr;      // [total_slots] bucket of each slot (0 = longest)
  uint32_t *cnt = nullptr;         // [num_lists * 2 * kBuckets] bucket totals per list
  uint32_t *blk = nullptr;         // per-block bucket counts, then output offsets
  uint32_t *hist = nullptr;        // prediction histogram
  uint32_t *thr = nullptr;         // bucket thresholds (prediction bins)
  float *pred = nullptr;           // [total_slots] predicted log2 length
  uint32_t *train_rows = nullptr;  // [kTrainRows]
  uint64_t train_rows_n = 0;
  uint64_t epochs_seen = 0, trainings = 0;
  float last_loss = 0.f;
  // Training after the first runs on its own stream beside the next
  // epochs: the epoch's features and targets are taken on the batch stream,
  // the trainer folds the result into the model's `next` weights, and
  // sched_poll publishes them between epochs once `trained` has fired.
  hipStream_t train_stream = nullptr;
  hipEvent_t taken = nullptr, trained = nullptr;
  bool pending = false;
  float *loss0 = nullptr;  // device: first-step loss of the last training
};

constexpr uint64_t kTrainRows = 8192;

// Identifies a list of slots of one language inside plain/order lists.
struct SchedList {
  int lang;
  uint64_t begin, count;
};

int sched_ensure(SchedState &s, const BatchDev &d, size_t num_lists);
void sched_free(SchedState &s);
// Predicts every slot from its tape in d.stage (halves == nullptr) or, for
// the next epoch, from the two staged halves named in d.halves_next,
// then writes each list's slots to `order` grouped by bucket, longest
// bucket first.
hipError_t sched_order(hipStream_t st, SchedState &s, const BatchDev &d, const uint32_t *halves,
                       const uint32_t *plain, uint32_t *order, const SchedList *lists,
                       size_t num_lists);
// Trains on the last epoch's staged tapes and recorded ops (host-driven,
// outside captured graphs) and publishes the weights before returning.
// Returns the loss of the first step.
hipError_t sched_train(hipStream_t st, SchedState &s, const BatchDev &d, int steps,
                       float lr, float *first_loss);
// The same on the training stream: returns once the work is queued; st only
// waits for the features and targets to be taken (and for a previous
// training still pending, which it publishes first).
hipError_t sched_train_async(hipStream_t st, SchedState &s, const BatchDev &d, int steps,
                             float lr);
// Publishes a pending training's weights on st if it has finished, or,
// with wait, in any case (st then waits for it).
hipError_t sched_poll(hipStream_t st, SchedState &s, bool wait);
// Loss of the first step of the last finished training (reads the device).
float sched_last_loss(SchedState &s);

}  // namespace cubff
