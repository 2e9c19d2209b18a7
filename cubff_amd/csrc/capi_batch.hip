// C-ABI of the batched-soup path (declared in include/cubff_hip.h).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/cubff_hip.h"
#include "batch.h"
#include "device_common.h"
#include "kernels.h"
#include "languages.h"
#include "sched.h"

using namespace cubff;

struct cubff_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  // Pairings are drawn two epochs ahead on stream2, and after them the
  // learned order one epoch ahead, forked from and joined back into each
  // epoch (also inside the captured graph).
  hipStream_t stream2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // Batches of several languages: every second language's interpreter
  // launch runs on stream_i beside the others (one per language otherwise
  // queue behind each other), forked after the stage and joined before the
  // epoch ends.
  hipStream_t stream_i = nullptr;
  hipEvent_t join_i = nullptr;
  uint64_t drawn = UINT64_MAX;  // the pairings of epochs drawn and drawn + 1 are drawn
  uint64_t last = 0;            // last epoch run
  uint32_t num_soups = 0;
  std::vector<int> langs;
  std::vector<uint64_t> n, prog_base, slot_base;
  BatchDev d{};
  // Device allocations owned by the batch (d's const pointers alias these).
  uint64_t *prog_base_d = nullptr, *slot_base_d = nullptr, *sm_seed_d = nullptr;
  uint64_t *num_indices_d = nullptr;
  uint32_t *prob_d = nullptr;
  // Per language present: the identity slot list and its scheduled order.
  std::vector<SchedList> lists;  // offsets into plain_list / order_list
  uint32_t *plain_list = nullptr;   // total_slots
  uint32_t *claims = nullptr;  // per list: the deferring kernel's next entry
  uint32_t *order_list = nullptr;   // total_slots (scheduled order)
  // Learned order of the next epoch, built on stream2 beside the
  // interpreter from the next pairs' programs as this epoch staged them
  // (allocated on the first learned run).
  uint32_t *order_next = nullptr;   // total_slots
  uint32_t *ahead = nullptr;        // total_slots * 2 staged halves
  uint64_t ordered = UINT64_MAX;    // epoch whose order is in order_list
  SchedState sched{};
  int wg = 64;  // launch shape: measured best for plain and learned order
  uint32_t per_lane = 2;
  bool have_seeds = false;
  hipGraphExec_t graph = nullptr;
  int graph_sched = -1, graph_wg = 0;
  uint32_t graph_per_lane = 0;
  // Deferring epochs (defer.hip) for the batch's first language that
  // batch_defers() names (lists[defer_li]): runs of two or more epochs
  // without graph replay go in segments whose long interactions a resolver
  // on stream_r finishes beside the later epochs.  CUBFF_FLOW_OFF keeps
  // every epoch to the interpreter rounds.
  int pipeline = CUBFF_FLOW_AUTO;
  int defer_li = -1;
  hipStream_t stream_r = nullptr;
  uint32_t *owner = nullptr;  // total_programs
  DeferItem *ditems = nullptr;
  uint32_t *dqueue = nullptr;
  DeferCtl *dctl = nullptr;
  uint64_t defer_cap = 0, dqueue_dirty = 0, defer_items = 0;
  uint32_t defer_gen = 0, defer_budget = 768;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess)                                                    \
      return fail(CUBFF_EHIP, std::string(#expr " failed: ") +              \
                                  hipGetErrorString(_e));                    \
  } while (0)

template <typename T>
int alloc(T **p, uint64_t count) {
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), (count ? count : 1) * sizeof(T));
  if (e != hipSuccess)
    return fail(CUBFF_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  return CUBFF_OK;
}

template <typename T>
void release(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

int check(cubff_batch *b) {
  if (!b) return fail(CUBFF_EINVAL, "null batch");
  HIP_TRY(hipSetDevice(b->device));
  return CUBFF_OK;
}

void drop_graph(cubff_batch *b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_sched = -1;
}

// One epoch on the batch stream: stage, interpret (in order_list's order
// under the learned scheduler), end-of-epoch bookkeeping.  interp_ev
// (nullable) brackets the interpreter launches with events.
// With da (a deferring segment), the deferring language's slots go to the
// deferring epoch kernel as epoch e_rel of the segment.
int enqueue_epoch(cubff_batch *b, int sched, hipEvent_t ev0, hipEvent_t ev1,
                  const DeferArgs *da = nullptr, uint32_t e_rel = 0) {
  const hipStream_t st = b->stream;
  if (sched != CUBFF_SCHED_PLAIN && sched != CUBFF_SCHED_LEARNED)
    return fail(CUBFF_EINVAL, "unknown scheduling mode");
  const bool learned = sched == CUBFF_SCHED_LEARNED;
  HIP_TRY(batch_stage(st, b->d));
  // The pairing of epoch e + 2, beside the rest of this one (it writes only
  // its own pairing buffer and the pairing scratch); forked after the stage,
  // whose memory traffic it would otherwise slow down.  Under the learned
  // scheduler the next epoch's order follows it on stream2: inference on its
  // pairs' tapes as pieces of this epoch's stage (named by the stage in
  // d.halves_next) and bucketing into order_next.  Started behind the
  // pairing, the inference meets the interpreter's tail, when most of its
  // workgroups are done, rather than its start, when they fill every CU
  // (measured: 1.8-2.0 ms behind the pairing vs 3.0 ms from the fork, 0.8 ms
  // alone).
  HIP_TRY(batch_fill_u32(st, b->claims, 0, b->lists.size()));
  HIP_TRY(hipEventRecord(b->fork, st));
  HIP_TRY(hipStreamWaitEvent(b->stream2, b->fork, 0));
  HIP_TRY(batch_pair_shuffle(b->stream2, b->d, 2));
  if (learned)
    HIP_TRY(sched_order(b->stream2, b->sched, b->d, b->ahead, b->plain_list, b->order_next,
                        b->lists.data(), b->lists.size()));
  HIP_TRY(hipEventRecord(b->join, b->stream2));
  const uint32_t *list = learned ? b->order_list : b->plain_list;
  if (ev0) HIP_TRY(hipEventRecord(ev0, st));
  // Languages' slots are disjoint, and so are their soups.
  bool side = false;
  for (size_t li = 0; li < b->lists.size(); li++) {
    const auto &l = b->lists[li];
    const bool deferring = da && (int)li == b->defer_li;
    const hipStream_t s = (!deferring && b->stream_i && (li & 1)) ? b->stream_i : st;
    if (s != st && !side) {
      HIP_TRY(hipStreamWaitEvent(s, b->fork, 0));
      side = true;
    }
    if (deferring) {
      DeferArgs ea = *da;
      ea.list = list + l.begin;
      ea.slots = (uint32_t)l.count;
      ea.claim = b->claims + li;
      HIP_TRY(launch_b_epoch_defer(st, l.lang, ea, e_rel, b->wg, b->per_lane));
    } else {
      HIP_TRY(batch_interp(s, l.lang, b->d, list + l.begin, l.count, b->wg, b->per_lane));
    }
  }
  if (side) {
    HIP_TRY(hipEventRecord(b->join_i, b->stream_i));
    HIP_TRY(hipStreamWaitEvent(st, b->join_i, 0));
  }
  if (ev1) HIP_TRY(hipEventRecord(ev1, st));
  HIP_TRY(hipStreamWaitEvent(st, b->join, 0));  // before the epoch counter moves
  if (learned) HIP_TRY(batch_copy_u32(st, b->order_list, b->order_next, b->d.total_slots));
  HIP_TRY(batch_end_epoch(st, b->d));
  return CUBFF_OK;
}

// The learned order of epoch *d.epoch from its own staged tapes, on the batch
// stream (outside captured graphs): for the first epoch of a run with no
// order built ahead, and after the first training, whose weights the order
// built during that epoch did not have yet.  The epoch's stage is repeated
// by the epoch itself (same inputs, same bytes).
int order_fresh(cubff_batch *b) {
  HIP_TRY(batch_stage(b->stream, b->d));
  HIP_TRY(sched_order(b->stream, b->sched, b->d, nullptr, b->plain_list, b->order_list,
                      b->lists.data(), b->lists.size()));
  return CUBFF_OK;
}

// After an epoch under the learned scheduler: train on its recorded ops on
// the first epoch (synchronously: the initial weights are random) and every
// kTrainInterval epochs after that (beside the next epochs, on the
// scheduler's training stream); publish finished trainings.
constexpr uint64_t kTrainInterval = 16;
constexpr int kTrainSteps = 16;
constexpr float kTrainLr = 2e-3f;

int after_epoch(cubff_batch *b, int sched) {
  if (sched != CUBFF_SCHED_LEARNED) return CUBFF_OK;
  if (b->sched.epochs_seen++ % kTrainInterval == 0) {
    if (b->sched.trainings == 0 && !b->sched.pending) {
      float loss = 0;
      HIP_TRY(sched_train(b->stream, b->sched, b->d, kTrainSteps, kTrainLr, &loss));
      int r;
      if ((r = order_fresh(b))) return r;
    } else {
      HIP_TRY(sched_train_async(b->stream, b->sched, b->d, kTrainSteps, kTrainLr));
    }
  } else {
    HIP_TRY(sched_poll(b->stream, b->sched, false));
  }
  return CUBFF_OK;
}

// End of a run: a training still in flight is waited for and published on
// the batch stream, so the run's work (and any timing of it) includes it.
int finish_run(cubff_batch *b, int sched) {
  if (sched == CUBFF_SCHED_LEARNED) HIP_TRY(sched_poll(b->stream, b->sched, true));
  return CUBFF_OK;
}

// Sets the device epoch counter to the first epoch of a run of `count` and
// draws the pairings of that epoch and the next unless the previous run drew
// them already; under the learned scheduler, orders that epoch unless the
// previous run did.
int start_run(cubff_batch *b, uint64_t epoch, uint64_t count, int sched) {
  HIP_TRY(hipMemcpyAsync(b->d.epoch, &epoch, 8, hipMemcpyHostToDevice, b->stream));
  if (count && b->drawn != epoch) {
    HIP_TRY(batch_pair_shuffle(b->stream, b->d, 0));
    HIP_TRY(batch_pair_shuffle(b->stream, b->d, 1));
  }
  const bool learned = sched == CUBFF_SCHED_LEARNED;
  int r;
  if (count && learned && b->ordered != epoch && (r = order_fresh(b))) return r;
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (count) {
    b->drawn = epoch + count;
    b->last = epoch + count - 1;
    b->ordered = learned ? epoch + count : UINT64_MAX;
  }
  return CUBFF_OK;
}

// Scheduler state and the buffers of the order built ahead.
int ensure_learned(cubff_batch *b) {
  if (sched_ensure(b->sched, b->d, b->lists.size()))
    return fail(CUBFF_ENOMEM, "scheduler allocation failed");
  int r;
  if (!b->order_next && (r = alloc(&b->order_next, b->d.total_slots))) return r;
  if (!b->ahead) {
    if ((r = alloc(&b->ahead, b->d.total_slots * 2))) return r;
    b->d.halves_next = b->ahead;
    drop_graph(b);  // captured with the old device view
  }
  return CUBFF_OK;
}

constexpr unsigned kBatchResolveWaves = 512;  // resolver wavefronts
constexpr double kBatchOweWaitS = 200e-6;       // as for one soup
constexpr double kBatchSpinS = 10.0;            // longest wait before the run is abandoned
constexpr double kBatchIdleS = 0.1;             // resolver leaves with nothing deferred
// List entries of one deferring segment: every slot of the deferring
// language in every epoch in the worst case (64-byte entry, 4-byte queue
// slot), within 2 GiB; at most 64 epochs.
constexpr uint64_t kBatchDeferBytes = 1ull << 31;

// Deferring segments: for the plain order by default, for both orders with
// CUBFF_FLOW_DEFER.  The learned order pays in the interpreter rounds,
// whose wavefronts it fills with pairs of like length; the deferring
// kernel's lanes take new pairs as they finish, which leaves the order
// little to gain while its inference still costs.
//
// By default a batch of several languages keeps to the rounds: its epoch
// lasts as long as its slowest language's launch (the Forth family's, in
// mixed bff/forth batches), which deferring another language does not
// shorten, while the persistent kernel and the resolver take CUs from it
// (32 mixed soups: 7.9-8.3 ms per epoch deferring vs 7.5-7.7 in rounds).
bool defer_active(const cubff_batch *b, int use_graph, int sched) {
  if (use_graph || b->defer_li < 0 || b->pipeline == CUBFF_FLOW_OFF) return false;
  if (b->pipeline == CUBFF_FLOW_DEFER) return true;
  return sched != CUBFF_SCHED_LEARNED && b->lists.size() == 1;
}

uint64_t defer_segment_epochs_b(const cubff_batch *b) {
  const uint64_t slots = std::max<uint64_t>(1, b->lists[b->defer_li].count);
  return std::max<uint64_t>(2, std::min<uint64_t>(64, kBatchDeferBytes / (68 * slots)));
}

// The resolver's stream: a CU-masked stream (every CU enabled), which the
// runtime gives a hardware queue of its own rather than one shared with
// another stream; a persistent resolver sharing the epoch stream's queue
// would hold back the epoch kernels it waits for (the resolver then leaves
// after its idle limit and the epochs run without deferral).
hipError_t create_resolver_stream(int device, hipStream_t *out) {
  int cus = 0;
  hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess || cus <= 0) return hipStreamCreateWithFlags(out, hipStreamNonBlocking);
  std::vector<uint32_t> mask((cus + 31) / 32, 0xffffffffu);
  if (cus % 32) mask.back() = (1u << (cus % 32)) - 1;
  e = hipExtStreamCreateWithCUMask(out, (uint32_t)mask.size(), mask.data());
  if (e != hipSuccess) {
    (void)hipGetLastError();
    e = hipStreamCreateWithFlags(out, hipStreamNonBlocking);
  }
  return e;
}

int ensure_defer_b(cubff_batch *b, uint64_t epochs) {
  const uint64_t need = epochs * b->lists[b->defer_li].count;
  int r;
  if (!b->stream_r) HIP_TRY(create_resolver_stream(b->device, &b->stream_r));
  if (!b->owner && (r = alloc(&b->owner, b->d.total_programs))) return r;
  if (!b->dctl && (r = alloc(&b->dctl, 1))) return r;
  if (b->defer_cap < need) {
    release(b->ditems);
    release(b->dqueue);
    b->defer_cap = 0;
    if ((r = alloc(&b->ditems, need)) || (r = alloc(&b->dqueue, need))) return r;
    // Generation words start at 0 (runs use generations >= 1); queue slots
    // start empty.
    HIP_TRY(hipMemsetAsync(b->ditems, 0, need * sizeof(DeferItem), b->stream));
    HIP_TRY(batch_fill_u32(b->stream, b->dqueue, kNone, need));
    b->dqueue_dirty = 0;
    b->defer_cap = need;
  }
  if (b->dqueue_dirty) HIP_TRY(batch_fill_u32(b->stream, b->dqueue, kNone, b->dqueue_dirty));
  b->dqueue_dirty = 0;
  return CUBFF_OK;
}

// Epochs [first, first + epochs) of the batch as one deferring segment.
// iev (nullable, 2 * epochs events) brackets each epoch's interpreter
// launches.
int run_defer_segment_b(cubff_batch *b, int sched, uint64_t first, uint32_t epochs,
                        hipEvent_t *iev) {
  int r;
  if ((r = ensure_defer_b(b, epochs))) return r;
  const SchedList &l = b->lists[b->defer_li];
  int khz = 0;
  HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
  const BatchDev &d = b->d;
  DeferArgs a{};
  a.soup = d.soup;
  a.owner = b->owner;
  a.items = b->ditems;
  a.queue = b->dqueue;
  a.ctl = b->dctl;
  a.slots = (uint32_t)l.count;
  a.epochs = epochs;
  a.capacity = (uint32_t)(epochs * l.count);
  a.budget = b->defer_budget;
  a.max_backlog = 2 * kBatchResolveWaves;
  a.gen = ++b->defer_gen;
  if (a.gen == 0) a.gen = b->defer_gen = 1;  // 0 marks never-written entries
  a.first_epoch = first;
  a.pair_ops = d.pair_ops;
  a.pair_steps = d.pair_steps;
  a.pair_work = d.pair_work;
  a.counters = d.counters;
  a.spin_ticks = (uint64_t)(kBatchSpinS * khz * 1000.0);
  a.idle_ticks = (uint64_t)(kBatchIdleS * khz * 1000.0);
  a.owe_ticks = (uint64_t)(kBatchOweWaitS * khz * 1000.0);
  a.num_soups = d.num_soups;
  a.prog_base = d.prog_base;
  a.slot_base = d.slot_base;
  a.sm_seeds = d.sm_seed;
  a.num_idx = d.num_indices;
  a.probs = d.prob;
  a.tgt = d.tgt;
  a.stage = d.stage;
  const hipStream_t st = b->stream;
  hipEvent_t started = nullptr, resolved = nullptr;
  auto cleanup = [&]() {
    if (started) (void)hipEventDestroy(started);
    if (resolved) (void)hipEventDestroy(resolved);
  };
  hipError_t e = hipEventCreateWithFlags(&started, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&resolved, hipEventDisableTiming);
  if (e == hipSuccess) e = batch_fill_u32(st, b->owner, kNone, d.total_programs);
  if (e == hipSuccess) e = launch_defer_start(st, b->dctl);
  b->dqueue_dirty = b->defer_cap;  // until this segment ends well
  if (e == hipSuccess) e = hipEventRecord(started, st);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->stream_r, started, 0);
  if (e == hipSuccess) e = launch_resolve(b->stream_r, l.lang, a, kBatchResolveWaves);
  if (e != hipSuccess) {
    cleanup();
    return fail(CUBFF_EHIP, std::string("deferred run setup failed: ") + hipGetErrorString(e));
  }
  for (uint32_t j = 0; r == CUBFF_OK && j < epochs; j++) {
    r = enqueue_epoch(b, sched, iev ? iev[2 * j] : nullptr, iev ? iev[2 * j + 1] : nullptr, &a,
                      j);
    if (r == CUBFF_OK) r = after_epoch(b, sched);
  }
  // Close the list whatever happened, so the resolver leaves.
  e = launch_defer_close(st, b->dctl);
  if (e == hipSuccess) e = hipEventRecord(resolved, b->stream_r);
  if (e == hipSuccess) e = hipStreamWaitEvent(st, resolved, 0);
  DeferCtl ctl{};
  if (e == hipSuccess) e = hipMemcpyAsync(&ctl, b->dctl, sizeof(ctl), hipMemcpyDeviceToHost, st);
  hipError_t e2 = hipStreamSynchronize(b->stream_r);
  if (e == hipSuccess) e = e2;
  e2 = hipStreamSynchronize(st);
  if (e == hipSuccess) e = e2;
  cleanup();
  if (r) return r;
  if (e != hipSuccess)
    return fail(CUBFF_EHIP, std::string("deferred run failed: ") + hipGetErrorString(e));
  const uint32_t listed = ctl.final_tail >= kDeferGone ? 0u : ctl.final_tail;
  b->defer_items += listed;
  b->dqueue_dirty = std::min<uint64_t>(std::max(listed, ctl.qtail), b->defer_cap);
  if (ctl.resolved != listed)
    return fail(CUBFF_EHIP, "deferred-interaction resolver stopped: a wait exceeded " +
                                std::to_string(kBatchSpinS) + " s; soup state is undefined");
  // The last epoch's per-slot counts of the interactions the resolver ran.
  HIP_TRY(launch_b_defer_fixup(st, a, listed));
  return CUBFF_OK;
}

// `count` epochs of a run without graph replay: deferring segments where
// they apply, else one enqueued epoch after another.  iev as above.
int run_epochs_plain(cubff_batch *b, int sched, uint64_t first, uint64_t count,
                     hipEvent_t *iev) {
  int r;
  uint64_t i = 0;
  // The learned schedule's first training runs synchronously after its
  // epoch; that epoch goes first on its own (a resolver beside it would see
  // the epoch stream stall and leave).
  if (defer_active(b, 0, sched) && count >= 3 && sched == CUBFF_SCHED_LEARNED &&
      b->sched.trainings == 0) {
    if ((r = enqueue_epoch(b, sched, iev ? iev[0] : nullptr, iev ? iev[1] : nullptr))) return r;
    if ((r = after_epoch(b, sched))) return r;
    i = 1;
  }
  if (defer_active(b, 0, sched) && count - i >= 2) {
    const uint64_t seg = defer_segment_epochs_b(b);
    while (count - i >= 2) {
      const uint64_t ep = std::min(seg, count - i);
      if ((r = run_defer_segment_b(b, sched, first + i, (uint32_t)ep, iev ? iev + 2 * i : nullptr)))
        return r;
      i += ep;
    }
  }
  for (; i < count; i++) {
    if ((r = enqueue_epoch(b, sched, iev ? iev[2 * i] : nullptr, iev ? iev[2 * i + 1] : nullptr)))
      return r;
    if ((r = after_epoch(b, sched))) return r;
  }
  return CUBFF_OK;
}

// The captured epoch of this schedule and launch shape is ready to replay.
bool graph_ready_of(const cubff_batch *b, int sched) {
  return b->graph && b->graph_sched == sched && b->graph_wg == b->wg &&
         b->graph_per_lane == b->per_lane;
}

// One epoch captured in a hipGraph (no hipMemsetAsync inside the capture;
// the epoch number lives on the device, so every replay is the next epoch).
int capture_epoch_graph(cubff_batch *b, int sched) {
  drop_graph(b);
  hipGraph_t g;
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int r = enqueue_epoch(b, sched, nullptr, nullptr);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (r) return r;
  if (e != hipSuccess) return fail(CUBFF_EHIP, "graph capture failed");
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(CUBFF_EHIP, "graph instantiate failed");
  b->graph_sched = sched;
  b->graph_wg = b->wg;
  b->graph_per_lane = b->per_lane;
  return CUBFF_OK;
}

// `count` epochs as replays of the captured epoch.
int run_epochs_graph(cubff_batch *b, int sched, uint64_t count) {
  int r;
  if (!graph_ready_of(b, sched) && (r = capture_epoch_graph(b, sched))) return r;
  for (uint64_t i = 0; i < count; i++) {
    HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    if ((r = after_epoch(b, sched))) return r;
  }
  return CUBFF_OK;
}

}  // namespace

extern "C" {

const char *cubff_batch_last_error(void) { return g_err.c_str(); }

int cubff_batch_create(cubff_batch **out, uint32_t num_soups, const int *langs,
                       const uint64_t *num_programs, int device) {
  if (!out) return fail(CUBFF_EINVAL, "null output pointer");
  *out = nullptr;
  if (num_soups == 0 || !langs || !num_programs)
    return fail(CUBFF_EINVAL, "need at least one soup");
  uint64_t total = 0;
  for (uint32_t s = 0; s < num_soups; s++) {
    if (langs[s] < 0 || langs[s] >= kNumLangs) return fail(CUBFF_EINVAL, "unknown language id");
    if (num_programs[s] == 0 || num_programs[s] % 2)
      return fail(CUBFF_EINVAL, "num_programs must be even and > 0");
    total += num_programs[s];
  }
  if (total >= (1ULL << 32)) return fail(CUBFF_EINVAL, "total programs must be < 2^32");
  HIP_TRY(hipSetDevice(device));
  cubff_batch *b = new cubff_batch();
  b->device = device;
  b->num_soups = num_soups;
  b->langs.assign(langs, langs + num_soups);
  b->n.assign(num_programs, num_programs + num_soups);
  b->prog_base.resize(num_soups + 1);
  b->slot_base.resize(num_soups + 1);
  b->prog_base[0] = b->slot_base[0] = 0;
  for (uint32_t s = 0; s < num_soups; s++) {
    b->prog_base[s + 1] = b->prog_base[s] + b->n[s];
    b->slot_base[s + 1] = b->slot_base[s] + b->n[s] / 2;
  }
  const uint64_t P = b->prog_base[num_soups], S = b->slot_base[num_soups];
  BatchDev &d = b->d;
  d.num_soups = num_soups;
  d.total_programs = P;
  d.total_slots = S;
  int r;
  uint32_t *cnt = nullptr;
  unsigned long long *ctr = nullptr;
  if ((r = alloc(&b->prog_base_d, num_soups + 1)) || (r = alloc(&b->slot_base_d, num_soups + 1)) ||
      (r = alloc(&b->sm_seed_d, num_soups)) || (r = alloc(&b->num_indices_d, num_soups)) ||
      (r = alloc(&b->prob_d, num_soups)) || (r = alloc(&d.soup, P * kTape)) ||
      (r = alloc(&d.shuf[0], P)) || (r = alloc(&d.shuf[1], P)) || (r = alloc(&d.shuf[2], P)) ||
      (r = alloc(&d.stage, S * kTape2)) ||
      (r = alloc(&d.tgt, 2 * S)) || (r = alloc(&d.inv[0], P)) || (r = alloc(&d.inv[1], P)) ||
      (r = alloc(&d.inv[2], P)) || (r = alloc(&d.pair_ops, S)) ||
      (r = alloc(&d.pair_steps, S)) || (r = alloc(&d.pair_work, S)) ||
      (r = alloc(&ctr, (uint64_t)num_soups * kNumBatchCounters)) ||
      (r = alloc(&d.epoch, 1)) || (r = alloc(&d.fy_j, P)) || (r = alloc(&cnt, P + 1)) ||
      (r = alloc(&d.fy_base, P)) || (r = alloc(&d.fy_cursor, P)) ||
      (r = alloc(&d.fy_items, P)) || (r = alloc(&d.fy_first, P)) ||
      (r = alloc(&d.fy_succ, P)) || (r = alloc(&b->plain_list, S)) ||
      (r = alloc(&b->claims, kNumLangs)) ||
      (r = alloc(&b->order_list, S))) {
    d.fy_cnt = cnt;
    d.counters = ctr;
    std::string keep = g_err;
    cubff_batch_destroy(b);
    g_err = keep;
    return r;
  }
  d.fy_cnt = cnt;
  d.counters = ctr;
  d.prog_base = b->prog_base_d;
  d.slot_base = b->slot_base_d;
  d.sm_seed = b->sm_seed_d;
  d.prob = b->prob_d;
  d.num_indices = b->num_indices_d;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
  if (e != hipSuccess) {
    cubff_batch_destroy(b);
    return fail(CUBFF_EHIP, "hipStreamCreate failed");
  }
  // Slot lists grouped by language, in soup order within each language.
  std::vector<uint32_t> plain;
  plain.reserve(S);
  for (int lang = 0; lang < kNumLangs; lang++) {
    const uint64_t begin = plain.size();
    for (uint32_t s = 0; s < num_soups; s++)
      if (b->langs[s] == lang)
        for (uint64_t g = b->slot_base[s]; g < b->slot_base[s + 1]; g++) plain.push_back((uint32_t)g);
    if (plain.size() > begin) b->lists.push_back({lang, begin, plain.size() - begin});
  }
  for (size_t li = 0; li < b->lists.size() && b->defer_li < 0; li++)
    if (batch_defers(b->lists[li].lang)) b->defer_li = (int)li;
  if (b->lists.size() > 1) {
    e = hipStreamCreateWithFlags(&b->stream_i, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join_i, hipEventDisableTiming);
    if (e != hipSuccess) {
      cubff_batch_destroy(b);
      return fail(CUBFF_EHIP, "hipStreamCreate failed");
    }
  }
  std::vector<uint64_t> ni(num_soups);
  for (uint32_t s = 0; s < num_soups; s++) ni[s] = b->n[s] / 2;
  const uint64_t zero = 0;
  e = hipMemcpy(b->prog_base_d, b->prog_base.data(), (num_soups + 1) * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->slot_base_d, b->slot_base.data(), (num_soups + 1) * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->num_indices_d, ni.data(), num_soups * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->plain_list, plain.data(), S * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d.epoch, &zero, 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = batch_fill_u32(b->stream, reinterpret_cast<uint32_t *>(d.counters), 0,
                       2ull * num_soups * kNumBatchCounters);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) {
    cubff_batch_destroy(b);
    return fail(CUBFF_EHIP, std::string("batch init failed: ") + hipGetErrorString(e));
  }
  *out = b;
  return CUBFF_OK;
}

int cubff_batch_destroy(cubff_batch *b) {
  if (!b) return CUBFF_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  sched_free(b->sched);
  release(b->prog_base_d);
  release(b->slot_base_d);
  release(b->sm_seed_d);
  release(b->num_indices_d);
  release(b->prob_d);
  release(b->d.soup);
  release(b->d.shuf[0]);
  release(b->d.shuf[1]);
  release(b->d.shuf[2]);
  release(b->d.stage);
  release(b->d.tgt);
  release(b->d.inv[0]);
  release(b->d.inv[1]);
  release(b->d.inv[2]);
  release(b->order_next);
  release(b->ahead);
  release(b->d.pair_ops);
  release(b->d.pair_steps);
  release(b->d.pair_work);
  release(b->d.counters);
  release(b->d.epoch);
  release(b->d.fy_j);
  release(b->d.fy_cnt);
  release(b->d.fy_base);
  release(b->d.fy_cursor);
  release(b->d.fy_items);
  release(b->d.fy_first);
  release(b->d.fy_succ);
  release(b->plain_list);
  release(b->claims);
  if (b->stream_r) (void)hipStreamSynchronize(b->stream_r);
  release(b->owner);
  release(b->ditems);
  release(b->dqueue);
  release(b->dctl);
  if (b->stream_r) (void)hipStreamDestroy(b->stream_r);
  release(b->order_list);
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  if (b->stream_i) (void)hipStreamSynchronize(b->stream_i);
  if (b->join_i) (void)hipEventDestroy(b->join_i);
  if (b->stream_i) (void)hipStreamDestroy(b->stream_i);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CUBFF_OK;
}

int cubff_batch_init(cubff_batch *b, const uint64_t *params_seeds,
                     const uint32_t *mutation_probs, int zero_init) {
  int r;
  if ((r = check(b))) return r;
  if (!params_seeds || !mutation_probs) return fail(CUBFF_EINVAL, "null seeds / probabilities");
  std::vector<uint64_t> sm(b->num_soups);
  for (uint32_t s = 0; s < b->num_soups; s++) sm[s] = splitmix64(params_seeds[s]);
  HIP_TRY(hipMemcpyAsync(b->sm_seed_d, sm.data(), b->num_soups * 8, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->prob_d, mutation_probs, b->num_soups * 4, hipMemcpyHostToDevice, b->stream));
  for (uint32_t s = 0; s < b->num_soups; s++)
    HIP_TRY(launch_init(b->stream, b->d.soup + b->prog_base[s] * kTape, b->n[s],
                        epoch_seed_pre(sm[s], 0), b->n[s], zero_init));
  HIP_TRY(batch_fill_u32(b->stream, reinterpret_cast<uint32_t *>(b->d.counters), 0,
                         2ull * b->num_soups * kNumBatchCounters));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->have_seeds = true;
  b->drawn = UINT64_MAX;  // pairings depend on the seeds
  b->ordered = UINT64_MAX;
  return CUBFF_OK;
}

int cubff_batch_set_launch(cubff_batch *b, int wg, uint32_t tapes_per_lane) {
  int r;
  if ((r = check(b))) return r;
  if (wg != 64 && wg != 128 && wg != 256) return fail(CUBFF_EINVAL, "workgroup size must be 64, 128 or 256");
  if (tapes_per_lane < 1 || tapes_per_lane > 64) return fail(CUBFF_EINVAL, "tapes_per_lane must be in [1, 64]");
  b->wg = wg;
  b->per_lane = tapes_per_lane;
  return CUBFF_OK;
}

int cubff_batch_get_launch(cubff_batch *b, int *wg, uint32_t *tapes_per_lane) {
  int r;
  if ((r = check(b))) return r;
  if (wg) *wg = b->wg;
  if (tapes_per_lane) *tapes_per_lane = b->per_lane;
  return CUBFF_OK;
}

int cubff_batch_run_epochs(cubff_batch *b, uint64_t first_epoch, uint64_t count,
                           int sched, int use_graph) {
  int r;
  if ((r = check(b))) return r;
  if (!b->have_seeds) return fail(CUBFF_ESTATE, "cubff_batch_init first");
  if (sched == CUBFF_SCHED_LEARNED && (r = ensure_learned(b))) return r;
  if ((r = start_run(b, first_epoch, count, sched))) return r;
  if ((r = use_graph ? run_epochs_graph(b, sched, count)
                     : run_epochs_plain(b, sched, first_epoch, count, nullptr)))
    return r;
  return finish_run(b, sched);
}

int cubff_batch_run_epochs_timed(cubff_batch *b, uint64_t first_epoch, uint64_t count,
                                 int sched, int use_graph, float *total_ms, float *interp_ms) {
  int r;
  if ((r = check(b))) return r;
  if (!b->have_seeds) return fail(CUBFF_ESTATE, "cubff_batch_init first");
  if (sched == CUBFF_SCHED_LEARNED && (r = ensure_learned(b))) return r;
  if ((r = start_run(b, first_epoch, count, sched))) return r;
  // Graph replay: the captured epoch is one graph launch, so only the whole
  // run is timed (interpreter time reported as -1).
  std::vector<hipEvent_t> ev(use_graph ? 2 : 2 * count + 2);
  for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
  if (use_graph && !graph_ready_of(b, sched)) {
    // Capture outside the timed region.
    if ((r = capture_epoch_graph(b, sched))) return r;
  }
  HIP_TRY(hipEventRecord(ev[0], b->stream));
  if ((r = use_graph ? run_epochs_graph(b, sched, count)
                     : run_epochs_plain(b, sched, first_epoch, count, ev.data() + 2)))
    return r;
  if ((r = finish_run(b, sched))) return r;
  HIP_TRY(hipEventRecord(ev[1], b->stream));
  HIP_TRY(hipEventSynchronize(ev[1]));
  float tot = 0, it = use_graph ? -1.f : 0.f;
  HIP_TRY(hipEventElapsedTime(&tot, ev[0], ev[1]));
  for (uint64_t i = 0; !use_graph && i < count; i++) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev[2 + 2 * i], ev[3 + 2 * i]));
    it += ms;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (total_ms) *total_ms = tot;
  if (interp_ms) *interp_ms = it;
  return CUBFF_OK;
}

int cubff_batch_autotune(cubff_batch *b, uint64_t first_epoch, uint64_t epochs_per_config,
                         int sched, int *best_wg, uint32_t *best_tapes_per_lane,
                         float *ms_table) {
  int r;
  if ((r = check(b))) return r;
  static const int kWg[3] = {64, 128, 256};
  constexpr int kShapes = 12;
  // Rounds of one epoch per shape, so that every shape sees the soups at
  // about the same stage of their evolution; times summed per shape.
  float sum[kShapes] = {};
  uint64_t e = first_epoch;
  // Runs of deferring segments take two epochs at least: each shape is
  // timed on two-epoch runs then (ms per epoch reported either way).
  const uint64_t per = defer_active(b, 0, sched) ? 2 : 1;
  for (uint64_t round = 0; round < epochs_per_config; round++) {
    for (int sh = 0; sh < kShapes; sh++) {
      b->wg = kWg[sh / 4];
      b->per_lane = (uint32_t)(sh % 4) + 1;
      float tot = 0, it = 0;
      if ((r = cubff_batch_run_epochs_timed(b, e, per, sched, 0, &tot, &it))) return r;
      e += per;
      sum[sh] += tot / per;
    }
  }
  int best = 0;
  for (int sh = 1; sh < kShapes; sh++)
    if (sum[sh] < sum[best]) best = sh;
  b->wg = kWg[best / 4];
  b->per_lane = (uint32_t)(best % 4) + 1;
  if (ms_table)
    for (int sh = 0; sh < kShapes; sh++)
      ms_table[sh] = epochs_per_config ? sum[sh] / epochs_per_config : 0.f;
  if (best_wg) *best_wg = b->wg;
  if (best_tapes_per_lane) *best_tapes_per_lane = b->per_lane;
  return CUBFF_OK;
}

int cubff_batch_set_pipeline(cubff_batch *b, int mode) {
  int r;
  if ((r = check(b))) return r;
  if (mode != CUBFF_FLOW_OFF && mode != CUBFF_FLOW_AUTO && mode != CUBFF_FLOW_DEFER)
    return fail(CUBFF_EINVAL, "batch pipeline must be CUBFF_FLOW_OFF, _AUTO or _DEFER");
  b->pipeline = mode;
  return CUBFF_OK;
}

int cubff_batch_next_epoch(cubff_batch *b, uint64_t *next) {
  int r;
  if ((r = check(b))) return r;
  if (next) *next = b->drawn == UINT64_MAX ? 0 : b->last + 1;
  return CUBFF_OK;
}

int cubff_batch_set_defer_budget(cubff_batch *b, uint32_t budget) {
  int r;
  if ((r = check(b))) return r;
  if (budget == 0) return fail(CUBFF_EINVAL, "defer budget must be >= 1");
  b->defer_budget = budget;
  return CUBFF_OK;
}

int cubff_batch_defer_stats(cubff_batch *b, uint64_t *deferred) {
  int r;
  if ((r = check(b))) return r;
  if (deferred) *deferred = b->defer_items;
  return CUBFF_OK;
}

int cubff_batch_sched_stats(cubff_batch *b, uint64_t *trainings, float *last_loss,
                            uint32_t *bucket_counts) {
  int r;
  if ((r = check(b))) return r;
  if (trainings) *trainings = b->sched.trainings;
  if (last_loss) *last_loss = sched_last_loss(b->sched);
  if (bucket_counts) {
    for (int i = 0; i < kBuckets; i++) bucket_counts[i] = 0;
    if (b->sched.ready) {
      std::vector<uint32_t> c(b->lists.size() * 2 * kBuckets);
      HIP_TRY(hipMemcpyAsync(c.data(), b->sched.cnt, c.size() * 4, hipMemcpyDeviceToHost,
                             b->stream));
      HIP_TRY(hipStreamSynchronize(b->stream));
      for (size_t li = 0; li < b->lists.size(); li++)
        for (int i = 0; i < kBuckets; i++) bucket_counts[i] += c[li * 2 * kBuckets + i];
    }
  }
  return CUBFF_OK;
}

int cubff_batch_sync(cubff_batch *b) {
  int r;
  if ((r = check(b))) return r;
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CUBFF_OK;
}

int cubff_batch_read_soup(cubff_batch *b, uint32_t soup, uint8_t *dst) {
  int r;
  if ((r = check(b))) return r;
  if (soup >= b->num_soups || !dst) return fail(CUBFF_EINVAL, "bad soup index or null destination");
  HIP_TRY(hipMemcpyAsync(dst, b->d.soup + b->prog_base[soup] * kTape, b->n[soup] * kTape,
                         hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CUBFF_OK;
}

int cubff_batch_write_soup(cubff_batch *b, uint32_t soup, const uint8_t *src) {
  int r;
  if ((r = check(b))) return r;
  if (soup >= b->num_soups || !src) return fail(CUBFF_EINVAL, "bad soup index or null source");
  HIP_TRY(hipMemcpyAsync(b->d.soup + b->prog_base[soup] * kTape, src, b->n[soup] * kTape,
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->ordered = UINT64_MAX;  // an order built ahead saw other programs
  return CUBFF_OK;
}

int cubff_batch_read_counters(cubff_batch *b, uint32_t soup, uint64_t *ops,
                              uint64_t *num_runs, uint64_t *steps) {
  int r;
  if ((r = check(b))) return r;
  if (soup >= b->num_soups) return fail(CUBFF_EINVAL, "bad soup index");
  unsigned long long h[kNumBatchCounters];
  HIP_TRY(hipMemcpyAsync(h, b->d.counters + (uint64_t)soup * kNumBatchCounters, sizeof(h),
                         hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (ops) *ops = h[0];
  if (num_runs) *num_runs = h[1];
  if (steps) *steps = h[2];
  return CUBFF_OK;
}

int cubff_batch_reset_counters(cubff_batch *b) {
  int r;
  if ((r = check(b))) return r;
  HIP_TRY(batch_fill_u32(b->stream, reinterpret_cast<uint32_t *>(b->d.counters), 0,
                         2ull * b->num_soups * kNumBatchCounters));
  return CUBFF_OK;
}

int cubff_batch_read_pair_stats(cubff_batch *b, uint32_t soup, uint32_t *ops,
                                uint32_t *steps, uint32_t *work) {
  int r;
  if ((r = check(b))) return r;
  if (soup >= b->num_soups) return fail(CUBFF_EINVAL, "bad soup index");
  const uint64_t s0 = b->slot_base[soup], ns = b->slot_base[soup + 1] - s0;
  if (ops) HIP_TRY(hipMemcpyAsync(ops, b->d.pair_ops + s0, ns * 4, hipMemcpyDeviceToHost, b->stream));
  if (steps) HIP_TRY(hipMemcpyAsync(steps, b->d.pair_steps + s0, ns * 4, hipMemcpyDeviceToHost, b->stream));
  if (work) HIP_TRY(hipMemcpyAsync(work, b->d.pair_work + s0, ns * 4, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CUBFF_OK;
}

int cubff_batch_read_shuffle(cubff_batch *b, uint32_t soup, uint32_t *s) {
  int r;
  if ((r = check(b))) return r;
  if (soup >= b->num_soups || !s) return fail(CUBFF_EINVAL, "bad soup index or null destination");
  HIP_TRY(hipMemcpyAsync(s, b->d.shuf[b->last % kPairBufs] + b->prog_base[soup], b->n[soup] * 4,
                         hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CUBFF_OK;
}

}  // extern "C"
