// C-ABI of the cubff HIP library (declared in include/cubff_hip.h).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/cubff_hip.h"
#include "device_common.h"
#include "kernels.h"
#include "languages.h"

using namespace cubff;

// Command budget of an interaction in the deferring epoch kernel: runs
// past it are resumed by the resolver (2^17 bff_noheads: 384-768 within 1%
// of each other, 512 best).
constexpr uint32_t kDeferBudget = 512;

struct cubff_ctx {
  int lang = 0;
  int device = 0;
  uint64_t n = 0;
  hipStream_t stream = nullptr;
  uint8_t *soup = nullptr;
  uint32_t *shuf = nullptr;
  uint64_t *num_indices = nullptr;         // device scalar
  unsigned long long *counters = nullptr;  // device [ops, num_runs, steps, -]
  uint32_t *pair_ops = nullptr;
  uint32_t *pair_steps = nullptr;
  uint64_t *selfrep = nullptr;
  SrScratch sr{};  // levelled self-rep check (allocated on first use)
  FyScratch fy{};
  G2Scratch g2{};
  uint64_t *row_off = nullptr;
  uint32_t *nbr = nullptr;
  // Multi-epoch dataflow path (flow.hip): CUBFF_FLOW_* mode; scratch for
  // flow_cap epochs; a single wait may last flow_wait_s seconds.
  int pipeline = CUBFF_FLOW_AUTO;
  double flow_wait_s = 10.0;
  uint64_t flow_cap = 0;
  uint32_t *flow_shuf = nullptr, *flow_pairof = nullptr, *flow_cnt = nullptr,
           *flow_queue = nullptr;
  FlowCtl *flow_ctl = nullptr;
  uint64_t flow_claims = 0, flow_items = 0;  // since creation
  // Deferred-interaction path (defer.hip): second stream, program owners,
  // the list (defer_cap entries), run generation, command budget.
  hipStream_t stream2 = nullptr;
  hipStream_t stream3 = nullptr;  // pairings drawn ahead
  // Pairing buffers drawn ahead of the current one, with their pair counts.
  uint32_t *shuf_ahead[2] = {nullptr, nullptr};
  uint64_t *nidx_ahead[2] = {nullptr, nullptr};
  // Fisher-Yates pairings of several epochs drawn at once: scratch and
  // results for fyb_epochs epochs, and a device scalar holding n / 2.
  FyScratch fyb{};
  uint32_t *shuf_many = nullptr;
  uint64_t *nidx_half = nullptr;
  uint32_t fyb_epochs = 0;
  unsigned long long *hist = nullptr;  // byte histogram, 256 entries
  uint32_t *owner = nullptr;
  DeferItem *ditems = nullptr;
  uint32_t *dqueue = nullptr;   // ready queue of the resolver, defer_cap slots
  uint64_t dqueue_dirty = 0;    // slots the last segment filled (cleared before the next)
  DeferCtl *dctl = nullptr;
  uint64_t defer_cap = 0;
  uint32_t defer_gen = 0;
  uint32_t defer_budget = kDeferBudget;
  uint64_t defer_items = 0;  // interactions deferred since creation
  double resolver_idle_s = 0.1;  // the resolver leaves after this long with nothing deferred
  uint64_t sr_chunk = kSelfrepChunk;  // programs per chunk of the levelled self-rep check
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
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
int dev_alloc(T **p, uint64_t count) {
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p),
                           count ? count * sizeof(T) : sizeof(T));
  if (e != hipSuccess)
    return fail(CUBFF_ENOMEM, std::string("hipMalloc failed: ") +
                                  hipGetErrorString(e));
  return CUBFF_OK;
}

template <typename T>
void dev_free(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int set_device(int device) {
  HIP_TRY(hipSetDevice(device));
  return CUBFF_OK;
}

int check_ctx(cubff_ctx *ctx) {
  if (!ctx) return fail(CUBFF_EINVAL, "null context");
  return set_device(ctx->device);
}

int ensure_2d_scratch(cubff_ctx *c) {
  if (c->g2.order) return CUBFF_OK;
  const uint64_t n = c->n;
  int r;
  if ((r = dev_alloc(&c->g2.order, n)) || (r = dev_alloc(&c->g2.eu, n)) ||
      (r = dev_alloc(&c->g2.ew, n)) || (r = dev_alloc(&c->g2.minpri, n)) ||
      (r = dev_alloc(&c->g2.rank, n)) || (r = dev_alloc(&c->g2.undecided, 1)) ||
      (r = dev_alloc(&c->g2.state, n)) ||
      (r = dev_alloc(&c->g2.used, (n + 3) / 4 * 4)) ||
      (r = dev_alloc(&c->g2.rest_total, 1)))
    return r;
  return CUBFF_OK;
}

// Multi-epoch segments: at most kFlowMaxEpochs epochs and kFlowBytes of
// scratch (pairings, pair-of-program maps, pending counts, queue: 12 bytes
// per program and epoch).
constexpr uint64_t kFlowMaxEpochs = 64;
constexpr uint64_t kFlowBytes = 1ULL << 30;
constexpr double kFlowPatienceS = 2e-6;
// Scratch and results of Fisher-Yates pairings drawn several epochs at once.
constexpr uint64_t kFyBatchBytes = 256ULL << 20;
// Resolver wavefronts (one interaction each); 256 measured 2.5% faster per
// epoch than 512 and 1,024 at 2^17 bff_noheads, 128 within 1%.
constexpr unsigned kResolveWaves = 256;
// Entries awaiting the resolver past which epoch kernels defer nothing more.
constexpr uint32_t kMaxBacklog = 1024;
constexpr uint64_t kLoopSegmentEpochs = 64;  // epochs per synchronised loop segment
// How long an epoch kernel's lane waits for the deferred interaction that
// owes one of its programs before deferring its own (~half an epoch kernel
// at 2^17 programs).
constexpr double kOweWaitS = 200e-6;
// A run of the deferring epoch kernel still going this long after the
// kernel's start stops as if its command budget were spent (the resolver
// goes on from it): the rare runs whose commands mostly take the slow path
// no longer hold their epoch (2^17 bff_noheads: 0.425 -> 0.401 ms per epoch;
// 150 us and less defer too much).
constexpr double kRunLimitS = 200e-6;
// Deferring segments end by waiting for the resolver to finish the last
// epochs' long interactions: long segments (bounded by kFlowBytes of list
// entries, 16 bytes per program and epoch) amortize that wait.
constexpr uint64_t kDeferMaxEpochs = 1024;

// Which multi-epoch path the default picks (measured on MI355X at 2^17
// programs, 32-64 epochs, against one launch per epoch).  Deferring epoch
// kernels: bff 2.0x, bff_perm 1.8x, bff_noheads 1.4-1.6x, subleq 1.3x,
// bff_selfmove 1.4x, rsubleq4 2.4x -- languages whose interactions rarely
// exceed the budget, or whose long runs a cycle skip ends early.
// Dataflow kernel: the Forth family 1.6-2.2x, bff_noheads_4bit 1.3x --
// long interactions are common there and would mostly be deferred.  bff8
// and bff8_noheads (whose capped runs the full-state cycle skip ends early)
// are fastest with one launch per epoch.  Up to 2^18 programs; at 2^22 the epochs are throughput-bound
// and one launch per epoch is used.
bool defer_pays(int lang, uint64_t n) {
  if (n > (1ULL << 18)) return false;
  return lang == kBff || lang == kBffNoheads || lang == kBffPerm || lang == kSubleq ||
         lang == kBffSelfmove || lang == kRsubleq4;
}

bool flow_pays(int lang, uint64_t n) {
  if (n > (1ULL << 18)) return false;
  switch (lang) {
    case kBffNoheads4bit:
    case kForth: case kForthcopy: case kForthtrivial: case kForthtrivialReset:
      return true;
    default:
      return false;
  }
}

// Deferred-interaction segments: the list holds every pair slot of the
// segment in the worst case (a 64-byte entry and a 4-byte ready-queue slot
// each), within kFlowBytes.
uint64_t defer_segment_epochs(uint64_t n) {
  return std::max<uint64_t>(2, std::min(kFlowBytes / (34 * n), kDeferMaxEpochs));
}

uint64_t flow_segment_epochs(uint64_t n) {
  return std::min(kFlowBytes / (12 * n), kFlowMaxEpochs);
}

int ensure_flow_scratch(cubff_ctx *c, uint64_t epochs) {
  if (c->flow_cap >= epochs) return CUBFF_OK;
  dev_free(c->flow_shuf);
  dev_free(c->flow_pairof);
  dev_free(c->flow_cnt);
  dev_free(c->flow_queue);
  c->flow_cap = 0;
  const uint64_t n = c->n;
  int r;
  if ((r = dev_alloc(&c->flow_shuf, epochs * n)) || (r = dev_alloc(&c->flow_pairof, epochs * n)) ||
      (r = dev_alloc(&c->flow_cnt, epochs * (n / 2))) ||
      (r = dev_alloc(&c->flow_queue, epochs * (n / 2))))
    return r;
  if (!c->flow_ctl && (r = dev_alloc(&c->flow_ctl, 1))) return r;
  c->flow_cap = epochs;
  return CUBFF_OK;
}

// Epochs whose Fisher-Yates pairings are drawn in one set of launches: up
// to 32, within kFyBatchBytes of scratch and results (32 B per program and
// epoch).
uint32_t fy_batch_epochs(uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(32, kFyBatchBytes / (32 * n + 4)));
}

void free_fy_batch(cubff_ctx *c) {
  for (uint32_t *p : {c->fyb.j, c->fyb.cnt, c->fyb.base, c->fyb.cursor, c->fyb.items,
                      c->fyb.first, c->fyb.succ, c->shuf_many})
    dev_free(p);
  dev_free(c->nidx_half);
  c->fyb = FyScratch{};
  c->shuf_many = nullptr;
  c->nidx_half = nullptr;
  c->fyb_epochs = 0;
}

int ensure_fy_batch(cubff_ctx *c, uint32_t epochs) {
  if (c->fyb_epochs >= epochs) return CUBFF_OK;
  free_fy_batch(c);
  const uint64_t m = c->n * epochs;
  int r;
  if ((r = dev_alloc(&c->fyb.j, m)) || (r = dev_alloc(&c->fyb.cnt, m + epochs)) ||
      (r = dev_alloc(&c->fyb.base, m)) || (r = dev_alloc(&c->fyb.cursor, m)) ||
      (r = dev_alloc(&c->fyb.items, m)) || (r = dev_alloc(&c->fyb.first, m)) ||
      (r = dev_alloc(&c->fyb.succ, m)) || (r = dev_alloc(&c->shuf_many, m)) ||
      (r = dev_alloc(&c->nidx_half, 1))) {
    free_fy_batch(c);
    return r;
  }
  c->fyb_epochs = epochs;
  return CUBFF_OK;
}

int pair_into(cubff_ctx *c, int mode, uint64_t params_seed, uint64_t epoch, uint32_t *dst,
              hipStream_t st = nullptr, uint64_t *num_indices = nullptr) {
  if (!st) st = c->stream;
  switch (mode) {
    case CUBFF_PAIR_2D:
      if (!c->row_off || !num_indices)
        return fail(CUBFF_ESTATE, "2-D pairing needs cubff_set_interactions first");
      HIP_TRY(launch_pair_2d(st, c->fy, c->g2, c->row_off, c->nbr, dst, c->n, params_seed, epoch,
                             num_indices));
      return CUBFF_OK;
    case CUBFF_PAIR_SHUFFLE:
      HIP_TRY(launch_pair_shuffle(st, c->fy, dst, c->n, params_seed, epoch));
      return CUBFF_OK;
    case CUBFF_PAIR_FIXED:
      HIP_TRY(launch_pair_fixed(st, dst, c->n, epoch));
      return CUBFF_OK;
    case CUBFF_PAIR_CYCLIC:
      HIP_TRY(launch_pair_cyclic(st, dst, c->n, epoch));
      return CUBFF_OK;
    default:
      return fail(CUBFF_EINVAL, "pairing mode without a closed form or shuffle");
  }
}

// Pairings of a run of epochs, one of two ways.  Fisher-Yates pairings are
// drawn `chunk` epochs at a time (one set of five launches per chunk) on the
// epoch stream, before the first epoch of each chunk: drawn one epoch at a
// time beside the epoch kernels instead, the pairing kernels (small and
// latency-bound) slowed those kernels by more than they cost alone.  Other
// modes are drawn two epochs ahead on a third stream into three buffers
// (the context's shuf and shuf_ahead, with their pair counts): no pairing
// depends on the soup.  Epoch j's kernel waits for its pairing; the pairing
// of epoch j+3 waits for epoch j's kernel, which reads the same buffer.
// Epoch j's kernel is queued before pairing j+2 is drawn, since the 2-D
// pairing blocks the host while its rounds converge.
struct PairAhead {
  static constexpr uint32_t kBufs = 3;  // pairings are drawn kBufs - 1 epochs ahead
  cubff_ctx *c;
  int mode;
  uint64_t seed, first;
  uint32_t *buf[kBufs] = {};
  uint64_t *nidx[kBufs] = {};  // pairs evaluated, per buffer
  hipEvent_t drawn[kBufs] = {}, used[kBufs] = {};
  uint32_t chunk = 0;  // > 0: Fisher-Yates pairings drawn this many epochs at a time

  int init(uint32_t epochs) {
    int r;
    if (mode == CUBFF_PAIR_SHUFFLE) {
      // The scratch is sized for the longest chunk at once: freeing it later
      // would wait for every stream, the deferring runs' resolver included.
      if ((r = ensure_fy_batch(c, fy_batch_epochs(c->n)))) return r;
      chunk = std::min(fy_batch_epochs(c->n), std::max(epochs, 1u));
      HIP_TRY(launch_set_u64(c->stream, c->nidx_half, c->n / 2));
      return CUBFF_OK;
    }
    if (!c->stream3) {
      hipError_t e = hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking);
      if (e != hipSuccess)
        return fail(CUBFF_EHIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    }
    for (uint32_t i = 0; i + 1 < kBufs; i++) {
      if (!c->shuf_ahead[i] && (r = dev_alloc(&c->shuf_ahead[i], c->n))) return r;
      if (!c->nidx_ahead[i] && (r = dev_alloc(&c->nidx_ahead[i], 1))) return r;
    }
    buf[0] = c->shuf;
    nidx[0] = c->num_indices;
    for (uint32_t i = 1; i < kBufs; i++) {
      buf[i] = c->shuf_ahead[i - 1];
      nidx[i] = c->nidx_ahead[i - 1];
    }
    for (uint32_t i = 0; i < kBufs; i++) {
      HIP_TRY(hipEventCreateWithFlags(&drawn[i], hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&used[i], hipEventDisableTiming));
    }
    // The stream of the first pairings must see everything queued so far.
    HIP_TRY(hipEventRecord(used[0], c->stream));
    HIP_TRY(hipStreamWaitEvent(c->stream3, used[0], 0));
    return CUBFF_OK;
  }
  // The pairing epoch first + j reads, and its pair count.
  uint32_t *shuf(uint32_t j) const {
    return chunk ? c->shuf_many + (uint64_t)(j % chunk) * c->n : buf[j % kBufs];
  }
  uint64_t *count(uint32_t j) const { return chunk ? c->nidx_half : nidx[j % kBufs]; }
  // Pairing of epoch first + j into buffer j % kBufs, after the kernel that
  // last read that buffer (epoch j - kBufs).
  int draw(uint32_t j) {
    int r;
    const uint32_t b = j % kBufs;
    if (j >= kBufs) HIP_TRY(hipStreamWaitEvent(c->stream3, used[b], 0));
    if ((r = pair_into(c, mode, seed, first + j, buf[b], c->stream3, nidx[b]))) return r;
    if (mode != CUBFF_PAIR_2D) HIP_TRY(launch_set_u64(c->stream3, nidx[b], c->n / 2));
    HIP_TRY(hipEventRecord(drawn[b], c->stream3));
    return CUBFF_OK;
  }
  // Draws of the first epochs, before any kernel.
  int start(uint32_t epochs) {
    int r;
    for (uint32_t j = 0; !chunk && j + 1 < kBufs && j < epochs; j++)
      if ((r = draw(j))) return r;
    return CUBFF_OK;
  }
  // After epoch j's kernel is queued: the pairing kBufs - 1 epochs on.
  int next(uint32_t j, uint32_t epochs) {
    const uint32_t k = j + kBufs - 1;
    return !chunk && k < epochs ? draw(k) : CUBFF_OK;
  }
  int before_kernel(uint32_t j, uint32_t epochs) {
    if (chunk) {
      if (j % chunk == 0)
        HIP_TRY(launch_pair_shuffle_many(c->stream, c->fyb, c->shuf_many, c->n, seed, first + j,
                                         std::min(chunk, epochs - j)));
      return CUBFF_OK;
    }
    HIP_TRY(hipStreamWaitEvent(c->stream, drawn[j % kBufs], 0));
    return CUBFF_OK;
  }
  int after_kernel(uint32_t j) {
    if (!chunk) HIP_TRY(hipEventRecord(used[j % kBufs], c->stream));
    return CUBFF_OK;
  }
  // The last epoch's pairing becomes the context's current one (copied out
  // of the chunk, or its buffer swapped in; the other buffers stay with the
  // context).
  int finish(uint32_t epochs) {
    if (!epochs) return CUBFF_OK;
    if (chunk) {
      HIP_TRY(hipMemcpyAsync(c->shuf, shuf(epochs - 1), c->n * sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, c->stream));
      HIP_TRY(launch_set_u64(c->stream, c->num_indices, c->n / 2));
      return CUBFF_OK;
    }
    const uint32_t last = (epochs - 1) % kBufs;
    c->shuf = buf[last];
    c->num_indices = nidx[last];
    for (uint32_t i = 1; i < kBufs; i++) {
      c->shuf_ahead[i - 1] = buf[(last + i) % kBufs];
      c->nidx_ahead[i - 1] = nidx[(last + i) % kBufs];
    }
    return CUBFF_OK;
  }
  ~PairAhead() {
    for (uint32_t i = 0; i < kBufs; i++) {
      if (drawn[i]) (void)hipEventDestroy(drawn[i]);
      if (used[i]) (void)hipEventDestroy(used[i]);
    }
  }
};

// Epochs [first, first + epochs), one launch each, pairings drawn ahead.
int run_loop_segment(cubff_ctx *c, int mode, uint64_t params_seed, uint64_t first,
                     uint32_t epochs, uint32_t prob, float *kern_ms) {
  int r;
  PairAhead pa{c, mode, params_seed, first};
  if ((r = pa.init(epochs))) return r;
  std::vector<hipEvent_t> kev(kern_ms ? 2 * epochs : 0, nullptr);
  auto cleanup = [&]() {
    for (auto &x : kev)
      if (x) (void)hipEventDestroy(x);
  };
  hipError_t e = hipSuccess;
  for (auto &x : kev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  if (e != hipSuccess) {
    cleanup();
    return fail(CUBFF_EHIP, std::string("epoch setup failed: ") + hipGetErrorString(e));
  }
  const uint64_t sm = splitmix64(params_seed);
  r = pa.start(epochs);
  for (uint32_t j = 0; r == CUBFF_OK && e == hipSuccess && j < epochs; j++) {
    if ((r = pa.before_kernel(j, epochs))) break;
    if (kern_ms) e = hipEventRecord(kev[2 * j], c->stream);
    if (e == hipSuccess)
      e = launch_epoch(c->stream, c->lang, c->soup, pa.shuf(j), c->n, pa.count(j),
                       epoch_seed_pre(sm, first + j), prob, c->pair_ops, c->pair_steps,
                       c->counters);
    if (e == hipSuccess && kern_ms) e = hipEventRecord(kev[2 * j + 1], c->stream);
    if (e == hipSuccess) r = pa.after_kernel(j);
    if (r == CUBFF_OK && e == hipSuccess) r = pa.next(j, epochs);
  }
  hipError_t e2 = hipStreamSynchronize(c->stream3);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && r == CUBFF_OK && kern_ms) {
    for (uint32_t j = 0; e == hipSuccess && j < epochs; j++) {
      float ms = 0;
      e = hipEventElapsedTime(&ms, kev[2 * j], kev[2 * j + 1]);
      *kern_ms += ms;
    }
  }
  cleanup();
  if (r) return r;
  if (e != hipSuccess) return fail(CUBFF_EHIP, std::string("epoch run failed: ") + hipGetErrorString(e));
  return pa.finish(epochs);
}

// Epochs [first, first + epochs) as one dataflow launch; leaves the last
// epoch's pairing in c->shuf and its per-pair results in pair_ops/steps.
// With ev, records ev[0]/ev[1] around the dataflow kernel.  Blocks until
// done to check the kernel's abort flag.
int run_flow_segment(cubff_ctx *c, int mode, uint64_t params_seed, uint64_t first,
                     uint32_t epochs, uint32_t prob, hipEvent_t *ev) {
  int r;
  if ((r = ensure_flow_scratch(c, epochs))) return r;
  const uint64_t n = c->n;
  if (mode == CUBFF_PAIR_SHUFFLE) {
    // Fisher-Yates pairings several epochs per set of launches.
    const uint32_t chunk = std::min(fy_batch_epochs(n), epochs);
    if ((r = ensure_fy_batch(c, fy_batch_epochs(n)))) return r;
    for (uint32_t j = 0; j < epochs; j += chunk)
      HIP_TRY(launch_pair_shuffle_many(c->stream, c->fyb, c->flow_shuf + (uint64_t)j * n, n,
                                       params_seed, first + j, std::min(chunk, epochs - j)));
  } else {
    for (uint32_t j = 0; j < epochs; j++)
      if ((r = pair_into(c, mode, params_seed, first + j, c->flow_shuf + (uint64_t)j * n)))
        return r;
  }
  HIP_TRY(launch_flow_pairof(c->stream, c->flow_shuf, c->flow_pairof, n, epochs));
  int khz = 0;
  HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
  FlowArgs a{};
  a.soup = c->soup;
  a.shuf = c->flow_shuf;
  a.pairof = c->flow_pairof;
  a.cnt = c->flow_cnt;
  a.queue = c->flow_queue;
  a.ctl = c->flow_ctl;
  a.n = n;
  a.slots = (uint32_t)(n / 2);
  a.epochs = epochs;
  a.sm_seed = splitmix64(params_seed);
  a.first_epoch = first;
  a.prob = prob;
  a.pair_ops = c->pair_ops;
  a.pair_steps = c->pair_steps;
  a.counters = c->counters;
  a.spin_ticks = (uint64_t)(c->flow_wait_s * khz * 1000.0);
  a.patience_ticks = (uint64_t)(kFlowPatienceS * khz * 1000.0);
  if (ev) HIP_TRY(hipEventRecord(ev[0], c->stream));
  HIP_TRY(launch_flow(c->stream, c->lang, a));
  if (ev) HIP_TRY(hipEventRecord(ev[1], c->stream));
  HIP_TRY(hipMemcpyAsync(c->shuf, c->flow_shuf + (uint64_t)(epochs - 1) * n, n * 4,
                         hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(launch_set_u64(c->stream, c->num_indices, n / 2));
  FlowCtl ctl{};
  HIP_TRY(hipMemcpyAsync(&ctl, c->flow_ctl, sizeof(ctl), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->flow_claims += ctl.claims;
  c->flow_items += (uint64_t)epochs * (n / 2);
  if (ctl.abort)
    return fail(CUBFF_EHIP, "multi-epoch kernel stopped: a wait exceeded " +
                                std::to_string(c->flow_wait_s) + " s; soup state is undefined");
  return CUBFF_OK;
}

int ensure_defer_scratch(cubff_ctx *c, uint64_t epochs) {
  const uint64_t need = epochs * (c->n / 2);
  int r;
  if (!c->stream2) {
    hipError_t e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
    if (e != hipSuccess)
      return fail(CUBFF_EHIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
  }
  if (!c->owner && (r = dev_alloc(&c->owner, c->n))) return r;
  if (!c->dctl && (r = dev_alloc(&c->dctl, 1))) return r;
  if (c->defer_cap < need) {
    dev_free(c->ditems);
    dev_free(c->dqueue);
    c->defer_cap = 0;
    if ((r = dev_alloc(&c->ditems, need)) || (r = dev_alloc(&c->dqueue, need))) return r;
    // Generation words start at 0; runs use generations >= 1.  Queue slots
    // start empty.
    HIP_TRY(hipMemsetAsync(c->ditems, 0, need * sizeof(DeferItem), c->stream));
    HIP_TRY(launch_fill_u32(c->stream, c->dqueue, kNone, need));
    c->dqueue_dirty = 0;
    c->defer_cap = need;
  }
  // Slots the last segment filled start empty again.
  if (c->dqueue_dirty) HIP_TRY(launch_fill_u32(c->stream, c->dqueue, kNone, c->dqueue_dirty));
  c->dqueue_dirty = 0;
  return CUBFF_OK;
}

// Epochs [first, first + epochs) with the deferring epoch kernel on the
// context stream and the resolver on the second stream.  With kern_ms, adds
// the summed duration of the epoch kernels (HIP events around each).
int run_defer_segment(cubff_ctx *c, int mode, uint64_t params_seed, uint64_t first,
                      uint32_t epochs, uint32_t prob, float *kern_ms) {
  int r;
  if ((r = ensure_defer_scratch(c, epochs))) return r;
  const uint64_t n = c->n;
  int khz = 0;
  HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
  DeferArgs a{};
  a.soup = c->soup;
  a.owner = c->owner;
  a.items = c->ditems;
  a.queue = c->dqueue;
  a.ctl = c->dctl;
  a.n = n;
  a.slots = (uint32_t)(n / 2);
  a.epochs = epochs;
  a.capacity = (uint32_t)(epochs * (n / 2));
  a.budget = c->defer_budget;
  a.max_backlog = kMaxBacklog;
  a.gen = ++c->defer_gen;
  if (a.gen == 0) a.gen = c->defer_gen = 1;  // 0 marks never-written entries
  a.prob = prob;
  a.sm_seed = splitmix64(params_seed);
  a.first_epoch = first;
  a.pair_ops = c->pair_ops;
  a.pair_steps = c->pair_steps;
  a.counters = c->counters;
  a.spin_ticks = (uint64_t)(c->flow_wait_s * khz * 1000.0);
  a.idle_ticks = (uint64_t)(std::min(c->resolver_idle_s, c->flow_wait_s) * khz * 1000.0);
  a.owe_ticks = (uint64_t)(std::min(kOweWaitS, c->flow_wait_s) * khz * 1000.0);
  a.run_ticks = (uint64_t)(std::min(kRunLimitS, c->flow_wait_s) * khz * 1000.0);
  hipEvent_t started = nullptr, resolved = nullptr;
  std::vector<hipEvent_t> kev(kern_ms ? 2 * epochs : 0, nullptr);
  auto cleanup = [&]() {
    if (started) (void)hipEventDestroy(started);
    if (resolved) (void)hipEventDestroy(resolved);
    for (auto &x : kev)
      if (x) (void)hipEventDestroy(x);
  };
  hipError_t e = hipSuccess;
  for (auto &x : kev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&started, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&resolved, hipEventDisableTiming);
  // Pairing buffers are allocated before the resolver starts: it waits for
  // the epoch kernels, and an allocation that waits for the device would
  // keep them from being queued.
  PairAhead pa{c, mode, params_seed, first};
  if (e == hipSuccess && (r = pa.init(epochs))) {
    cleanup();
    return r;
  }
  if (e == hipSuccess) e = launch_fill_u32(c->stream, c->owner, kNone, n);
  if (e == hipSuccess) e = launch_defer_start(c->stream, c->dctl);
  // Until this segment ends well, every queue slot counts as used.
  c->dqueue_dirty = c->defer_cap;
  if (e == hipSuccess) e = hipEventRecord(started, c->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->stream2, started, 0);
  if (e == hipSuccess) e = launch_resolve(c->stream2, c->lang, a, kResolveWaves);
  if (e != hipSuccess) {
    cleanup();
    return fail(CUBFF_EHIP, std::string("deferred run setup failed: ") + hipGetErrorString(e));
  }
  r = pa.start(epochs);
  for (uint32_t j = 0; r == CUBFF_OK && e == hipSuccess && j < epochs; j++) {
    if ((r = pa.before_kernel(j, epochs))) break;
    if (kern_ms) e = hipEventRecord(kev[2 * j], c->stream);
    if (e == hipSuccess)
      e = launch_epoch_defer(c->stream, c->lang, a, pa.shuf(j), pa.count(j), j);
    if (e == hipSuccess && kern_ms) e = hipEventRecord(kev[2 * j + 1], c->stream);
    if (e == hipSuccess) r = pa.after_kernel(j);
    if (r == CUBFF_OK && e == hipSuccess) r = pa.next(j, epochs);
  }
  // Close the list whatever happened, so the resolver leaves.
  hipError_t e2 = launch_defer_close(c->stream, c->dctl);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipEventRecord(resolved, c->stream2);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, resolved, 0);
  DeferCtl ctl{};
  if (e == hipSuccess)
    e = hipMemcpyAsync(&ctl, c->dctl, sizeof(ctl), hipMemcpyDeviceToHost, c->stream);
  hipError_t e3 = hipStreamSynchronize(c->stream2);
  if (e == hipSuccess) e = e3;
  if (c->stream3) {
    e3 = hipStreamSynchronize(c->stream3);
    if (e == hipSuccess) e = e3;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && r == CUBFF_OK && kern_ms) {
    for (uint32_t j = 0; e == hipSuccess && j < epochs; j++) {
      float ms = 0;
      e = hipEventElapsedTime(&ms, kev[2 * j], kev[2 * j + 1]);
      *kern_ms += ms;
    }
  }
  cleanup();
  if (r) return r;
  if (e != hipSuccess) return fail(CUBFF_EHIP, std::string("deferred run failed: ") + hipGetErrorString(e));
  if ((r = pa.finish(epochs))) return r;
  // The resolver may leave early when it cannot run beside the epoch
  // kernels (the list then reads kDeferGone or more and holds nothing);
  // the run is complete as long as every deferred entry was run.
  const uint32_t listed = ctl.final_tail >= kDeferGone ? 0u : ctl.final_tail;
  c->defer_items += listed;
  c->dqueue_dirty = std::min<uint64_t>(std::max(listed, ctl.qtail), c->defer_cap);
  if (ctl.resolved != listed)
    return fail(CUBFF_EHIP, "deferred-interaction resolver stopped: a wait exceeded " +
                                std::to_string(c->flow_wait_s) + " s; soup state is undefined");
  return CUBFF_OK;
}

// The body of cubff_run_epochs / cubff_run_epochs_timed.  Closed-form and
// shuffle pairings of two or more epochs go through the dataflow kernel in
// segments; otherwise one pairing + epoch launch per epoch.  timing (2
// floats, nullable) receives the total and the epoch-kernel milliseconds.
int run_epochs_impl(cubff_ctx *c, int mode, uint64_t params_seed, uint64_t first_epoch,
                    uint64_t count, uint32_t prob, float *timing) {
  // The fixed shuffle of a soup whose size is not a power of two is not a
  // permutation (((i * offset) % n) ^ flip repeats programs: n = 6 at epoch
  // 1 gives 1,2,1,2,1,2).  The reference CLI rejects it (main.cc:233-237);
  // its library runs it with slots racing on shared programs.  Here it runs
  // one launch per epoch, whose lanes then race the same way; the dataflow
  // and deferring kernels, which track one writer per program, never see it.
  const bool permutation = mode != CUBFF_PAIR_FIXED || (c->n & (c->n - 1)) == 0;
  const bool closed_form = permutation && (mode == CUBFF_PAIR_SHUFFLE || mode == CUBFF_PAIR_FIXED ||
                                           mode == CUBFF_PAIR_CYCLIC);
  // Pairings drawn ahead: every device mode (none depends on the soup).
  const bool ahead = closed_form || (mode == CUBFF_PAIR_2D && c->row_off);
  const bool defer = (c->pipeline == CUBFF_FLOW_DEFER ||
                      (c->pipeline == CUBFF_FLOW_AUTO && defer_pays(c->lang, c->n))) &&
                     count >= 2 && ahead;
  const uint64_t seg = defer ? defer_segment_epochs(c->n) : flow_segment_epochs(c->n);
  const bool flow = !defer &&
                    (c->pipeline == CUBFF_FLOW_ON ||
                     (c->pipeline == CUBFF_FLOW_AUTO && flow_pays(c->lang, c->n))) &&
                    count >= 2 && seg >= 2 && closed_form;
  const uint64_t sm = splitmix64(params_seed);
  std::vector<hipEvent_t> ev;
  auto destroy = [&]() {
    for (auto &e : ev) (void)hipEventDestroy(e);
  };
  int r = CUBFF_OK;
  if (timing) {
    const uint64_t launches = (flow || defer) ? (count + seg - 1) / seg : count;  // upper bound
    ev.resize(2 + 2 * launches);
    for (auto &e : ev) {
      hipError_t he = hipEventCreate(&e);
      if (he != hipSuccess) {
        destroy();
        return fail(CUBFF_EHIP, std::string("hipEventCreate failed: ") + hipGetErrorString(he));
      }
    }
    hipError_t he = hipEventRecord(ev[0], c->stream);
    if (he != hipSuccess) r = fail(CUBFF_EHIP, hipGetErrorString(he));
  }
  uint64_t launch = 0;
  float defer_kern = 0;
  std::vector<bool> untimed(timing ? ev.size() / 2 : 0, false);
  for (uint64_t e = first_epoch; r == CUBFF_OK && e < first_epoch + count; launch++) {
    hipEvent_t *kev = timing ? &ev[2 + 2 * launch] : nullptr;
    const uint64_t left = first_epoch + count - e;
    if (flow && left >= 2) {
      const uint32_t ep = (uint32_t)std::min(left, seg);
      r = run_flow_segment(c, mode, params_seed, e, ep, prob, kev);
      e += ep;
      continue;
    }
    if (defer && left >= 2) {
      const uint32_t ep = (uint32_t)std::min(left, seg);
      r = run_defer_segment(c, mode, params_seed, e, ep, prob, timing ? &defer_kern : nullptr);
      if (timing) untimed[launch] = true;  // timed inside
      e += ep;
      continue;
    }
    if (!flow && !defer && ahead && left >= 2) {
      const uint32_t ep = (uint32_t)std::min<uint64_t>(left, kLoopSegmentEpochs);
      r = run_loop_segment(c, mode, params_seed, e, ep, prob, timing ? &defer_kern : nullptr);
      if (timing) untimed[launch] = true;  // timed inside
      e += ep;
      continue;
    }
    if ((r = cubff_pair(c, mode, params_seed, e))) break;
    hipError_t he = hipSuccess;
    if (kev) he = hipEventRecord(kev[0], c->stream);
    if (he == hipSuccess)
      he = launch_epoch(c->stream, c->lang, c->soup, c->shuf, c->n, c->num_indices,
                        epoch_seed_pre(sm, e), prob, c->pair_ops, c->pair_steps, c->counters);
    if (he == hipSuccess && kev) he = hipEventRecord(kev[1], c->stream);
    if (he != hipSuccess) r = fail(CUBFF_EHIP, std::string("epoch launch failed: ") + hipGetErrorString(he));
    e++;
  }
  if (r == CUBFF_OK && timing) {
    float tot = 0, kern = 0;
    hipError_t he = hipEventRecord(ev[1], c->stream);
    if (he == hipSuccess) he = hipEventSynchronize(ev[1]);
    if (he == hipSuccess) he = hipEventElapsedTime(&tot, ev[0], ev[1]);
    for (uint64_t i = 0; he == hipSuccess && i < launch; i++) {
      if (untimed[i]) continue;
      float ms = 0;
      he = hipEventElapsedTime(&ms, ev[2 + 2 * i], ev[3 + 2 * i]);
      kern += ms;
    }
    kern += defer_kern;
    if (he != hipSuccess) r = fail(CUBFF_EHIP, std::string("event timing failed: ") + hipGetErrorString(he));
    timing[0] = tot;
    timing[1] = kern;
  }
  destroy();
  return r;
}

}  // namespace

namespace cubff {
// Failures of entry points defined in other files of the library, reported
// through cubff_last_error().
int report_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace cubff

extern "C" {

const char *cubff_last_error(void) { return g_last_error.c_str(); }

int cubff_num_langs(void) { return kNumLangs; }

const char *cubff_lang_name(int lang) {
  return (lang >= 0 && lang < kNumLangs) ? kLangNames[lang] : nullptr;
}

int cubff_lang_id(const char *name) {
  if (!name) return -1;
  for (int i = 0; i < kNumLangs; i++)
    if (strcmp(name, kLangNames[i]) == 0) return i;
  return -1;
}

uint64_t cubff_splitmix64(uint64_t x) { return splitmix64(x); }

uint64_t cubff_epoch_seed(uint64_t params_seed, uint64_t x) {
  return epoch_seed_pre(splitmix64(params_seed), x);
}

int cubff_ctx_create(cubff_ctx **out, int lang, uint64_t n, int device) {
  if (!out) return fail(CUBFF_EINVAL, "null output pointer");
  *out = nullptr;
  if (lang < 0 || lang >= kNumLangs) return fail(CUBFF_EINVAL, "unknown language id");
  // common_language.h:365 requires an even count; indices are 32-bit.
  if (n == 0 || n % 2 != 0) return fail(CUBFF_EINVAL, "num_programs must be even and > 0");
  if (n >= (1ULL << 32)) return fail(CUBFF_EINVAL, "num_programs must be < 2^32");
  int r;
  if ((r = set_device(device))) return r;
  cubff_ctx *c = new cubff_ctx();
  c->lang = lang;
  c->device = device;
  c->n = n;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(CUBFF_EHIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
  }
  if ((r = dev_alloc(&c->soup, n * kTape)) || (r = dev_alloc(&c->shuf, n)) ||
      (r = dev_alloc(&c->num_indices, 1)) || (r = dev_alloc(&c->counters, kNumCounters)) ||
      (r = dev_alloc(&c->pair_ops, n / 2)) ||
      (r = dev_alloc(&c->pair_steps, n / 2)) ||
      (r = dev_alloc(&c->fy.j, n)) || (r = dev_alloc(&c->fy.cnt, n + 1)) ||
      (r = dev_alloc(&c->fy.base, n)) || (r = dev_alloc(&c->fy.cursor, n)) ||
      (r = dev_alloc(&c->fy.items, n)) || (r = dev_alloc(&c->fy.first, n)) ||
      (r = dev_alloc(&c->fy.succ, n))) {
    std::string keep = g_last_error;
    cubff_ctx_destroy(c);
    g_last_error = keep;
    return r;
  }
  // Identity pairing with every slot evaluated until the first pairing call
  // (the state of RunSimulation after :409-411).
  hipError_t e2 = hipSuccess;
  {
    uint32_t *host = new uint32_t[n];
    for (uint64_t i = 0; i < n; i++) host[i] = (uint32_t)i;
    e2 = hipMemcpyAsync(c->shuf, host, n * 4, hipMemcpyHostToDevice, c->stream);
    if (e2 == hipSuccess) e2 = hipStreamSynchronize(c->stream);
    delete[] host;
  }
  if (e2 == hipSuccess) e2 = launch_set_u64(c->stream, c->num_indices, n / 2);
  if (e2 == hipSuccess)
    e2 = launch_fill_u32(c->stream, reinterpret_cast<uint32_t *>(c->counters), 0, 2 * kNumCounters);
  if (e2 == hipSuccess) e2 = launch_fill_u32(c->stream, c->pair_ops, 0, n / 2);
  if (e2 == hipSuccess) e2 = launch_fill_u32(c->stream, c->pair_steps, 0, n / 2);
  if (e2 == hipSuccess) e2 = hipStreamSynchronize(c->stream);
  if (e2 != hipSuccess) {
    cubff_ctx_destroy(c);
    return fail(CUBFF_EHIP, std::string("context init failed: ") + hipGetErrorString(e2));
  }
  *out = c;
  return CUBFF_OK;
}

int cubff_ctx_destroy(cubff_ctx *c) {
  if (!c) return CUBFF_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dev_free(c->soup);
  dev_free(c->shuf);
  dev_free(c->num_indices);
  dev_free(c->counters);
  dev_free(c->pair_ops);
  dev_free(c->pair_steps);
  dev_free(c->selfrep);
  dev_free(c->sr.finals);
  dev_free(c->sr.list[0]);
  dev_free(c->sr.list[1]);
  dev_free(c->sr.counts);
  dev_free(c->fy.j);
  dev_free(c->fy.cnt);
  dev_free(c->fy.base);
  dev_free(c->fy.cursor);
  dev_free(c->fy.items);
  dev_free(c->fy.first);
  dev_free(c->fy.succ);
  free_fy_batch(c);
  dev_free(c->g2.order);
  dev_free(c->g2.eu);
  dev_free(c->g2.ew);
  dev_free(c->g2.minpri);
  dev_free(c->g2.rank);
  dev_free(c->g2.undecided);
  dev_free(c->g2.state);
  dev_free(c->g2.used);
  dev_free(c->g2.rest_total);
  dev_free(c->row_off);
  dev_free(c->nbr);
  dev_free(c->flow_shuf);
  dev_free(c->flow_pairof);
  dev_free(c->flow_cnt);
  dev_free(c->flow_queue);
  dev_free(c->flow_ctl);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  dev_free(c->owner);
  dev_free(c->ditems);
  dev_free(c->dqueue);
  dev_free(c->dctl);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream3) (void)hipStreamSynchronize(c->stream3);
  if (c->stream3) (void)hipStreamDestroy(c->stream3);
  for (int i = 0; i < 2; i++) {
    dev_free(c->shuf_ahead[i]);
    dev_free(c->nidx_ahead[i]);
  }
  dev_free(c->hist);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CUBFF_OK;
}

int cubff_init_programs(cubff_ctx *c, uint64_t seed_value, int zero_init,
                        uint64_t count) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (count > c->n) return fail(CUBFF_EINVAL, "count exceeds num_programs");
  HIP_TRY(launch_init(c->stream, c->soup, c->n, seed_value, count, zero_init));
  return CUBFF_OK;
}

int cubff_write_soup(cubff_ctx *c, uint64_t offset, const uint8_t *src,
                     uint64_t len) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (offset > c->n * kTape || len > c->n * kTape - offset)
    return fail(CUBFF_EINVAL, "write past the end of the soup");
  if (len == 0) return CUBFF_OK;
  if (!src) return fail(CUBFF_EINVAL, "null source");
  HIP_TRY(hipMemcpyAsync(c->soup + offset, src, len, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUBFF_OK;
}

int cubff_read_soup(cubff_ctx *c, uint64_t offset, uint8_t *dst, uint64_t len) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (offset > c->n * kTape || len > c->n * kTape - offset)
    return fail(CUBFF_EINVAL, "read past the end of the soup");
  if (len == 0) return CUBFF_OK;
  if (!dst) return fail(CUBFF_EINVAL, "null destination");
  HIP_TRY(hipMemcpyAsync(dst, c->soup + offset, len, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUBFF_OK;
}

int cubff_set_shuffle(cubff_ctx *c, const uint32_t *s, uint64_t num_indices) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (!s) return fail(CUBFF_EINVAL, "null shuffle");
  if (num_indices > c->n / 2) return fail(CUBFF_EINVAL, "num_indices exceeds num_programs/2");
  for (uint64_t i = 0; i < c->n; i++)
    if (s[i] >= c->n) return fail(CUBFF_EINVAL, "shuffle index out of range");
  HIP_TRY(hipMemcpyAsync(c->shuf, s, c->n * 4, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(launch_set_u64(c->stream, c->num_indices, num_indices));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUBFF_OK;
}

int cubff_set_interactions(cubff_ctx *c, const uint64_t *row_off,
                           const uint32_t *nbr) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (!row_off) return fail(CUBFF_EINVAL, "null row_off");
  const uint64_t n = c->n;
  if (row_off[0] != 0) return fail(CUBFF_EINVAL, "row_off[0] must be 0");
  for (uint64_t i = 0; i < n; i++)
    if (row_off[i + 1] < row_off[i]) return fail(CUBFF_EINVAL, "row_off must be non-decreasing");
  const uint64_t e = row_off[n];
  if (e && !nbr) return fail(CUBFF_EINVAL, "null nbr");
  for (uint64_t x = 0; x < e; x++)
    if (nbr[x] >= n)
      return fail(CUBFF_EINVAL, "invalid interaction pattern: programs not in [0, n) range.");
  if ((r = ensure_2d_scratch(c))) return r;
  dev_free(c->row_off);
  dev_free(c->nbr);
  if ((r = dev_alloc(&c->row_off, n + 1)) || (r = dev_alloc(&c->nbr, e ? e : 1))) return r;
  HIP_TRY(hipMemcpyAsync(c->row_off, row_off, (n + 1) * 8, hipMemcpyHostToDevice, c->stream));
  if (e) HIP_TRY(hipMemcpyAsync(c->nbr, nbr, e * 4, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUBFF_OK;
}

int cubff_pair(cubff_ctx *c, int mode, uint64_t params_seed, uint64_t epoch) {
  int r;
  if ((r = check_ctx(c))) return r;
  const uint64_t n = c->n;
  switch (mode) {
    case CUBFF_PAIR_SHUFFLE:
      HIP_TRY(launch_pair_shuffle(c->stream, c->fy, c->shuf, n, params_seed, epoch));
      HIP_TRY(launch_set_u64(c->stream, c->num_indices, n / 2));
      return CUBFF_OK;
    case CUBFF_PAIR_FIXED:
      HIP_TRY(launch_pair_fixed(c->stream, c->shuf, n, epoch));
      HIP_TRY(launch_set_u64(c->stream, c->num_indices, n / 2));
      return CUBFF_OK;
    case CUBFF_PAIR_CYCLIC:
      HIP_TRY(launch_pair_cyclic(c->stream, c->shuf, n, epoch));
      HIP_TRY(launch_set_u64(c->stream, c->num_indices, n / 2));
      return CUBFF_OK;
    case CUBFF_PAIR_2D:
      if (!c->row_off) return fail(CUBFF_ESTATE, "2-D pairing needs cubff_set_interactions first");
      HIP_TRY(launch_pair_2d(c->stream, c->fy, c->g2, c->row_off, c->nbr, c->shuf, n,
                             params_seed, epoch, c->num_indices));
      return CUBFF_OK;
    default:
      return fail(CUBFF_EINVAL, "unknown pairing mode");
  }
}

int cubff_run_epoch(cubff_ctx *c, uint64_t seed_value, uint32_t prob) {
  int r;
  if ((r = check_ctx(c))) return r;
  HIP_TRY(launch_epoch(c->stream, c->lang, c->soup, c->shuf, c->n, c->num_indices,
                       seed_value, prob, c->pair_ops, c->pair_steps, c->counters));
  return CUBFF_OK;
}

int cubff_run_epochs(cubff_ctx *c, int mode, uint64_t params_seed,
                     uint64_t first_epoch, uint64_t count, uint32_t prob) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (mode == CUBFF_PAIR_HOST) return fail(CUBFF_EINVAL, "run_epochs needs a device pairing mode");
  return run_epochs_impl(c, mode, params_seed, first_epoch, count, prob, nullptr);
}

int cubff_run_epochs_timed(cubff_ctx *c, int mode, uint64_t params_seed,
                           uint64_t first_epoch, uint64_t count, uint32_t prob,
                           float *total_ms, float *epoch_kernel_ms) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (mode == CUBFF_PAIR_HOST) return fail(CUBFF_EINVAL, "run_epochs needs a device pairing mode");
  if (mode == CUBFF_PAIR_2D) return fail(CUBFF_EINVAL, "timed runs support the closed-form and shuffle pairings");
  float t[2] = {0, 0};
  if ((r = run_epochs_impl(c, mode, params_seed, first_epoch, count, prob, t))) return r;
  if (total_ms) *total_ms = t[0];
  if (epoch_kernel_ms) *epoch_kernel_ms = t[1];
  return CUBFF_OK;
}

int cubff_byte_histogram(cubff_ctx *c, uint64_t *counts) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (!counts) return fail(CUBFF_EINVAL, "null output");
  if (!c->hist && (r = dev_alloc(&c->hist, 256))) return r;
  HIP_TRY(launch_fill_u32(c->stream, reinterpret_cast<uint32_t *>(c->hist), 0, 512));
  HIP_TRY(launch_byte_hist(c->stream, c->soup, c->n * kTape, c->hist));
  HIP_TRY(hipMemcpyAsync(counts, c->hist, 256 * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUBFF_OK;
}

int cubff_flow_stats(cubff_ctx *c, uint64_t *items, uint64_t *claims, uint64_t *deferred) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (items) *items = c->flow_items;
  if (claims) *claims = c->flow_claims;
  if (deferred) *deferred = c->defer_items;
  return CUBFF_OK;
}

int cubff_set_resolver_idle(cubff_ctx *c, double seconds) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (!(seconds >= 0)) return fail(CUBFF_EINVAL, "idle time must be >= 0");
  c->resolver_idle_s = seconds;
  return CUBFF_OK;
}

int cubff_set_selfrep_chunk(cubff_ctx *c, uint64_t programs) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (programs > kSelfrepChunk)
    return fail(CUBFF_EINVAL, "self-rep chunk exceeds 2^22 programs");
  c->sr_chunk = programs ? programs : kSelfrepChunk;
  // The next check allocates the scratch for the new size.
  HIP_TRY(hipStreamSynchronize(c->stream));
  dev_free(c->sr.finals);
  dev_free(c->sr.list[0]);
  dev_free(c->sr.list[1]);
  dev_free(c->sr.counts);
  c->sr.chunk = 0;
  return CUBFF_OK;
}

int cubff_set_defer_budget(cubff_ctx *c, uint32_t budget) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (budget == 0) return fail(CUBFF_EINVAL, "budget must be > 0 commands");
  c->defer_budget = budget;
  return CUBFF_OK;
}

int cubff_set_pipeline(cubff_ctx *c, int enable, double wait_limit_s) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (!(wait_limit_s >= 0)) return fail(CUBFF_EINVAL, "wait limit must be >= 0 seconds");
  if (enable != CUBFF_FLOW_OFF && enable != CUBFF_FLOW_ON && enable != CUBFF_FLOW_AUTO &&
      enable != CUBFF_FLOW_DEFER)
    return fail(CUBFF_EINVAL, "pipeline mode must be CUBFF_FLOW_OFF, _ON, _AUTO or _DEFER");
  c->pipeline = enable;
  c->flow_wait_s = wait_limit_s;
  return CUBFF_OK;
}

int cubff_read_counters(cubff_ctx *c, uint64_t *ops, uint64_t *num_runs,
                        uint64_t *steps) {
  int r;
  if ((r = check_ctx(c))) return r;
  unsigned long long h[kNumCounters];
  HIP_TRY(hipMemcpyAsync(h, c->counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (ops) *ops = h[0];
  if (num_runs) *num_runs = h[1];
  if (steps) *steps = h[2];
  return CUBFF_OK;
}

int cubff_reset_counters(cubff_ctx *c) {
  int r;
  if ((r = check_ctx(c))) return r;
  HIP_TRY(launch_fill_u32(c->stream, reinterpret_cast<uint32_t *>(c->counters), 0, 2 * kNumCounters));
  return CUBFF_OK;
}

int cubff_read_pair_ops(cubff_ctx *c, uint32_t *ops, uint32_t *steps,
                        uint64_t num_slots) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (num_slots > c->n / 2) return fail(CUBFF_EINVAL, "num_slots exceeds num_programs/2");
  if (ops) HIP_TRY(hipMemcpyAsync(ops, c->pair_ops, num_slots * 4, hipMemcpyDeviceToHost, c->stream));
  if (steps) HIP_TRY(hipMemcpyAsync(steps, c->pair_steps, num_slots * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUBFF_OK;
}

int cubff_read_shuffle(cubff_ctx *c, uint32_t *s, uint64_t *num_indices) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (s) HIP_TRY(hipMemcpyAsync(s, c->shuf, c->n * 4, hipMemcpyDeviceToHost, c->stream));
  uint64_t k = 0;
  HIP_TRY(hipMemcpyAsync(&k, c->num_indices, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (num_indices) *num_indices = k;
  return CUBFF_OK;
}

int cubff_selfrep(cubff_ctx *c, uint64_t seed_value, uint64_t count,
                  uint64_t *result) {
  int r;
  if ((r = check_ctx(c))) return r;
  if (count > c->n) return fail(CUBFF_EINVAL, "count exceeds num_programs");
  if (count && !result) return fail(CUBFF_EINVAL, "null result");
  if (!c->selfrep && (r = dev_alloc(&c->selfrep, c->n))) return r;
  if (selfrep_levels_supported(c->lang)) {
    if (!c->sr.finals) {
      const uint64_t chunk = std::min<uint64_t>(c->n, c->sr_chunk);
      if ((r = dev_alloc(&c->sr.finals, chunk * 13 * 128)) ||
          (r = dev_alloc(&c->sr.list[0], chunk * 13)) ||
          (r = dev_alloc(&c->sr.list[1], chunk * 13)) || (r = dev_alloc(&c->sr.counts, 2)))
        return r;
      c->sr.chunk = chunk;
    }
    HIP_TRY(launch_selfrep_levels(c->stream, c->lang, c->soup, c->n, count, seed_value,
                                  c->selfrep, c->sr));
  } else {
    HIP_TRY(launch_selfrep(c->stream, c->lang, c->soup, c->n, count, seed_value, c->selfrep));
  }
  if (count)
    HIP_TRY(hipMemcpyAsync(result, c->selfrep, count * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUBFF_OK;
}

int cubff_eval_tapes(int lang, int device, uint8_t *tapes, uint64_t count,
                     uint64_t steps, uint32_t *ops, uint32_t *steps_out) {
  if (lang < 0 || lang >= kNumLangs) return fail(CUBFF_EINVAL, "unknown language id");
  if (count && !tapes) return fail(CUBFF_EINVAL, "null tapes");
  if (steps > 0xffffffffULL) return fail(CUBFF_EINVAL, "steps must fit in 32 bits");
  int r;
  if ((r = set_device(device))) return r;
  if (count == 0) return CUBFF_OK;
  uint8_t *d_t = nullptr;
  uint32_t *d_o = nullptr, *d_s = nullptr;
  if ((r = dev_alloc(&d_t, count * kTape2)) || (r = dev_alloc(&d_o, count)) ||
      (r = dev_alloc(&d_s, count))) {
    dev_free(d_t);
    dev_free(d_o);
    return r;
  }
  hipError_t e = hipMemcpy(d_t, tapes, count * kTape2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_eval_tapes(nullptr, lang, d_t, count, (uint32_t)steps, d_o, d_s);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(tapes, d_t, count * kTape2, hipMemcpyDeviceToHost);
  if (e == hipSuccess && ops) e = hipMemcpy(ops, d_o, count * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && steps_out) e = hipMemcpy(steps_out, d_s, count * 4, hipMemcpyDeviceToHost);
  dev_free(d_t);
  dev_free(d_o);
  dev_free(d_s);
  if (e != hipSuccess)
    return fail(CUBFF_EHIP, std::string("eval_tapes failed: ") + hipGetErrorString(e));
  return CUBFF_OK;
}

int cubff_trace_bff(int lang, int device, const uint8_t *tape, uint64_t steps, uint8_t *out,
                    uint64_t out_bytes, uint64_t *file_bytes, uint64_t *ops) {
  if (lang < kBff || lang > kBffSelfmove) return fail(CUBFF_EINVAL, "step traces cover the BFF family");
  if (!tape || !out) return fail(CUBFF_EINVAL, "null tape or output");
  if (steps > 0xffffffffULL) return fail(CUBFF_EINVAL, "steps must fit in 32 bits");
  const uint64_t rec = 12 + kTape2;
  if (out_bytes < 12 + rec * steps)
    return fail(CUBFF_EINVAL, "output must hold 12 + 140 * steps bytes");
  int r;
  if ((r = set_device(device))) return r;
  uint8_t *d_t = nullptr, *d_o = nullptr;
  uint64_t *d_m = nullptr;
  if ((r = dev_alloc(&d_t, kTape2)) || (r = dev_alloc(&d_o, rec * std::max<uint64_t>(steps, 1))) ||
      (r = dev_alloc(&d_m, 3))) {
    dev_free(d_t);
    dev_free(d_o);
    return r;
  }
  uint64_t meta[3] = {0, 0, 0};
  hipError_t e = hipMemcpy(d_t, tape, kTape2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_trace_bff(nullptr, lang, d_t, (uint32_t)steps, d_o, steps, d_m);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(meta, d_m, sizeof(meta), hipMemcpyDeviceToHost);
  if (e == hipSuccess && meta[0] > steps) e = hipErrorUnknown;
  if (e == hipSuccess && meta[0]) e = hipMemcpy(out + 12, d_o, rec * meta[0], hipMemcpyDeviceToHost);
  dev_free(d_t);
  dev_free(d_o);
  dev_free(d_m);
  if (e != hipSuccess) return fail(CUBFF_EHIP, std::string("trace failed: ") + hipGetErrorString(e));
  // Header: magic "BFF\0", format version 1, tape size (u32 little-endian).
  const uint32_t head[3] = {0x00464642u, 1u, (uint32_t)kTape2};
  for (int q = 0; q < 3; q++)
    for (int b = 0; b < 4; b++) out[4 * q + b] = (uint8_t)(head[q] >> (8 * b));
  if (file_bytes) *file_bytes = 12 + rec * meta[0];
  if (ops) *ops = meta[1];
  return CUBFF_OK;
}

static_assert(kStepRecordBytes == CUBFF_STEP_RECORD_BYTES, "step record size");

int cubff_trace_steps(int lang, int device, uint8_t *tape, uint64_t steps, uint8_t *records,
                      uint64_t cap, uint64_t *num_records, uint64_t *ops, uint64_t *steps_run) {
  if (lang < 0 || lang >= kNumLangs) return fail(CUBFF_EINVAL, "unknown language id");
  if (!tape || (cap && !records)) return fail(CUBFF_EINVAL, "null tape or records");
  if (steps > 0xffffffffULL) return fail(CUBFF_EINVAL, "steps must fit in 32 bits");
  int r;
  if ((r = set_device(device))) return r;
  uint8_t *d_t = nullptr, *d_o = nullptr;
  uint64_t *d_m = nullptr;
  if ((r = dev_alloc(&d_t, kTape2)) ||
      (r = dev_alloc(&d_o, (uint64_t)kStepRecordBytes * std::max<uint64_t>(cap, 1))) ||
      (r = dev_alloc(&d_m, 3))) {
    dev_free(d_t);
    dev_free(d_o);
    return r;
  }
  uint64_t meta[3] = {0, 0, 0};
  hipError_t e = hipMemcpy(d_t, tape, kTape2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_trace_steps(nullptr, lang, d_t, (uint32_t)steps, d_o, cap, d_m);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(meta, d_m, sizeof(meta), hipMemcpyDeviceToHost);
  if (e == hipSuccess && meta[0] > cap) e = hipErrorUnknown;
  if (e == hipSuccess && meta[0])
    e = hipMemcpy(records, d_o, (uint64_t)kStepRecordBytes * meta[0], hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(tape, d_t, kTape2, hipMemcpyDeviceToHost);
  dev_free(d_t);
  dev_free(d_o);
  dev_free(d_m);
  if (e != hipSuccess) return fail(CUBFF_EHIP, std::string("trace failed: ") + hipGetErrorString(e));
  if (num_records) *num_records = meta[0];
  if (ops) *ops = meta[1];
  if (steps_run) *steps_run = meta[2];
  return CUBFF_OK;
}

int cubff_eval_selfrep(int lang, int device, const uint8_t *prog, uint64_t len,
                       uint64_t epoch, uint64_t seed, uint64_t *result) {
  return cubff_eval_selfrep_trace(lang, device, prog, len, epoch, seed, result, nullptr);
}

int cubff_eval_selfrep_trace(int lang, int device, const uint8_t *prog, uint64_t len,
                             uint64_t epoch, uint64_t seed, uint64_t *result, uint8_t *tapes) {
  if (lang < 0 || lang >= kNumLangs) return fail(CUBFF_EINVAL, "unknown language id");
  if (!result) return fail(CUBFF_EINVAL, "null result");
  int r;
  if ((r = set_device(device))) return r;
  uint8_t padded[kTape] = {};
  if (len > kTape) len = kTape;
  if (len) memcpy(padded, prog, len);
  uint8_t *d_p = nullptr, *d_tr = nullptr;
  uint64_t *d_r = nullptr;
  if ((r = dev_alloc(&d_p, kTape)) || (r = dev_alloc(&d_r, 1)) ||
      (tapes && (r = dev_alloc(&d_tr, CUBFF_SELFREP_TRACE_BYTES)))) {
    dev_free(d_p);
    dev_free(d_r);
    return r;
  }
  const uint64_t seed_value = epoch_seed_pre(splitmix64(seed), epoch);  // :334
  hipError_t e = hipMemcpy(d_p, padded, kTape, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_selfrep(nullptr, lang, d_p, 1, 1, seed_value, d_r, d_tr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(result, d_r, 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess && tapes)
    e = hipMemcpy(tapes, d_tr, CUBFF_SELFREP_TRACE_BYTES, hipMemcpyDeviceToHost);
  dev_free(d_p);
  dev_free(d_r);
  dev_free(d_tr);
  if (e != hipSuccess)
    return fail(CUBFF_EHIP, std::string("eval_selfrep failed: ") + hipGetErrorString(e));
  return CUBFF_OK;
}

int cubff_sync(cubff_ctx *c) {
  int r;
  if ((r = check_ctx(c))) return r;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUBFF_OK;
}

}  // extern "C"
