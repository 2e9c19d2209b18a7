// Learned scheduler: MLP inference over the staged tapes, 8 length buckets
// (longest first), online training on recorded ops.
#include <hip/hip_runtime.h>

#include <vector>

#include "sched.h"

namespace cubff {

// Buckets are quantiles of the current epoch's predictions, longest first:
// bucket b holds the slots ranked in (F[b-1], F[b]] of the batch, F = 0.5%,
// 2%, 5%, 10%, 20%, 40%, 70%, 100%.  Ranking is what the interpreter needs;
// quantiles keep the top buckets small whatever the model's calibration.
constexpr int kPredBins = 1024;
constexpr float kPredScale = kPredBins / 16.0f;  // predictions in [0, 16)

__device__ __forceinline__ uint32_t pred_bin(float y) {
  const float b = y * kPredScale;
  return b <= 0.f ? 0u : (b >= kPredBins - 1 ? kPredBins - 1 : (uint32_t)b);
}

__global__ void k_pred_hist(const float *__restrict__ y, uint64_t count,
                            uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[kPredBins];
  for (int i = threadIdx.x; i < kPredBins; i += blockDim.x) h[i] = 0;
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&h[pred_bin(y[i])], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < kPredBins; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// thr[b] = lowest prediction bin of bucket b: the largest bin whose suffix
// count (predictions in that bin or above) reaches F[b] of the batch.  One
// block of kPredBins threads: a suffix scan in LDS, then each bin checks the
// seven crossings.
__global__ __launch_bounds__(kPredBins) void k_pred_thresholds(const uint32_t *__restrict__ hist,
                                                               uint64_t count,
                                                               uint32_t *__restrict__ thr) {
  __shared__ uint32_t suf[kPredBins + 1];
  const int bin = threadIdx.x;
  suf[bin] = hist[bin];
  if (bin == 0) suf[kPredBins] = 0;
  __syncthreads();
  for (int o = 1; o < kPredBins; o <<= 1) {  // suf[bin] = sum of hist[bin..]
    const uint32_t v = bin + o < kPredBins ? suf[bin + o] : 0u;
    __syncthreads();
    suf[bin] += v;
    __syncthreads();
  }
  constexpr double F[kBuckets - 1] = {0.005, 0.02, 0.05, 0.1, 0.2, 0.4, 0.7};
#pragma unroll
  for (int j = 0; j < kBuckets - 1; j++) {
    const double need = F[j] * (double)count;
    const bool here = (double)suf[bin] >= need && (double)suf[bin + 1] < need;
    if (here) thr[j] = (uint32_t)bin;
    if (bin == 0 && !(need > 0.0)) thr[j] = 0;  // empty batch
  }
}

__device__ __forceinline__ uint32_t bucket_of(float y, const uint32_t *__restrict__ thr) {
  const uint32_t bin = pred_bin(y);
  uint32_t k = 0;
#pragma unroll
  for (int i = 0; i < kBuckets - 1; i++) k += bin < thr[i];
  return k;
}

// Bucketing without contended atomics: each block of 256 slots counts its
// slots per bucket (blk[block][bucket]), one block per bucket turns the
// counts into offsets within the bucket, and each block scatters its slots
// to (earlier buckets' totals + its offset + rank).
constexpr int kSchedBlock = 256;

__global__ __launch_bounds__(kSchedBlock) void k_sched_count(
    const uint32_t *__restrict__ list, uint64_t count, const float *__restrict__ y,
    const uint32_t *__restrict__ thr, uint32_t *__restrict__ bucket, uint32_t *__restrict__ blk) {
  __shared__ uint32_t local[kBuckets];
  if (threadIdx.x < kBuckets) local[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * kSchedBlock + threadIdx.x;
  if (i < count) {
    const uint32_t g = list[i];
    const uint32_t b = bucket_of(y[g], thr);
    bucket[g] = b;
    atomicAdd(&local[b], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kBuckets) blk[(uint64_t)blockIdx.x * kBuckets + threadIdx.x] = local[threadIdx.x];
}

// blk[block][b] <- position of that block's first bucket-b slot within
// bucket b; tot[b] <- slots in bucket b.  One block of 1024 threads per
// bucket.
__global__ __launch_bounds__(1024) void k_sched_offsets(uint32_t *__restrict__ blk,
                                                        uint32_t nblocks,
                                                        uint32_t *__restrict__ tot) {
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x, b = blockIdx.x;
  const uint32_t per = (nblocks + 1023) / 1024;
  const uint32_t lo = min(nblocks, tid * per), hi = min(nblocks, lo + per);
  uint32_t c = 0;
  for (uint32_t k = lo; k < hi; k++) c += blk[(uint64_t)k * kBuckets + b];
  part[tid] = c;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t v = tid >= o ? part[tid - o] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = part[tid] - c;
  for (uint32_t k = lo; k < hi; k++) {
    const uint32_t n = blk[(uint64_t)k * kBuckets + b];
    blk[(uint64_t)k * kBuckets + b] = run;
    run += n;
  }
  if (tid == 1023) tot[b] = part[1023];
}

__global__ __launch_bounds__(kSchedBlock) void k_sched_scatter(
    const uint32_t *__restrict__ list, uint64_t count, const uint32_t *__restrict__ bucket,
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ tot,
    uint32_t *__restrict__ order) {
  __shared__ uint32_t local[kBuckets];
  if (threadIdx.x < kBuckets) local[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * kSchedBlock + threadIdx.x;
  if (i < count) {
    const uint32_t g = list[i];
    const uint32_t b = bucket[g];
    const uint32_t rank = atomicAdd(&local[b], 1u);
    uint32_t base = 0;  // slots in the buckets before b
    for (uint32_t j = 0; j < b; j++) base += tot[j];
    order[base + blk[(uint64_t)blockIdx.x * kBuckets + b] + rank] = g;
  }
}

__global__ void k_sched_fill(uint32_t *__restrict__ p, uint32_t v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

namespace {

int sched_alloc(SchedState &s, const BatchDev &d, size_t num_lists) {
  // Training sample: kTrainRows slots spread evenly over the batch.
  s.train_rows_n = d.total_slots < kTrainRows ? d.total_slots / 64 * 64 : kTrainRows;
  if (s.train_rows_n == 0) s.train_rows_n = d.total_slots;
  // Activation buffers are only needed for the training rows; inference is
  // fused (mlp_predict).
  if (mlp_alloc(s.mlp, s.train_rows_n) != hipSuccess) return -1;
  if (hipMalloc(&s.pred, d.total_slots * 4) != hipSuccess) return -1;
  if (hipMalloc(&s.bucket, d.total_slots * 4) != hipSuccess) return -1;
  if (hipMalloc(&s.cnt, (num_lists ? num_lists : 1) * 2 * kBuckets * 4) != hipSuccess) return -1;
  if (hipMalloc(&s.blk, (d.total_slots / kSchedBlock + 2) * kBuckets * 4) != hipSuccess) return -1;
  if (hipMalloc(&s.hist, kPredBins * 4) != hipSuccess) return -1;
  if (hipMalloc(&s.thr, kBuckets * 4) != hipSuccess) return -1;
  std::vector<uint32_t> rows(s.train_rows_n);
  for (uint64_t i = 0; i < s.train_rows_n; i++)
    rows[i] = (uint32_t)((i * d.total_slots) / s.train_rows_n);
  if (hipMalloc(&s.train_rows, s.train_rows_n * 4) != hipSuccess) return -1;
  if (hipMemcpy(s.train_rows, rows.data(), s.train_rows_n * 4, hipMemcpyHostToDevice) !=
      hipSuccess)
    return -1;
  if (hipStreamCreateWithFlags(&s.train_stream, hipStreamNonBlocking) != hipSuccess) return -1;
  if (hipEventCreateWithFlags(&s.taken, hipEventDisableTiming) != hipSuccess) return -1;
  if (hipEventCreateWithFlags(&s.trained, hipEventDisableTiming) != hipSuccess) return -1;
  if (hipMalloc(&s.loss0, 4) != hipSuccess) return -1;
  hipStream_t st = nullptr;
  if (mlp_init(s.mlp, 0x5eed, st) != hipSuccess) return -1;
  if (mlp_fold(s.mlp, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return -1;
  s.ready = true;
  return 0;
}

}  // namespace

int sched_ensure(SchedState &s, const BatchDev &d, size_t num_lists) {
  if (s.ready) return 0;
  if (sched_alloc(s, d, num_lists) == 0) return 0;
  sched_free(s);  // whatever the failed attempt allocated
  return -1;
}

void sched_free(SchedState &s) {
  if (s.train_stream) (void)hipStreamSynchronize(s.train_stream);
  mlp_free(s.mlp);
  if (s.loss0) (void)hipFree(s.loss0);
  if (s.taken) (void)hipEventDestroy(s.taken);
  if (s.trained) (void)hipEventDestroy(s.trained);
  if (s.train_stream) (void)hipStreamDestroy(s.train_stream);
  if (s.bucket) (void)hipFree(s.bucket);
  if (s.cnt) (void)hipFree(s.cnt);
  if (s.blk) (void)hipFree(s.blk);
  if (s.hist) (void)hipFree(s.hist);
  if (s.thr) (void)hipFree(s.thr);
  if (s.pred) (void)hipFree(s.pred);
  if (s.train_rows) (void)hipFree(s.train_rows);
  s = SchedState{};
}

hipError_t sched_order(hipStream_t st, SchedState &s, const BatchDev &d, const uint32_t *halves,
                       const uint32_t *plain, uint32_t *order, const SchedList *lists,
                       size_t num_lists) {
  if (!s.ready) return hipErrorNotReady;
  hipError_t e;
  if ((e = mlp_predict_live(s.mlp, d.stage, d.total_slots, s.pred, st, halves))) return e;
  hipLaunchKernelGGL(k_sched_fill, dim3(1), dim3(256), 0, st, s.cnt, 0u,
                     (uint64_t)num_lists * 2 * kBuckets);
  hipLaunchKernelGGL(k_sched_fill, dim3(kPredBins / 256), dim3(256), 0, st, s.hist, 0u,
                     (uint64_t)kPredBins);
  hipLaunchKernelGGL(k_pred_hist, dim3(256), dim3(256), 0, st, s.pred, d.total_slots, s.hist);
  hipLaunchKernelGGL(k_pred_thresholds, dim3(1), dim3(kPredBins), 0, st, s.hist, d.total_slots,
                     s.thr);
  for (size_t li = 0; li < num_lists; li++) {
    const uint64_t n = lists[li].count;
    if (n == 0) continue;
    const unsigned g = (unsigned)((n + kSchedBlock - 1) / kSchedBlock);
    uint32_t *cnt = s.cnt + li * 2 * kBuckets;
    hipLaunchKernelGGL(k_sched_count, dim3(g), dim3(kSchedBlock), 0, st, plain + lists[li].begin,
                       n, s.pred, s.thr, s.bucket, s.blk);
    hipLaunchKernelGGL(k_sched_offsets, dim3(kBuckets), dim3(1024), 0, st, s.blk, g, cnt);
    hipLaunchKernelGGL(k_sched_scatter, dim3(g), dim3(kSchedBlock), 0, st,
                       plain + lists[li].begin, n, s.bucket, s.blk, cnt,
                       order + lists[li].begin);
  }
  return hipGetLastError();
}

namespace {

// Features and targets of the training sample, from the epoch just run.
hipError_t take_sample(hipStream_t st, SchedState &s, const BatchDev &d) {
  hipError_t e;
  const uint64_t n = s.train_rows_n;
  if ((e = mlp_features_rows(s.mlp, d.stage, s.train_rows, n, st))) return e;
  // Targets: the commands each slot's interpreter actually ran (the cost
  // the schedule balances), not the reference's op count.
  return mlp_targets_from_ops(s.mlp, d.pair_work, s.train_rows, n, st);
}

// `steps` Adam steps on the taken sample; the first step's loss to loss0.
hipError_t train_steps(hipStream_t st, SchedState &s, int steps, float lr) {
  hipError_t e;
  const uint64_t n = s.train_rows_n;
  for (int i = 0; i < steps; i++) {
    if ((e = mlp_forward(s.mlp, n, st))) return e;
    if ((e = mlp_backward(s.mlp, n, st))) return e;
    if (i == 0 && (e = hipMemcpyAsync(s.loss0, s.mlp.loss, 4, hipMemcpyDeviceToDevice, st)))
      return e;
    if ((e = mlp_adam(s.mlp, lr, st))) return e;
  }
  return hipSuccess;
}

}  // namespace

hipError_t sched_poll(hipStream_t st, SchedState &s, bool wait) {
  if (!s.pending) return hipSuccess;
  if (!wait) {
    const hipError_t q = hipEventQuery(s.trained);
    if (q == hipErrorNotReady) return hipSuccess;
    if (q != hipSuccess) return q;
  }
  hipError_t e;
  if ((e = hipStreamWaitEvent(st, s.trained, 0))) return e;
  if ((e = mlp_publish(s.mlp, st))) return e;
  s.pending = false;
  s.trainings++;
  return hipSuccess;
}

float sched_last_loss(SchedState &s) {
  if (!s.ready || s.trainings == 0) return s.last_loss;
  float v = 0.f;
  if (s.train_stream && hipStreamSynchronize(s.train_stream) == hipSuccess &&
      hipMemcpy(&v, s.loss0, 4, hipMemcpyDeviceToHost) == hipSuccess)
    s.last_loss = v;
  return s.last_loss;
}

hipError_t sched_train_async(hipStream_t st, SchedState &s, const BatchDev &d, int steps,
                             float lr) {
  if (!s.ready) return hipErrorNotReady;
  hipError_t e;
  // The sample buffers belong to a pending training until it is done.
  if ((e = sched_poll(st, s, true))) return e;
  if ((e = take_sample(st, s, d))) return e;
  if ((e = hipEventRecord(s.taken, st))) return e;
  if ((e = hipStreamWaitEvent(s.train_stream, s.taken, 0))) return e;
  if ((e = train_steps(s.train_stream, s, steps, lr))) return e;
  if ((e = mlp_fold_next(s.mlp, s.train_stream))) return e;
  if ((e = hipEventRecord(s.trained, s.train_stream))) return e;
  s.pending = true;
  return hipSuccess;
}

hipError_t sched_train(hipStream_t st, SchedState &s, const BatchDev &d, int steps, float lr,
                       float *first_loss) {
  if (!s.ready) return hipErrorNotReady;
  hipError_t e;
  if ((e = sched_poll(st, s, true))) return e;
  if ((e = take_sample(st, s, d))) return e;
  // The first round starts from random weights: train it longer.
  if (s.trainings == 0) steps *= 8;
  if ((e = train_steps(st, s, steps, lr))) return e;
  // Inference (possibly replayed from a captured graph) reads the live
  // weights: refresh them now, outside any capture.
  if ((e = mlp_fold(s.mlp, st))) return e;
  float loss = 0.f;
  if ((e = hipMemcpyAsync(&loss, s.loss0, 4, hipMemcpyDeviceToHost, st))) return e;
  if ((e = hipStreamSynchronize(st))) return e;
  s.trainings++;
  s.last_loss = loss;
  if (first_loss) *first_loss = loss;
  return hipSuccess;
}

}  // namespace cubff
