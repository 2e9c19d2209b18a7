// MLP cost model of the learned scheduler.
//
//   x   = byte histogram / 128     (256, bf16)
//   p   = x E                      (E: 256 x 256, pooled byte embedding)
//   a1  = relu(p W1 + b1)          (256)
//   a2  = relu(a1 W2 + b2)         (256)
//   y   = a2 . w3 + b3             predicted log2 length
//
// bf16 weights and activations, fp32 accumulation on MFMA; fp32 master
// weights with Adam.  All GEMMs are hand-written (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubff {

constexpr int kMlpIn = 256, kMlpHid = 256;
// Offsets of the fp32 parameter vector.
constexpr uint64_t kOffE = 0;
constexpr uint64_t kOffW1 = kOffE + (uint64_t)kMlpIn * kMlpHid;
constexpr uint64_t kOffB1 = kOffW1 + (uint64_t)kMlpHid * kMlpHid;
constexpr uint64_t kOffW2 = kOffB1 + kMlpHid;
constexpr uint64_t kOffB2 = kOffW2 + (uint64_t)kMlpHid * kMlpHid;
constexpr uint64_t kOffW3 = kOffB2 + kMlpHid;
constexpr uint64_t kOffB3 = kOffW3 + kMlpHid;
constexpr uint64_t kNumParams = kOffB3 + 1;

// Weights of the fused inference kernel: folded E W1 / 128 and W2, both
// transposed bf16 [256][256], and vec = b1, b2, w3 (256 each), b3.
constexpr int kInfVec = 3 * kMlpHid + 1;
struct MlpInfer {
  uint16_t *ew1t = nullptr, *w2t = nullptr;
  float *vec = nullptr;
};

struct Mlp {
  uint64_t capacity = 0;  // rows the activation buffers hold (multiple of 128)
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  // bf16 weights: W (row-major [in][out]) and WT ([out][in]) per layer.
  uint16_t *e_w = nullptr, *e_wt = nullptr, *w1 = nullptr, *w1t = nullptr;
  uint16_t *w2 = nullptr, *w2t = nullptr;
  // Inference weights: `live` is what mlp_predict reads; `next` receives a
  // fold made on another stream, published into `live` later.
  MlpInfer live, next;
  bool folded_valid = false;
  uint16_t *x = nullptr;               // [capacity, 256] features
  uint16_t *p = nullptr, *a1 = nullptr, *a2 = nullptr;  // [capacity, 256]
  uint16_t *d2 = nullptr, *d1 = nullptr, *dp = nullptr; // [capacity, 256]
  float *y = nullptr, *target = nullptr;                 // [capacity]
  float *loss = nullptr;               // device scalar
  unsigned int *claim = nullptr;       // inference: passes claimed so far
  uint64_t adam_t = 0;
};

hipError_t mlp_alloc(Mlp &m, uint64_t capacity);
void mlp_free(Mlp &m);
// Deterministic initialisation from a seed (SplitMix64 uniform, scaled per
// layer); uploads fp32 and bf16 copies.
hipError_t mlp_init(Mlp &m, uint64_t seed, hipStream_t st);
hipError_t mlp_set_params(Mlp &m, const float *host, hipStream_t st);
hipError_t mlp_get_params(const Mlp &m, float *host, hipStream_t st);
// Features of `count` 128-byte tapes (device, 128 bytes apart) into m.x.
hipError_t mlp_features(Mlp &m, const uint8_t *tapes, uint64_t count, hipStream_t st);
// Same for selected slots of a stage buffer: row i <- tape rows[i].
hipError_t mlp_features_rows(Mlp &m, const uint8_t *stage, const uint32_t *rows,
                             uint64_t count, hipStream_t st);
// Forward over m.x[0..count) into m.y (keeps p, a1, a2 for backward).
hipError_t mlp_forward(Mlp &m, uint64_t count, hipStream_t st);
// Folds the first two linear maps (E W1) and copies W2, the biases and w3
// for inference, into `live`; recomputed after every update.
hipError_t mlp_fold(Mlp &m, hipStream_t st);
// The same into `next` (training on its own stream) and next -> live.
hipError_t mlp_fold_next(Mlp &m, hipStream_t st);
hipError_t mlp_publish(Mlp &m, hipStream_t st);
// Fused inference of `count` tapes (device, 128 bytes apart) straight into
// y[0..count): features, both layers and the output in one kernel, with the
// folded first layer.  Needs no activation buffers.
hipError_t mlp_predict(Mlp &m, const uint8_t *tapes, uint64_t count, float *y, hipStream_t st,
                       const uint32_t *halves = nullptr);
// mlp_predict with the `live` weights as they are (never folds).  With
// `halves` (2 * count entries), tape r is the two 64-byte pieces at
// tapes + 64 * halves[2r] and tapes + 64 * halves[2r + 1] instead.
hipError_t mlp_predict_live(const Mlp &m, const uint8_t *tapes, uint64_t count, float *y,
                            hipStream_t st, const uint32_t *halves = nullptr);
// Loss 0.5 * mean((y - target)^2) and all gradients into m.grads for
// m.x / m.target rows [0..count); *m.loss receives the loss.
hipError_t mlp_backward(Mlp &m, uint64_t count, hipStream_t st);
hipError_t mlp_adam(Mlp &m, float lr, hipStream_t st);
// target[i] = log2(ops[rows[i]] + 1).
hipError_t mlp_targets_from_ops(Mlp &m, const uint32_t *ops, const uint32_t *rows,
                                uint64_t count, hipStream_t st);

}  // namespace cubff
