// C-ABI of the MLP cost model on its own (declared in include/cubff_hip.h):
// used to check its gradients against a CPU reference and to measure what
// training gains.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cubff_hip.h"
#include "device_common.h"
#include "mlp.h"

using namespace cubff;

struct cubff_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  Mlp m;
  uint8_t *tapes = nullptr;
  uint32_t *index = nullptr;  // 2 * cap piece indices
  uint64_t cap = 0;
};

namespace {
thread_local std::string g_err;
int fail(int code, const std::string &s) {
  g_err = s;
  return code;
}
#define MLP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      return fail(CUBFF_EHIP, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
  } while (0)

// Uploads `count` tapes and computes their features into m.x.
int load_tapes(cubff_mlp *h, const uint8_t *tapes, uint64_t count) {
  if (count == 0 || count > h->cap) return fail(CUBFF_EINVAL, "count must be in [1, capacity]");
  MLP_TRY(hipMemcpyAsync(h->tapes, tapes, count * kTape2, hipMemcpyHostToDevice, h->stream));
  MLP_TRY(mlp_features(h->m, h->tapes, count, h->stream));
  return CUBFF_OK;
}
}  // namespace

extern "C" {

const char *cubff_mlp_last_error(void) { return g_err.c_str(); }

uint64_t cubff_mlp_num_params(void) { return kNumParams; }

int cubff_mlp_create(cubff_mlp **out, uint64_t capacity, uint64_t init_seed, int device) {
  if (!out) return fail(CUBFF_EINVAL, "null output pointer");
  *out = nullptr;
  MLP_TRY(hipSetDevice(device));
  cubff_mlp *h = new cubff_mlp();
  h->device = device;
  h->cap = capacity;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = mlp_alloc(h->m, capacity);
  if (e == hipSuccess) e = hipMalloc(&h->tapes, (capacity ? capacity : 1) * kTape2);
  if (e == hipSuccess) e = hipMalloc(&h->index, (capacity ? capacity : 1) * 2 * 4);
  if (e == hipSuccess) e = mlp_init(h->m, init_seed, h->stream);
  if (e != hipSuccess) {
    cubff_mlp_destroy(h);
    return fail(CUBFF_EHIP, std::string("mlp create failed: ") + hipGetErrorString(e));
  }
  *out = h;
  return CUBFF_OK;
}

int cubff_mlp_destroy(cubff_mlp *h) {
  if (!h) return CUBFF_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  mlp_free(h->m);
  if (h->tapes) (void)hipFree(h->tapes);
  if (h->index) (void)hipFree(h->index);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CUBFF_OK;
}

int cubff_mlp_set_params(cubff_mlp *h, const float *params) {
  if (!h || !params) return fail(CUBFF_EINVAL, "null argument");
  MLP_TRY(hipSetDevice(h->device));
  MLP_TRY(mlp_set_params(h->m, params, h->stream));
  return CUBFF_OK;
}

int cubff_mlp_get_params(cubff_mlp *h, float *params) {
  if (!h || !params) return fail(CUBFF_EINVAL, "null argument");
  MLP_TRY(hipSetDevice(h->device));
  MLP_TRY(mlp_get_params(h->m, params, h->stream));
  return CUBFF_OK;
}

int cubff_mlp_forward(cubff_mlp *h, const uint8_t *tapes, uint64_t count, float *pred) {
  if (!h || !tapes || !pred) return fail(CUBFF_EINVAL, "null argument");
  MLP_TRY(hipSetDevice(h->device));
  int r;
  if ((r = load_tapes(h, tapes, count))) return r;
  MLP_TRY(mlp_forward(h->m, count, h->stream));
  MLP_TRY(hipMemcpyAsync(pred, h->m.y, count * 4, hipMemcpyDeviceToHost, h->stream));
  MLP_TRY(hipStreamSynchronize(h->stream));
  return CUBFF_OK;
}

int cubff_mlp_predict(cubff_mlp *h, const uint8_t *tapes, uint64_t count, float *pred) {
  if (!h || !tapes || !pred) return fail(CUBFF_EINVAL, "null argument");
  if (count == 0 || count > h->cap) return fail(CUBFF_EINVAL, "count must be in [1, capacity]");
  MLP_TRY(hipSetDevice(h->device));
  MLP_TRY(hipMemcpyAsync(h->tapes, tapes, count * kTape2, hipMemcpyHostToDevice, h->stream));
  MLP_TRY(mlp_predict(h->m, h->tapes, count, h->m.y, h->stream));
  MLP_TRY(hipMemcpyAsync(pred, h->m.y, count * 4, hipMemcpyDeviceToHost, h->stream));
  MLP_TRY(hipStreamSynchronize(h->stream));
  return CUBFF_OK;
}

int cubff_mlp_predict_pieces(cubff_mlp *h, const uint8_t *pieces, uint64_t num_pieces,
                             const uint32_t *index, uint64_t count, float *pred) {
  if (!h || !pieces || !index || !pred) return fail(CUBFF_EINVAL, "null argument");
  if (count == 0 || count > h->cap) return fail(CUBFF_EINVAL, "count must be in [1, capacity]");
  if (num_pieces == 0 || num_pieces > 2 * h->cap)
    return fail(CUBFF_EINVAL, "num_pieces must be in [1, 2 * capacity]");
  for (uint64_t i = 0; i < 2 * count; i++)
    if (index[i] >= num_pieces) return fail(CUBFF_EINVAL, "piece index out of range");
  MLP_TRY(hipSetDevice(h->device));
  MLP_TRY(hipMemcpyAsync(h->tapes, pieces, num_pieces * kTape, hipMemcpyHostToDevice, h->stream));
  MLP_TRY(hipMemcpyAsync(h->index, index, 2 * count * 4, hipMemcpyHostToDevice, h->stream));
  MLP_TRY(mlp_predict(h->m, h->tapes, count, h->m.y, h->stream, h->index));
  MLP_TRY(hipMemcpyAsync(pred, h->m.y, count * 4, hipMemcpyDeviceToHost, h->stream));
  MLP_TRY(hipStreamSynchronize(h->stream));
  return CUBFF_OK;
}

int cubff_mlp_gradients(cubff_mlp *h, const uint8_t *tapes, const float *targets,
                        uint64_t count, float *loss, float *grads) {
  if (!h || !tapes || !targets) return fail(CUBFF_EINVAL, "null argument");
  MLP_TRY(hipSetDevice(h->device));
  int r;
  if ((r = load_tapes(h, tapes, count))) return r;
  MLP_TRY(hipMemcpyAsync(h->m.target, targets, count * 4, hipMemcpyHostToDevice, h->stream));
  MLP_TRY(mlp_forward(h->m, count, h->stream));
  MLP_TRY(mlp_backward(h->m, count, h->stream));
  if (loss) MLP_TRY(hipMemcpyAsync(loss, h->m.loss, 4, hipMemcpyDeviceToHost, h->stream));
  if (grads)
    MLP_TRY(hipMemcpyAsync(grads, h->m.grads, kNumParams * 4, hipMemcpyDeviceToHost, h->stream));
  MLP_TRY(hipStreamSynchronize(h->stream));
  return CUBFF_OK;
}

int cubff_mlp_train(cubff_mlp *h, const uint8_t *tapes, const float *targets, uint64_t count,
                    uint32_t steps, float lr, float *losses) {
  if (!h || !tapes || !targets) return fail(CUBFF_EINVAL, "null argument");
  MLP_TRY(hipSetDevice(h->device));
  int r;
  if ((r = load_tapes(h, tapes, count))) return r;
  MLP_TRY(hipMemcpyAsync(h->m.target, targets, count * 4, hipMemcpyHostToDevice, h->stream));
  for (uint32_t i = 0; i < steps; i++) {
    MLP_TRY(mlp_forward(h->m, count, h->stream));
    MLP_TRY(mlp_backward(h->m, count, h->stream));
    if (losses) MLP_TRY(hipMemcpyAsync(losses + i, h->m.loss, 4, hipMemcpyDeviceToHost, h->stream));
    MLP_TRY(mlp_adam(h->m, lr, h->stream));
  }
  MLP_TRY(hipStreamSynchronize(h->stream));
  return CUBFF_OK;
}

}  // extern "C"
