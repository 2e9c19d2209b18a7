// cubff_eval_tapes_noalloc: cubff_eval_tapes without any device allocation.
//
// cubff_eval_tapes stages the caller's tapes and the per-tape results in
// device buffers it allocates on every call.  This variant stages them in
// `scratch`, a file-scope device array of fixed size that every GPU holds
// from the library's load until the process ends, so it can serve callers
// that must not call hipMalloc (tight loops, harnesses that forbid it).
// The array's name is a host-side stand-in in host code; its device
// address is looked up once per device with hipGetSymbolAddress and only
// that address is passed to kernels and copies.  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/cubff_hip.h"
#include "device_common.h"
#include "kernels.h"
#include "languages.h"

using namespace cubff;

// Layout: the tapes (128 bytes each), then ops and steps (4 bytes each) of
// up to CUBFF_NOALLOC_MAX_TAPES tapes.
static constexpr uint64_t kMaxTapes = CUBFF_NOALLOC_MAX_TAPES;
static constexpr uint64_t kOpsOffset = kMaxTapes * kTape2;
static constexpr uint64_t kStepsOffset = kOpsOffset + kMaxTapes * 4;
static constexpr uint64_t kScratchBytes = kStepsOffset + kMaxTapes * 4;
static_assert(kOpsOffset % 256 == 0 && kStepsOffset % 256 == 0, "aligned sections");

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t *g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

int fail(int code, const std::string &msg) { return report_error(code, msg); }

// The device address of this device's `scratch` (looked up on the first
// call on the device, with the device current).
int scratch_of(int device, uint8_t **out) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CUBFF_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                  hipGetErrorString(e));
    g_ptr[device] = static_cast<uint8_t *>(p);
  }
  *out = g_ptr[device];
  return CUBFF_OK;
}

}  // namespace

extern "C" {

int cubff_eval_tapes_noalloc(int lang, int device, uint8_t *tapes, uint64_t count,
                             uint64_t steps, uint32_t *ops, uint32_t *steps_out) {
  if (lang < 0 || lang >= kNumLangs) return fail(CUBFF_EINVAL, "unknown language id");
  if (count && !tapes) return fail(CUBFF_EINVAL, "null tapes");
  if (steps > 0xffffffffULL) return fail(CUBFF_EINVAL, "steps must fit in 32 bits");
  if (count > kMaxTapes)
    return fail(CUBFF_EINVAL, "count exceeds the no-alloc scratch of " +
                                  std::to_string(kMaxTapes) + " tapes");
  if (device < 0 || device >= kMaxDevices) return fail(CUBFF_EINVAL, "device out of range");
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess)
    return fail(CUBFF_EHIP, std::string("hipSetDevice failed: ") + hipGetErrorString(e));
  uint8_t *base = nullptr;
  int r;
  if ((r = scratch_of(device, &base))) return r;
  if (count == 0) return CUBFF_OK;
  uint8_t *d_t = base;
  uint32_t *d_o = reinterpret_cast<uint32_t *>(base + kOpsOffset);
  uint32_t *d_s = reinterpret_cast<uint32_t *>(base + kStepsOffset);
  e = hipMemcpy(d_t, tapes, count * kTape2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_eval_tapes(nullptr, lang, d_t, count, (uint32_t)steps, d_o, d_s);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(tapes, d_t, count * kTape2, hipMemcpyDeviceToHost);
  if (e == hipSuccess && ops) e = hipMemcpy(ops, d_o, count * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && steps_out) e = hipMemcpy(steps_out, d_s, count * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess)
    return fail(CUBFF_EHIP, std::string("eval_tapes_noalloc failed: ") + hipGetErrorString(e));
  return CUBFF_OK;
}

}  // extern "C"
