// Per-wavefront clock probe of the BFF interpreters on one GPU.
//
//   interp_probe <tapes file> [budget]
//
// The tapes file holds 128-byte bff_noheads interaction tapes back to back
// (bench.py --interp-probe writes the pairs of a real epoch).  Each kernel
// evaluates every tape, one lane per tape and 256 lanes per workgroup as in
// the epoch kernels, and records per lane the commands it executed and per
// wavefront the shader clocks (s_memtime) its evaluation took.  Printed as
// one JSON object per interpreter:
//   cycles_per_command  least-squares slope of wavefront cycles over the
//                       largest executed-command count of its lanes,
//   lone_lane           the same for tapes 0, 64, 128, ... alone in their
//                       wavefronts (every other lane given an empty tape):
//                       the cost of one lane's command, nothing else issuing,
//   kernel_ms           the launch (best of three).
// Interpreters: the word-table interpreter (bff_lean.h) budgeted as in the
// deferring epoch kernel, complete without its cycle skip (256 lanes: the
// LDS holds no anchor copies) and complete with it as the resolver runs it
// (128 lanes), and the byte-tape interpreters it replaced there (bff_fast.h
// budgeted, languages.h complete with its cycle skip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bff_fast.h"
#include "bff_lean.h"
#include "kernels.h"
#include "languages.h"
#include "tape_io.h"

using namespace cubff;

namespace {

constexpr int kLanes = 256;

enum Kind { kLeanBudget, kLeanFull, kFastBudget, kLangFull, kLeanCycle };

__global__ __launch_bounds__(kLanes) void k_probe_lean(const uint8_t *tapes, uint32_t n,
                                                       uint32_t budget, int full,
                                                       uint32_t *work_o, uint64_t *cyc_o) {
  using F = BffLean<kMapAscii, false>;
  using W = LeanLds<kLanes>;
  __shared__ __align__(16) uint8_t lds[W::kRegionBytes];
  __shared__ uint8_t lut[256];
  fill_lean_lut<kBffNoheads>(lut);
  __syncthreads();
  const int lane = threadIdx.x;
  const uint32_t k = blockIdx.x * kLanes + lane;
  const W w = W::at(lds, lane);
  const LdsByteLut blut{(const lds_u8 *)lut};
  if (k < n) {
    load_raw_row(w, 0, tapes + (uint64_t)k * kTape2);
    load_raw_row(w, 1, tapes + (uint64_t)k * kTape2 + kTape);
  }
  uint32_t steps = 0, work = 0;
  // With 256 lanes the LDS holds the word tables and no anchor copies, so
  // `full` runs the complete run without the cycle skip.
  const uint64_t c0 = __builtin_amdgcn_s_memtime();
  if (k < n) F::evaluate(w, blut, kEvalSteps, &steps, &work, full ? 0xffffffffu : budget);
  const uint64_t c1 = __builtin_amdgcn_s_memtime();
  if (k < n) work_o[k] = steps == kOverBudget ? budget : work;
  if ((lane & 63) == 0) cyc_o[k / 64] = c1 - c0;
}

// The complete run with the full-state cycle skip, as the resolver runs it
// (128 lanes per workgroup: the LDS then also holds the anchors).
__global__ __launch_bounds__(kLanes / 2) void k_probe_lean_cycle(const uint8_t *tapes, uint32_t n,
                                                                uint32_t *work_o, uint64_t *cyc_o) {
  constexpr int kL = kLanes / 2;
  using F = BffLean<kMapAscii, false>;
  using W = LeanLds<kL>;
  using A = LeanAnchorLds<kL>;
  __shared__ __align__(16) uint8_t lds[W::kRegionBytes + A::kRegionBytes];
  __shared__ uint8_t lut[256];
  fill_lean_lut<kBffNoheads>(lut);
  __syncthreads();
  const int lane = threadIdx.x;
  const uint32_t k = blockIdx.x * kL + lane;
  const W w = W::at(lds, lane);
  const A anchor = A::at(lds + W::kRegionBytes, lane);
  const LdsByteLut blut{(const lds_u8 *)lut};
  if (k < n) {
    load_raw_row(w, 0, tapes + (uint64_t)k * kTape2);
    load_raw_row(w, 1, tapes + (uint64_t)k * kTape2 + kTape);
  }
  uint32_t steps = 0, work = 0;
  const uint64_t c0 = __builtin_amdgcn_s_memtime();
  if (k < n) F::evaluate_full(w, anchor, blut, kEvalSteps, &steps, &work);
  const uint64_t c1 = __builtin_amdgcn_s_memtime();
  if (k < n) work_o[k] = work;
  if ((lane & 63) == 0) cyc_o[k / 64] = c1 - c0;
}

__global__ __launch_bounds__(kLanes) void k_probe_byte(const uint8_t *tapes, uint32_t n,
                                                       uint32_t budget, int full,
                                                       uint32_t *work_o, uint64_t *cyc_o) {
  using L = typename LangOf<kBffNoheads>::T;
  using Tape = LinTape<kLanes>;
  __shared__ __align__(16) uint8_t lds[2 * Tape::kRegionBytes];
  __shared__ uint8_t lut[256];
  fill_byte_lut<kMapAscii>(lut);
  __syncthreads();
  const int lane = threadIdx.x;
  const uint32_t k = blockIdx.x * kLanes + lane;
  const auto t = Tape::at(lds, lane);
  const auto second = Tape::at(lds + Tape::kRegionBytes, lane);
  if (k < n) {
    load_row(t, 0, tapes + (uint64_t)k * kTape2);
    load_row(t, 1, tapes + (uint64_t)k * kTape2 + kTape);
  }
  uint32_t steps = 0, work = 0;
  const uint64_t c0 = __builtin_amdgcn_s_memtime();
  if (k < n) {
    if (full)
      L::evaluate(t, second, kEvalSteps, &steps, &work);
    else
      BffFast<kMapAscii, false>::evaluate(t, second, LdsByteLut{(const lds_u8 *)lut}, kEvalSteps,
                                          &steps, &work, budget);
  }
  const uint64_t c1 = __builtin_amdgcn_s_memtime();
  if (k < n) work_o[k] = steps == kOverBudget ? budget : work;
  if ((lane & 63) == 0) cyc_o[k / 64] = c1 - c0;
}

#define CK(x)                                                                   \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                   \
      exit(1);                                                                  \
    }                                                                           \
  } while (0)

struct Fit {
  double slope, kernel_ms;
  uint32_t max_work;
};

Fit run(int kind, const uint8_t *d_t, uint32_t n, uint32_t budget, uint32_t *d_w, uint64_t *d_c) {
  const uint32_t blocks = (n + kLanes - 1) / kLanes, waves = blocks * 4;
  const uint32_t half_blocks = (n + kLanes / 2 - 1) / (kLanes / 2);
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  const int full = kind == kLeanFull || kind == kLangFull;
  float best = 1e30f;
  for (int rep = 0; rep < 3; rep++) {
    CK(hipEventRecord(a, 0));
    if (kind == kLeanCycle)
      hipLaunchKernelGGL(k_probe_lean_cycle, dim3(half_blocks), dim3(kLanes / 2), 0, 0, d_t, n,
                         d_w, d_c);
    else if (kind == kLeanBudget || kind == kLeanFull)
      hipLaunchKernelGGL(k_probe_lean, dim3(blocks), dim3(kLanes), 0, 0, d_t, n, budget, full,
                         d_w, d_c);
    else
      hipLaunchKernelGGL(k_probe_byte, dim3(blocks), dim3(kLanes), 0, 0, d_t, n, budget, full,
                         d_w, d_c);
    CK(hipGetLastError());
    CK(hipEventRecord(b, 0));
    CK(hipEventSynchronize(b));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms);
  }
  std::vector<uint32_t> w(n);
  std::vector<uint64_t> c(waves);
  CK(hipMemcpy(w.data(), d_w, n * 4, hipMemcpyDeviceToHost));
  CK(hipMemcpy(c.data(), d_c, waves * 8, hipMemcpyDeviceToHost));
  double sx = 0, sy = 0, sxx = 0, sxy = 0;
  uint32_t mx = 0;
  int m = 0;
  for (uint32_t v = 0; v < waves; v++) {
    uint32_t mw = 0;
    for (uint32_t l = 0; l < 64; l++)
      if (v * 64 + l < n) mw = std::max(mw, w[v * 64 + l]);
    if (v * 64 >= n) continue;
    mx = std::max(mx, mw);
    const double x = mw, y = (double)c[v];
    sx += x;
    sy += y;
    sxx += x * x;
    sxy += x * y;
    m++;
  }
  const double den = m * sxx - sx * sx;
  return Fit{den > 0 ? (m * sxy - sx * sy) / den : 0.0, best, mx};
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: interp_probe <tapes file> [budget]\n");
    return 2;
  }
  const uint32_t budget = argc > 2 ? (uint32_t)atoi(argv[2]) : 768;
  FILE *f = fopen(argv[1], "rb");
  if (!f) {
    perror(argv[1]);
    return 2;
  }
  std::vector<uint8_t> h;
  uint8_t buf[kTape2];
  while (fread(buf, 1, kTape2, f) == kTape2) h.insert(h.end(), buf, buf + kTape2);
  fclose(f);
  const uint32_t n = (uint32_t)(h.size() / kTape2);
  if (n == 0) {
    fprintf(stderr, "no tapes\n");
    return 2;
  }
  // The lone-lane set: tape 64 j at lane 0 of wavefront j, empty tapes (no
  // command: the run ends at once) elsewhere.
  std::vector<uint8_t> lone(h.size(), 0);
  for (uint32_t k = 0; k < n; k += 64)
    std::copy(&h[(size_t)k * kTape2], &h[(size_t)k * kTape2] + kTape2, &lone[(size_t)k * kTape2]);
  uint8_t *d_t, *d_lone;
  uint32_t *d_w;
  uint64_t *d_c;
  CK(hipMalloc(&d_t, h.size()));
  CK(hipMalloc(&d_lone, h.size()));
  CK(hipMalloc(&d_w, (size_t)n * 4));
  CK(hipMalloc(&d_c, (size_t)((n + kLanes - 1) / kLanes) * 4 * 8));
  CK(hipMemcpy(d_t, h.data(), h.size(), hipMemcpyHostToDevice));
  CK(hipMemcpy(d_lone, lone.data(), lone.size(), hipMemcpyHostToDevice));
  static const char *kNames[5] = {"word_table_budgeted", "word_table_complete_no_cycle_skip",
                                  "byte_tape_budgeted", "byte_tape_complete_cycle_skip",
                                  "word_table_complete_cycle_skip"};
  printf("{\"tapes\": %u, \"budget\": %u", n, budget);
  for (int kind = 0; kind < 5; kind++) {
    const Fit all = run(kind, d_t, n, budget, d_w, d_c);
    const Fit one = run(kind, d_lone, n, budget, d_w, d_c);
    printf(", \"%s\": {\"cycles_per_command\": %.1f, \"kernel_ms\": %.4f, \"max_commands\": %u, "
           "\"lone_lane_cycles_per_command\": %.1f, \"lone_lane_kernel_ms\": %.4f}",
           kNames[kind], all.slope, all.kernel_ms, all.max_work, one.slope, one.kernel_ms);
  }
  printf("}\n");
  return 0;
}
