// Step trace of one BFF-family interaction: the state before every step --
// pc, both heads, the 128-byte tape -- in the layout the reference's Python
// interpreter writes for its visualiser (python/bff_interpreter.py
// evaluate_and_save, :327-393): u32 little-endian pc, head0, head1, then
// the tape, 140 bytes per state.  The loop is Evaluate (bff.inc.h:356-389)
// with EvaluateOne (bff.inc.h:289-349, bff8.cu:23-65, bff_selfmove.cu:44-98)
// taken one step at a time, with no skipping.
//
// One wavefront per tape: every lane runs the same steps on the shared LDS
// tape (identical values to identical addresses), so after each step all
// lanes hold the state and each stores two tape bytes of the record.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"
#include "languages.h"

namespace cubff {

namespace {

template <int MAP, int LOGIC>
__device__ __forceinline__ bool is_null(uint32_t v) {
  if constexpr (LOGIC == kLogicStd) return bff_kind<MAP>(v) == kNull;
  else return v == 0;
}

template <int MAP>
__device__ int scan_forward(const uint8_t *tape, int pc) {
  int depth = 1, p = pc + 1;
  for (; p < kTape2 && depth > 0; p++) {
    const int k = bff_kind<MAP>(tape[p]);
    if (k == kLoopEnd) depth--;
    if (k == kLoopStart) depth++;
  }
  p--;
  return depth != 0 ? kTape2 : p;
}

template <int MAP>
__device__ int scan_backward(const uint8_t *tape, int pc) {
  int depth = 1, p = pc - 1;
  for (; p >= 0 && depth > 0; p--) {
    const int k = bff_kind<MAP>(tape[p]);
    if (k == kLoopEnd) depth++;
    if (k == kLoopStart) depth--;
  }
  p++;
  return depth != 0 ? -1 : p;
}

// One EvaluateOne; false for a no-op.
template <int MAP, int LOGIC>
__device__ bool step(uint8_t *tape, int &h0, int &h1, int &pc) {
  const int k = bff_kind<MAP>(tape[pc]);
  switch (k) {
    case kDec0: h0--; return true;
    case kInc0: h0++; return true;
    case kPlus: tape[h0]++; return true;
    case kMinus: tape[h0]--; return true;
    case kCopy10:
      tape[h0] = tape[h1];
      if constexpr (LOGIC == kLogicSelfmove) h1++;
      return true;
    case kLoopStart:
      if constexpr (LOGIC == kLogicBff8) {
        if (!tape[h0]) pc += 8;
      } else if (is_null<MAP, LOGIC>(tape[h0])) {
        pc = scan_forward<MAP>(tape, pc);
      }
      return true;
    case kLoopEnd:
      if constexpr (LOGIC == kLogicBff8) {
        if (tape[h0]) pc -= 8;
      } else if (!is_null<MAP, LOGIC>(tape[h0])) {
        pc = scan_backward<MAP>(tape, pc);
      }
      return true;
    default: break;
  }
  if constexpr (LOGIC != kLogicSelfmove) {
    switch (k) {
      case kDec1: h1--; return true;
      case kInc1: h1++; return true;
      case kCopy01: tape[h1] = tape[h0]; return true;
      default: break;
    }
  }
  return false;
}

}  // namespace

template <int MAP, bool HEADS, int LOGIC>
__global__ __launch_bounds__(64) void k_trace(const uint8_t *__restrict__ tape_in,
                                              uint32_t steps, uint8_t *__restrict__ out,
                                              uint64_t cap, uint64_t *__restrict__ meta) {
  __shared__ uint8_t tape[kTape2];
  const int lane = threadIdx.x;
  tape[2 * lane] = tape_in[2 * lane];
  tape[2 * lane + 1] = tape_in[2 * lane + 1];
  __syncthreads();
  int h0, h1, pc;
  if constexpr (HEADS) {  // InitialState, bff.inc.h:269-275
    h0 = tape[0] % kTape2;
    h1 = tape[1] % kTape2;
    pc = 2;
  } else {
    h0 = kTape2;
    h1 = kTape2;
    pc = 0;
  }
  uint64_t i = 0, nskip = 0, rec = 0;
  for (; i < steps; i++) {
    h0 &= kTape2 - 1;
    h1 &= kTape2 - 1;
    if (rec < cap) {
      uint8_t *o = out + rec * (12 + kTape2);
      if (lane < 3) {
        const uint32_t f = lane == 0 ? (uint32_t)pc : (lane == 1 ? (uint32_t)h0 : (uint32_t)h1);
        *reinterpret_cast<uint32_t *>(o + 4 * lane) = f;
      }
      *reinterpret_cast<uint16_t *>(o + 12 + 2 * lane) =
          (uint16_t)(tape[2 * lane] | (uint32_t)tape[2 * lane + 1] << 8);
      rec++;
    }
    __syncthreads();  // every lane has read the tape before it changes
    if (!step<MAP, LOGIC>(tape, h0, h1, pc)) nskip++;
    __syncthreads();
    if (pc < 0) {
      i++;
      break;
    }
    pc++;
    if (pc >= kTape2) {
      i++;
      break;
    }
  }
  if (lane == 0) {
    meta[0] = rec;
    meta[1] = i - nskip;
    meta[2] = i;
  }
}

namespace {
template <int MAP, bool HEADS, int LOGIC>
hipError_t trace_one(hipStream_t st, const uint8_t *tape, uint32_t steps, uint8_t *out,
                     uint64_t cap, uint64_t *meta) {
  hipLaunchKernelGGL((k_trace<MAP, HEADS, LOGIC>), dim3(1), dim3(64), 0, st, tape, steps, out,
                     cap, meta);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_trace_bff(hipStream_t st, int lang, const uint8_t *tape, uint32_t steps,
                            uint8_t *out, uint64_t cap, uint64_t *meta) {
  switch (lang) {
    case kBff: return trace_one<kMapAscii, true, kLogicStd>(st, tape, steps, out, cap, meta);
    case kBffNoheads: return trace_one<kMapAscii, false, kLogicStd>(st, tape, steps, out, cap, meta);
    case kBff8: return trace_one<kMapAscii, true, kLogicBff8>(st, tape, steps, out, cap, meta);
    case kBff8Noheads:
      return trace_one<kMapAscii, false, kLogicBff8>(st, tape, steps, out, cap, meta);
    case kBffNoheads4bit:
      return trace_one<kMap4bit, false, kLogicStd>(st, tape, steps, out, cap, meta);
    case kBffPerm: return trace_one<kMapPerm, true, kLogicStd>(st, tape, steps, out, cap, meta);
    case kBffSelfmove:
      return trace_one<kMapSelfmove, true, kLogicSelfmove>(st, tape, steps, out, cap, meta);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cubff
