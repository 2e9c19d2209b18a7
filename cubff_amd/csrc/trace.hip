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

// ---------------------------------------------------------------------------
// Step records of every language: the state each language's Evaluate prints
// before every step under `debug` (bff.inc.h:368-372, forth.inc.h:385-388,
// subleq.cu:114-118, rsubleq4.cu:122-126).  Record: u32 pc, u32 r0, r1, r2,
// the 128 tape bytes, 128 stack bytes, with r0/r1 = head0/head1 (BFF), r0 =
// stack depth (Forth, stack bytes [0, r0) valid), r0/r1 = the two operand
// addresses (SUBLEQ) or r0/r1/r2 = the three (RSUBLEQ4).  One thread: a
// debugging aid, not a hot path.  meta = {records, ops, steps}; the final
// tape is written back.

namespace {

struct PlainTape {  // get/set view for ForthLang::reset
  uint8_t *p;
  __device__ uint32_t get(int i) const { return p[i]; }
  __device__ void set(int i, uint8_t v) const { p[i] = v; }
};

// EvaluateOne of the Forth family (forth.inc.h:251-373; the reset checks of
// forthtrivial_reset.cu:166-176 and :221-229).
template <int LANG>
__device__ void forth_step(uint8_t *tape, uint8_t *stk, int &sp, bool &ovf, int &pos,
                           uint32_t &nops) {
  auto push = [&](uint32_t v) {
    if (sp != kTape2) stk[sp++] = (uint8_t)v;
    else ovf = true;
  };
  auto pop = [&]() -> uint32_t { return sp == 0 ? 0u : stk[--sp]; };
  constexpr bool kReset = LANG == kForthtrivialReset;
  const uint32_t c = tape[pos];
  switch (forth_kind<LANG>(c)) {
    case fRead0: { const uint32_t a = pop() % kTape; push(tape[a]); break; }
    case fRead1: { const uint32_t a = pop() % kTape; push(tape[kTape + a]); break; }
    case fRead: { const uint32_t a = pop() % kTape2; push(tape[a]); break; }
    case fWrite: {
      const uint32_t v = pop(), a = pop() % kTape2;
      tape[a] = (uint8_t)v;
      break;
    }
    case fWrite0: {
      const uint32_t v = pop(), a = pop() % kTape;
      tape[a] = (uint8_t)v;
      break;
    }
    case fWrite1: {
      const uint32_t v = pop(), a = pop() % kTape;
      if (kReset && tape[kTape + a] == 0x13) ForthLang<LANG>::reset(PlainTape{tape}, pos);
      else tape[kTape + a] = (uint8_t)v;
      break;
    }
    case fDup: { const uint32_t v = pop(); push(v); push(v); break; }
    case fDrop: pop(); break;
    case fSwap: { const uint32_t x = pop(), y = pop(); push(x); push(y); break; }
    case fIf0: {
      const uint32_t v = pop();
      if (v) pos++;
      push(v);
      break;
    }
    case fInc: push(pop() + 1); break;
    case fDec: push(pop() - 1); break;
    case fAdd: { const uint32_t x = pop(), y = pop(); push(x + y); break; }
    case fSub: { const uint32_t x = pop(), y = pop(); push(x - y); break; }
    case fConst: push(c & 63); break;
    case fCopy0: {
      const uint32_t a = pop() % kTape;
      if (kReset && tape[kTape + a] == 0x13) ForthLang<LANG>::reset(PlainTape{tape}, pos);
      else tape[kTape + a] = tape[a];
      break;
    }
    case fCopy1: { const uint32_t a = pop() % kTape; tape[a] = tape[kTape + a]; break; }
    case fCopy: {
      const uint32_t to = pop() % kTape2, from = pop() % kTape2;
      tape[to] = tape[from];
      break;
    }
    case fXor: push(pop() ^ 64); break;
    case fJmp: {
      const int d = (int)(c & 63) + 1;
      pos += ((c & 64) ? -d : d) - 1;
      break;
    }
    default: nops++; break;
  }
  pos++;
}

template <int LANG>
__global__ __launch_bounds__(64) void k_steps(uint8_t *__restrict__ tape_io, uint32_t steps,
                                              uint8_t *__restrict__ out, uint64_t cap,
                                              uint64_t *__restrict__ meta) {
  __shared__ uint8_t tape[kTape2], stk[kTape2];
  if (threadIdx.x != 0) return;
  for (int i = 0; i < kTape2; i++) {
    tape[i] = tape_io[i];
    stk[i] = 0;
  }
  uint64_t rec = 0;
  auto record = [&](int pc, uint32_t r0, uint32_t r1, uint32_t r2) {
    if (rec >= cap) return;
    uint32_t *o = reinterpret_cast<uint32_t *>(out + rec * kStepRecordBytes);
    o[0] = (uint32_t)pc;
    o[1] = r0;
    o[2] = r1;
    o[3] = r2;
    for (int w = 0; w < kTape2 / 4; w++) {
      o[4 + w] = reinterpret_cast<const uint32_t *>(tape)[w];
      o[4 + kTape2 / 4 + w] = reinterpret_cast<const uint32_t *>(stk)[w];
    }
    rec++;
  };
  uint32_t i = 0, nskip = 0;
  if constexpr (LANG >= kBff && LANG <= kBffSelfmove) {
    constexpr int MAP = LANG == kBffNoheads4bit ? kMap4bit
                        : LANG == kBffPerm      ? kMapPerm
                        : LANG == kBffSelfmove  ? kMapSelfmove
                                                : kMapAscii;
    constexpr int LOGIC = (LANG == kBff8 || LANG == kBff8Noheads) ? kLogicBff8
                          : LANG == kBffSelfmove                  ? kLogicSelfmove
                                                                  : kLogicStd;
    constexpr bool HEADS = !(LANG == kBffNoheads || LANG == kBff8Noheads ||
                             LANG == kBffNoheads4bit);
    int h0 = HEADS ? tape[0] % kTape2 : kTape2, h1 = HEADS ? tape[1] % kTape2 : kTape2;
    int pc = HEADS ? 2 : 0;
    for (; i < steps; i++) {  // bff.inc.h:356-389
      h0 &= kTape2 - 1;
      h1 &= kTape2 - 1;
      record(pc, (uint32_t)h0, (uint32_t)h1, 0);
      if (!step<MAP, LOGIC>(tape, h0, h1, pc)) nskip++;
      if (pc < 0 || ++pc >= kTape2) {
        i++;
        break;
      }
    }
  } else if constexpr (LANG >= kForth && LANG <= kForthtrivialReset) {
    int sp = 0, pos = 0;
    bool ovf = false;
    for (; i < steps; i++) {  // forth.inc.h:379-397
      record(pos, (uint32_t)sp, 0, 0);
      forth_step<LANG>(tape, stk, sp, ovf, pos, nskip);
      if (pos >= kTape2 || pos < 0 || ovf) {
        i++;
        break;
      }
    }
  } else if constexpr (LANG == kSubleq) {
    int pos = 0;
    for (; i < steps; i++) {  // subleq.cu:109-135
      const uint32_t sa = tape[pos] % kTape2, sb = tape[pos + 1] % kTape2;
      record(pos, sa, sb, 0);
      tape[sa] = (uint8_t)(tape[sa] - tape[sb]);
      if ((tape[sa] & 0x80) || tape[sa] == 0) pos = tape[pos + 2];
      else pos += 3;
      if (pos + 3 > kTape2 || pos < 0) {
        i++;
        break;
      }
    }
  } else {
    int pos = 0;
    for (; i < steps; i++) {  // rsubleq4.cu:116-142
      const uint32_t sa = (pos + tape[pos]) % kTape2, sb = (pos + tape[pos + 1]) % kTape2;
      const uint32_t sc = (pos + tape[pos + 2]) % kTape2;
      record(pos, sa, sb, sc);
      tape[sa] = (uint8_t)(tape[sb] - tape[sc]);
      if ((tape[sa] & 0x80) || tape[sa] == 0) pos += (int8_t)tape[pos + 3];
      else pos += 4;
      if (pos + 4 > kTape2 || pos < 0) {
        i++;
        break;
      }
    }
  }
  for (int q = 0; q < kTape2; q++) tape_io[q] = tape[q];
  meta[0] = rec;
  meta[1] = i - nskip;
  meta[2] = i;
}

template <int LANG>
hipError_t steps_one(hipStream_t st, uint8_t *tape, uint32_t steps, uint8_t *out, uint64_t cap,
                     uint64_t *meta) {
  hipLaunchKernelGGL(k_steps<LANG>, dim3(1), dim3(64), 0, st, tape, steps, out, cap, meta);
  return hipGetLastError();
}

template <int MAP, bool HEADS, int LOGIC>
hipError_t trace_one(hipStream_t st, const uint8_t *tape, uint32_t steps, uint8_t *out,
                     uint64_t cap, uint64_t *meta) {
  hipLaunchKernelGGL((k_trace<MAP, HEADS, LOGIC>), dim3(1), dim3(64), 0, st, tape, steps, out,
                     cap, meta);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_trace_steps(hipStream_t st, int lang, uint8_t *tape, uint32_t steps,
                              uint8_t *out, uint64_t cap, uint64_t *meta) {
  switch (lang) {
#define CUBFF_STEPS_CASE(L) \
  case L: return steps_one<L>(st, tape, steps, out, cap, meta);
    CUBFF_STEPS_CASE(kBff)
    CUBFF_STEPS_CASE(kBffNoheads)
    CUBFF_STEPS_CASE(kBff8)
    CUBFF_STEPS_CASE(kBff8Noheads)
    CUBFF_STEPS_CASE(kBffNoheads4bit)
    CUBFF_STEPS_CASE(kBffPerm)
    CUBFF_STEPS_CASE(kBffSelfmove)
    CUBFF_STEPS_CASE(kForth)
    CUBFF_STEPS_CASE(kForthcopy)
    CUBFF_STEPS_CASE(kForthtrivial)
    CUBFF_STEPS_CASE(kForthtrivialReset)
    CUBFF_STEPS_CASE(kSubleq)
    CUBFF_STEPS_CASE(kRsubleq4)
#undef CUBFF_STEPS_CASE
    default: return hipErrorInvalidValue;
  }
}

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
