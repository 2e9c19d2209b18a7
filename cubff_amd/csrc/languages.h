// Device interpreters of the 13 cubff languages, one 128-byte tape per lane.
//
// Each language is a small policy type with a static `evaluate(tape, aux,
// stepcount, &steps)` returning ops (loop iterations minus no-op iterations),
// the value the reference's Language::Evaluate returns.  Tapes are accessed
// through an LDS accessor (LdsTape); `aux` is a second per-lane LDS region
// (the Forth stack).
#pragma once

#include "device_common.h"

namespace cubff {

enum LangId : int {
  kBff = 0,
  kBffNoheads,
  kBff8,
  kBff8Noheads,
  kBffNoheads4bit,
  kBffPerm,
  kBffSelfmove,
  kForth,
  kForthcopy,
  kForthtrivial,
  kForthtrivialReset,
  kSubleq,
  kRsubleq4,
  kNumLangs
};

// Registry names, as REGISTERed by the reference's language files.
inline const char *const kLangNames[kNumLangs] = {
    "bff",          "bff_noheads",        "bff8",    "bff8_noheads",
    "bff_noheads_4bit", "bff_perm",       "bff_selfmove", "forth",
    "forthcopy",    "forthtrivial",       "forthtrivial_reset", "subleq",
    "rsubleq4"};

constexpr bool lang_needs_aux(int lang) {
  return lang >= kForth && lang <= kForthtrivialReset;
}

// ---------------------------------------------------------------------------
// BFF family (bff.inc.h:26-389 and the variant files).

enum BffKind : int {
  kLoopStart, kLoopEnd, kPlus, kMinus, kCopy01, kCopy10, kDec0, kInc0, kDec1,
  kInc1, kNull, kNoop
};

enum BffLogic : int { kLogicStd, kLogicBff8, kLogicSelfmove };

// Op maps: bff.inc.h:58-85 (ASCII; bytes >= 128 are comments because the
// reference switches on a signed char), bff_noheads_4bit.cu:22-49,
// bff_perm.cu:23-50, bff_selfmove.cu:24-42.
enum BffMap : int { kMapAscii, kMap4bit, kMapPerm, kMapSelfmove };

template <int MAP>
__device__ __forceinline__ int bff_kind(uint32_t c) {
  if constexpr (MAP == kMapAscii) {
    switch (c) {
      case '[': return kLoopStart;
      case ']': return kLoopEnd;
      case '+': return kPlus;
      case '-': return kMinus;
      case '.': return kCopy01;
      case ',': return kCopy10;
      case '<': return kDec0;
      case '>': return kInc0;
      case '{': return kDec1;
      case '}': return kInc1;
      case 0: return kNull;
      default: return kNoop;
    }
  } else if constexpr (MAP == kMap4bit) {
    uint32_t m = c & 15;
    if (m == 0) return kNull;
    if (m < 6) return kNoop;
    // 6..15 -> [ ] + - . , < > { }
    return static_cast<int>(m - 6);
  } else if constexpr (MAP == kMapPerm) {
    switch (c) {
      case 0: return kNull;
      case 1: return kDec0;
      case 2: return kInc0;
      case 3: return kDec1;
      case 4: return kInc1;
      case 5: return kPlus;
      case 6: return kMinus;
      case 7: return kCopy01;
      case 8: return kCopy10;
      case 9: return kLoopStart;
      case 10: return kLoopEnd;
      default: return kNoop;
    }
  } else {
    switch (c) {
      case 0: return kInc0;
      case 1: return kDec0;
      case 2: return kPlus;
      case 3: return kMinus;
      case 4: return kCopy10;
      case 5: return kLoopStart;
      case 6: return kLoopEnd;
      default: return kNoop;
    }
  }
}

template <int MAP, bool HEADS, int LOGIC>
struct BffLang {
  static constexpr int kMap = MAP;
  static constexpr bool kHeads = HEADS;
  static constexpr int kLogic = LOGIC;

  // Forward scan to the matching loop end (bff.inc.h:319-328).
  template <class T>
  __device__ __forceinline__ static int scan_forward(const T &t, int pc) {
    int depth = 1;
    int p = pc + 1;
    for (; p < kTape2 && depth > 0; p++) {
      int k = bff_kind<MAP>(t.get(p));
      depth += (k == kLoopStart) - (k == kLoopEnd);
    }
    return depth != 0 ? kTape2 : p - 1;
  }

  // Backward scan to the matching loop start (bff.inc.h:333-342).
  template <class T>
  __device__ __forceinline__ static int scan_backward(const T &t, int pc) {
    int depth = 1;
    int p = pc - 1;
    for (; p >= 0 && depth > 0; p--) {
      int k = bff_kind<MAP>(t.get(p));
      depth += (k == kLoopEnd) - (k == kLoopStart);
    }
    return depth != 0 ? -1 : p + 1;
  }

  // Evaluate (bff.inc.h:356-389) with EvaluateOne (bff.inc.h:289-349,
  // bff8.cu:23-65, bff_selfmove.cu:44-98).
  template <class T>
  __device__ static uint32_t evaluate(const T &t, const T & /*aux*/,
                                      uint32_t stepcount, uint32_t *steps) {
    int h0, h1, pc;
    if constexpr (HEADS) {  // InitialState, bff.inc.h:269-275
      h0 = t.get(0) & (kTape2 - 1);
      h1 = t.get(1) & (kTape2 - 1);
      pc = 2;
    } else {  // bff.inc.h:277-282 (128 masks to 0)
      h0 = 0;
      h1 = 0;
      pc = 0;
    }
    uint32_t i = 0, nskip = 0;
    while (i < stepcount) {
      h0 &= kTape2 - 1;
      h1 &= kTape2 - 1;
      i++;
      int k = bff_kind<MAP>(t.get(pc));
      switch (k) {
        case kDec0: h0--; break;
        case kInc0: h0++; break;
        case kDec1:
          if (LOGIC == kLogicSelfmove) nskip++; else h1--;
          break;
        case kInc1:
          if (LOGIC == kLogicSelfmove) nskip++; else h1++;
          break;
        case kPlus: t.set(h0, t.get(h0) + 1); break;
        case kMinus: t.set(h0, t.get(h0) - 1); break;
        case kCopy01:
          if (LOGIC == kLogicSelfmove) nskip++; else t.set(h1, t.get(h0));
          break;
        case kCopy10:
          t.set(h0, t.get(h1));
          if (LOGIC == kLogicSelfmove) h1++;
          break;
        case kLoopStart: {
          uint32_t v = t.get(h0);
          if (LOGIC == kLogicBff8) {
            if (v == 0) pc += 8;
          } else if (LOGIC == kLogicSelfmove) {
            if (v == 0) pc = scan_forward(t, pc);
          } else if (bff_kind<MAP>(v) == kNull) {
            pc = scan_forward(t, pc);
          }
          break;
        }
        case kLoopEnd: {
          uint32_t v = t.get(h0);
          if (LOGIC == kLogicBff8) {
            if (v != 0) pc -= 8;
          } else if (LOGIC == kLogicSelfmove) {
            if (v != 0) pc = scan_backward(t, pc);
          } else if (bff_kind<MAP>(v) != kNull) {
            pc = scan_backward(t, pc);
          }
          break;
        }
        default: nskip++; break;
      }
      if (pc < 0) break;
      pc++;
      if (pc >= kTape2) break;
    }
    *steps = i;
    return i - nskip;
  }
};

// ---------------------------------------------------------------------------
// Forth family (forth.inc.h:26-397 and the variant files).

enum ForthKind : int {
  fWrite0, fWrite1, fWrite, fRead, fRead0, fRead1, fDup, fDrop, fSwap, fIf0,
  fInc, fDec, fAdd, fSub, fCopy, fCopy0, fCopy1, fXor, fConst, fJmp, fNoop
};

// forth.inc.h:54-84, forthcopy.cu:21-51, forthtrivial.cu:21-55.
template <int LANG>
__device__ __forceinline__ int forth_kind(uint32_t c) {
  if (c >= 128) return fJmp;
  if (c >= 64) return fConst;
  if (LANG == kForthcopy) {
    switch (c) {
      case 0: return fRead;
      case 1: return fWrite;
      case 2: return fCopy;
      case 3: return fXor;
      default: break;
    }
  } else {
    switch (c) {
      case 0: return fRead0;
      case 1: return fRead1;
      case 2: return fWrite0;
      case 3: return fWrite1;
      default: break;
    }
  }
  switch (c) {
    case 4: return fDup;
    case 5: return fDrop;
    case 6: return fSwap;
    case 7: return fIf0;
    case 8: return fInc;
    case 9: return fDec;
    case 10: return fAdd;
    case 11: return fSub;
    default: break;
  }
  if (LANG == kForthtrivial || LANG == kForthtrivialReset) {
    if (c == 12) return fCopy0;
    if (c == 13) return fCopy1;
  }
  return fNoop;
}

template <int LANG>
struct ForthLang {
  // Reset of forthtrivial_reset.cu:58-68.
  template <class T>
  __device__ static void reset(const T &t, int pos) {
    uint64_t seed = 0;
    for (int i = 0; i < kTape2; i++) seed = splitmix64(seed ^ t.get(i));
    const uint64_t base = seed * (uint64_t)(kTape * kTape) +
                          (uint64_t)(int64_t)(pos * kTape);
    for (int i = 0; i < kTape; i++)
      t.set(i, (uint8_t)splitmix64(base + (uint64_t)i));
  }

  // Evaluate (forth.inc.h:379-397) with EvaluateOne (forth.inc.h:251-373,
  // forthtrivial_reset.cu:123-252).  The stack (forth.inc.h:236-249) lives in
  // the lane's aux LDS region.
  template <class T>
  __device__ static uint32_t evaluate(const T &t, const T &stk,
                                      uint32_t stepcount, uint32_t *steps) {
    int sp = 0;
    bool overflow = false;
    int pos = 0;
    uint32_t i = 0, nops = 0;
    auto push = [&](uint32_t v) {
      if (sp != 128) stk.set(sp++, (uint8_t)v);
      else overflow = true;
    };
    auto pop = [&]() -> uint32_t { return sp == 0 ? 0u : stk.get(--sp); };
    constexpr bool kReset = LANG == kForthtrivialReset;
    while (i < stepcount) {
      i++;
      const uint32_t cmd = t.get(pos);
      switch (forth_kind<LANG>(cmd)) {
        case fRead0: { uint32_t a = pop() % kTape; push(t.get(a)); break; }
        case fRead1: { uint32_t a = pop() % kTape; push(t.get(kTape + a)); break; }
        case fRead: { uint32_t a = pop() % kTape2; push(t.get(a)); break; }
        case fWrite: {
          uint32_t v = pop();
          uint32_t a = pop() % kTape2;
          t.set(a, (uint8_t)v);
          break;
        }
        case fWrite0: {
          uint32_t v = pop();
          uint32_t a = pop() % kTape;
          t.set(a, (uint8_t)v);
          break;
        }
        case fWrite1: {
          uint32_t v = pop();
          uint32_t a = pop() % kTape;
          if (kReset && t.get(kTape + a) == 0x13) reset(t, pos);
          else t.set(kTape + a, (uint8_t)v);
          break;
        }
        case fDup: { uint32_t v = pop(); push(v); push(v); break; }
        case fDrop: pop(); break;
        case fSwap: {
          uint32_t a = pop();
          uint32_t b = pop();
          push(a);
          push(b);
          break;
        }
        case fIf0: {
          uint32_t v = pop();
          if (v) pos++;
          push(v);
          break;
        }
        case fInc: push(pop() + 1); break;
        case fDec: push(pop() - 1); break;
        case fAdd: { uint32_t a = pop(); uint32_t b = pop(); push(a + b); break; }
        case fSub: { uint32_t a = pop(); uint32_t b = pop(); push(a - b); break; }
        case fConst: push(cmd & 63); break;
        case fCopy0: {
          uint32_t a = pop() % kTape;
          if (kReset && t.get(kTape + a) == 0x13) reset(t, pos);
          else t.set(kTape + a, t.get(a));
          break;
        }
        case fCopy1: { uint32_t a = pop() % kTape; t.set(a, t.get(kTape + a)); break; }
        case fCopy: {
          uint32_t to = pop() % kTape2;
          uint32_t from = pop() % kTape2;
          t.set(to, t.get(from));
          break;
        }
        case fXor: push(pop() ^ 64); break;
        case fJmp: {
          int a = (int)(cmd & 63) + 1;
          pos += (cmd & 64) ? -a : a;
          pos--;
          break;
        }
        default: nops++; break;
      }
      pos++;
      if (pos >= kTape2 || pos < 0 || overflow) break;
    }
    *steps = i;
    return i - nops;
  }
};

// ---------------------------------------------------------------------------
// SUBLEQ (subleq.cu:109-133) and RSUBLEQ4 (rsubleq4.cu:116-141).  Every
// iteration counts as an op.

struct SubleqLang {
  template <class T>
  __device__ static uint32_t evaluate(const T &t, const T &, uint32_t stepcount,
                                      uint32_t *steps) {
    int pos = 0;
    uint32_t i = 0;
    while (i < stepcount) {
      i++;
      uint32_t sa = t.get(pos) % kTape2;
      uint32_t sb = t.get(pos + 1) % kTape2;
      uint8_t r = (uint8_t)(t.get(sa) - t.get(sb));
      t.set(sa, r);
      if ((r & 0x80) || r == 0) pos = t.get(pos + 2);  // read after the write
      else pos += 3;
      if (pos + 3 > kTape2 || pos < 0) break;
    }
    *steps = i;
    return i;
  }
};

struct Rsubleq4Lang {
  template <class T>
  __device__ static uint32_t evaluate(const T &t, const T &, uint32_t stepcount,
                                      uint32_t *steps) {
    int pos = 0;
    uint32_t i = 0;
    while (i < stepcount) {
      i++;
      uint32_t sa = (uint32_t)(pos + t.get(pos)) % kTape2;
      uint32_t sb = (uint32_t)(pos + t.get(pos + 1)) % kTape2;
      uint32_t sc = (uint32_t)(pos + t.get(pos + 2)) % kTape2;
      uint8_t r = (uint8_t)(t.get(sb) - t.get(sc));
      t.set(sa, r);
      if ((r & 0x80) || r == 0) pos += (int8_t)t.get(pos + 3);  // after write
      else pos += 4;
      if (pos + 4 > kTape2 || pos < 0) break;
    }
    *steps = i;
    return i;
  }
};

// Compile-time language table.
template <int LANG> struct LangOf;
template <> struct LangOf<kBff> { using T = BffLang<kMapAscii, true, kLogicStd>; };
template <> struct LangOf<kBffNoheads> { using T = BffLang<kMapAscii, false, kLogicStd>; };
template <> struct LangOf<kBff8> { using T = BffLang<kMapAscii, true, kLogicBff8>; };
template <> struct LangOf<kBff8Noheads> { using T = BffLang<kMapAscii, false, kLogicBff8>; };
template <> struct LangOf<kBffNoheads4bit> { using T = BffLang<kMap4bit, false, kLogicStd>; };
template <> struct LangOf<kBffPerm> { using T = BffLang<kMapPerm, true, kLogicStd>; };
template <> struct LangOf<kBffSelfmove> { using T = BffLang<kMapSelfmove, true, kLogicSelfmove>; };
template <> struct LangOf<kForth> { using T = ForthLang<kForth>; };
template <> struct LangOf<kForthcopy> { using T = ForthLang<kForthcopy>; };
template <> struct LangOf<kForthtrivial> { using T = ForthLang<kForthtrivial>; };
template <> struct LangOf<kForthtrivialReset> { using T = ForthLang<kForthtrivialReset>; };
template <> struct LangOf<kSubleq> { using T = SubleqLang; };
template <> struct LangOf<kRsubleq4> { using T = Rsubleq4Lang; };

}  // namespace cubff
