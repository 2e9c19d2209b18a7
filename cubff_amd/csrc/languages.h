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

// 128-bit per-lane position masks (two 64-bit halves held in VGPRs).
struct Mask128 {
  uint64_t lo, hi;
  __device__ __forceinline__ bool test(int p) const {
    return ((p < 64 ? lo : hi) >> (p & 63)) & 1;
  }
  __device__ __forceinline__ void flip(int p, bool f) {
    const uint64_t b = (uint64_t)f << (p & 63);
    if (p < 64) lo ^= b; else hi ^= b;
  }
  // Smallest set position >= p (0 <= p <= 128), or 128.
  __device__ __forceinline__ int next_from(int p) const {
    const uint64_t l = p >= 64 ? 0 : lo & (~0ULL << p);
    const uint64_t h = p <= 64 ? hi : (p >= 128 ? 0 : hi & (~0ULL << (p - 64)));
    return l ? __builtin_ctzll(l) : (h ? 64 + __builtin_ctzll(h) : 128);
  }
  // Largest set position <= p (-1 <= p <= 127), or -1.
  __device__ __forceinline__ int prev_upto(int p) const {
    const uint64_t h = p < 64 ? 0 : (p >= 127 ? hi : hi & (~0ULL >> (127 - p)));
    const uint64_t l = p < 0 ? 0 : (p >= 63 ? lo : lo & (~0ULL >> (63 - p)));
    return h ? 127 - __builtin_clzll(h) : (l ? 63 - __builtin_clzll(l) : -1);
  }
};

// Byte classes of a BFF-family map: command (not a comment and not the
// null op), loop start, loop end.
template <int MAP>
struct BffClass {
  __device__ __forceinline__ static bool cmd(uint32_t v) {
    if constexpr (MAP == kMapAscii) {
      // + , - . < > (0x2b-0x2e, 0x3c, 0x3e) and [ ] { } (0x5b, 0x5d, 0x7b, 0x7d).
      constexpr uint64_t kLo = (0xfULL << 0x2b) | (1ULL << 0x3c) | (1ULL << 0x3e);
      constexpr uint64_t kHi = (1ULL << (0x5b - 64)) | (1ULL << (0x5d - 64)) |
                               (1ULL << (0x7b - 64)) | (1ULL << (0x7d - 64));
      return v < 128 && (((v < 64 ? kLo : kHi) >> (v & 63)) & 1);
    } else if constexpr (MAP == kMap4bit) {
      return (v & 15) >= 6;
    } else if constexpr (MAP == kMapPerm) {
      return v - 1 < 10u;
    } else {
      return v < 7;
    }
  }
  __device__ __forceinline__ static bool ls(uint32_t v) {
    if constexpr (MAP == kMapAscii) return v == '[';
    else if constexpr (MAP == kMap4bit) return (v & 15) == 6;
    else if constexpr (MAP == kMapPerm) return v == 9;
    else return v == 5;
  }
  __device__ __forceinline__ static bool le(uint32_t v) {
    if constexpr (MAP == kMapAscii) return v == ']';
    else if constexpr (MAP == kMap4bit) return (v & 15) == 7;
    else if constexpr (MAP == kMapPerm) return v == 10;
    else return v == 6;
  }
  // Loop-test "byte counts as zero": kind == kNull for the standard logic
  // (bff.inc.h:318, :332), the byte itself for bff8 / selfmove.
  __device__ __forceinline__ static bool null(uint32_t v) {
    if constexpr (MAP == kMap4bit) return (v & 15) == 0;
    else return v == 0;
  }
};

template <int MAP, bool HEADS, int LOGIC>
struct BffLang {
  static constexpr int kMap = MAP;
  static constexpr bool kHeads = HEADS;
  static constexpr int kLogic = LOGIC;
  using C = BffClass<MAP>;

  // Evaluate (bff.inc.h:356-389) with EvaluateOne (bff.inc.h:289-349,
  // bff8.cu:23-65, bff_selfmove.cu:44-98).
  //
  // Runs of comment / null bytes change nothing but the step counters, so
  // the lane advances over a whole run at once, found with a 128-bit mask
  // of command positions; the step cap still binds inside a run.  Bracket
  // scans (bff.inc.h:317-344) walk 128-bit masks of `[` and `]` positions
  // instead of every byte.  The masks are kept equal to the tape: every
  // write that changes a byte's class flips the affected bits.
  template <class T>
  __device__ static uint32_t evaluate(const T &t, const T & /*aux*/,
                                      uint32_t stepcount, uint32_t *steps) {
    Mask128 mc{0, 0}, ml{0, 0}, me{0, 0};
#pragma unroll
    for (int w = 0; w < kTape2 / 4; w++) {
      const uint32_t word = t.word(w);
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const uint32_t v = (word >> (8 * b)) & 0xff;
        const int p = w * 4 + b;
        const uint64_t bit = 1ULL << (p & 63);
        if (p < 64) {
          mc.lo |= C::cmd(v) ? bit : 0;
          ml.lo |= C::ls(v) ? bit : 0;
          me.lo |= C::le(v) ? bit : 0;
        } else {
          mc.hi |= C::cmd(v) ? bit : 0;
          ml.hi |= C::ls(v) ? bit : 0;
          me.hi |= C::le(v) ? bit : 0;
        }
      }
    }
    int h0, h1, pc;
    if constexpr (HEADS) {  // InitialState, bff.inc.h:269-275
      h0 = t.get(0) & (kTape2 - 1);
      h1 = t.get(1) & (kTape2 - 1);
      pc = 2;
    } else {  // bff.inc.h:277-282; 128 masks to 0
      h0 = 0;
      h1 = 0;
      pc = 0;
    }
    uint32_t i = 0, nskip = 0;
    for (;;) {
      // Comment run [pc, nxt): each byte is one step that only counts.
      const int nxt = mc.next_from(pc);
      const uint32_t adv = min((uint32_t)(nxt - pc), stepcount - i);
      i += adv;
      nskip += adv;
      pc += adv;
      if (i >= stepcount || pc >= kTape2) break;
      // One command at pc.
      const uint32_t c = t.get(pc);
      const uint32_t v0 = t.get(h0);
      const uint32_t v1 = t.get(h1);
      i++;
      const int k = bff_kind<MAP>(c);
      if (k >= kDec0) {  // head moves (kNull/kNoop cannot occur here)
        h0 = (h0 + (k == kInc0) - (k == kDec0)) & (kTape2 - 1);
        h1 = (h1 + (k == kInc1) - (k == kDec1)) & (kTape2 - 1);
      } else if (k >= kPlus) {  // + - . ,
        int addr = h0;
        uint32_t old = v0, val;
        if (k == kPlus) val = v0 + 1;
        else if (k == kMinus) val = v0 - 1;
        else if (k == kCopy01) { addr = h1; old = v1; val = v0; }
        else val = v1;  // kCopy10
        val &= 0xff;
        if (LOGIC == kLogicSelfmove && k == kCopy10) h1 = (h1 + 1) & (kTape2 - 1);
        if (val != old) {
          t.set(addr, (uint8_t)val);
          mc.flip(addr, C::cmd(val) != C::cmd(old));
          ml.flip(addr, C::ls(val) != C::ls(old));
          me.flip(addr, C::le(val) != C::le(old));
        }
      } else if (k == kLoopStart) {
        if (LOGIC == kLogicBff8) {
          if (v0 == 0) pc += 8;
        } else if (C::null(v0)) {
          // Forward scan to the matching loop end; unmatched -> 128.
          int depth = 1, p = pc + 1;
          for (;;) {
            const int ql = ml.next_from(p), qe = me.next_from(p);
            const int q = min(ql, qe);
            if (q >= kTape2) { pc = kTape2; break; }
            depth += (q == ql) ? 1 : -1;
            if (depth == 0) { pc = q; break; }
            p = q + 1;
          }
        }
      } else {  // kLoopEnd
        if (LOGIC == kLogicBff8) {
          if (v0 != 0) pc -= 8;
        } else if (!C::null(v0)) {
          // Backward scan to the matching loop start; unmatched -> -1.
          int depth = 1, p = pc - 1;
          for (;;) {
            const int ql = ml.prev_upto(p), qe = me.prev_upto(p);
            const int q = max(ql, qe);
            if (q < 0) { pc = -1; break; }
            depth += (q == qe) ? 1 : -1;
            if (depth == 0) { pc = q; break; }
            p = q - 1;
          }
        }
      }
      if (pc < 0) break;
      pc++;
      if (pc >= kTape2 || i >= stepcount) break;
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
