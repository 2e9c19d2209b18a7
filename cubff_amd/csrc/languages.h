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

// evaluate<true>(..., budget) stops once it has executed more than `budget`
// commands and reports *steps = kOverBudget (the tape is then partly run and
// must be discarded).
constexpr uint32_t kOverBudget = 0xffffffffu;

// Languages whose interpreter uses a second per-lane LDS region: the Forth
// stack, and the tape snapshot of the BFF cycle check.
constexpr bool lang_needs_aux(int lang) {
  return (lang >= kForth && lang <= kForthtrivialReset) || lang == kBff ||
         lang == kBffNoheads || lang == kBffNoheads4bit || lang == kBffPerm ||
         lang == kBff8 || lang == kBff8Noheads || lang == kRsubleq4;
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

// Byte-linear per-lane tape in LDS: byte i of lane l sits at i * kStride + l.
// One multiply-add forms any byte address.  The stride of LANES + 4 bytes
// moves each row by one bank, so lanes reading different positions rarely
// collide.  Row kTape2 is a per-lane dummy byte that absorbs writes which
// must not land on the tape.
template <int LANES>
struct LinTape {
  static constexpr int kStride = LANES + 4;
  static constexpr int kRegionBytes = (kTape2 + 1) * kStride;
  static constexpr int kDummy = kTape2;
  lds_u8 *p;

  __device__ __forceinline__ static LinTape at(uint8_t *region, int lane) {
    return LinTape{(lds_u8 *)region + lane};
  }
  __device__ __forceinline__ uint8_t get(uint32_t i) const { return p[i * kStride]; }
  __device__ __forceinline__ void set(uint32_t i, uint8_t v) const {
    p[i * kStride] = v;
  }
  __device__ __forceinline__ uint32_t word(uint32_t w) const {
    return (uint32_t)get(4 * w) | ((uint32_t)get(4 * w + 1) << 8) |
           ((uint32_t)get(4 * w + 2) << 16) | ((uint32_t)get(4 * w + 3) << 24);
  }
  __device__ __forceinline__ void set_word(uint32_t w, uint32_t v) const {
    set(4 * w, (uint8_t)v);
    set(4 * w + 1, (uint8_t)(v >> 8));
    set(4 * w + 2, (uint8_t)(v >> 16));
    set(4 * w + 3, (uint8_t)(v >> 24));
  }
};

// 128-bit per-lane position masks (two 64-bit halves held in VGPRs).
struct Mask128 {
  uint64_t lo, hi;
  __host__ __device__ __forceinline__ void flip(int p, bool f) {
    const uint64_t b = (uint64_t)f << (p & 63);
    lo ^= p < 64 ? b : 0ULL;
    hi ^= p < 64 ? 0ULL : b;
  }
  // Smallest set position >= p (0 <= p <= 128), or 128.  Branch-free.
  __host__ __device__ __forceinline__ int next_from(int p) const {
    const uint64_t l = lo & (p >= 64 ? 0ULL : ~0ULL << (p & 63));
    const uint64_t h = hi & (p <= 64 ? ~0ULL : (p >= 128 ? 0ULL : ~0ULL << (p & 63)));
    const int cl = __builtin_ctzg(l, 64), ch = __builtin_ctzg(h, 64);
    return l ? cl : 64 + ch;
  }
  // Largest set position <= p (-1 <= p <= 127), or -1.
  __host__ __device__ __forceinline__ int prev_upto(int p) const {
    const uint64_t h = p < 64 ? 0 : (p >= 127 ? hi : hi & (~0ULL >> (127 - p)));
    const uint64_t l = p < 0 ? 0 : (p >= 63 ? lo : lo & (~0ULL >> (63 - p)));
    return h ? 127 - __builtin_clzll(h) : (l ? 63 - __builtin_clzll(l) : -1);
  }
};

// Byte classes and opcode decode of a BFF-family map.
//
// At a command position the opcode is decoded without a compare tree: a
// 4-bit hash of the byte indexes a 64-bit table of 4-bit kinds.  For the
// ASCII map the hash ((c * 7) >> 2) & 15 is collision-free on [ ] + - . , <
// > { }; the other maps index by the byte (or its low nibble) directly.
template <int MAP>
struct BffClass {
  __host__ __device__ __forceinline__ static bool cmd(uint32_t v) {
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
  __host__ __device__ __forceinline__ static bool ls(uint32_t v) {
    if constexpr (MAP == kMapAscii) return v == '[';
    else if constexpr (MAP == kMap4bit) return (v & 15) == 6;
    else if constexpr (MAP == kMapPerm) return v == 9;
    else return v == 5;
  }
  __host__ __device__ __forceinline__ static bool le(uint32_t v) {
    if constexpr (MAP == kMapAscii) return v == ']';
    else if constexpr (MAP == kMap4bit) return (v & 15) == 7;
    else if constexpr (MAP == kMapPerm) return v == 10;
    else return v == 6;
  }
  // Loop-test "byte counts as zero": kind == kNull for the standard logic
  // (bff.inc.h:318, :332), the byte itself for bff8 / selfmove.
  __host__ __device__ __forceinline__ static bool null(uint32_t v) {
    if constexpr (MAP == kMap4bit) return (v & 15) == 0;
    else return v == 0;
  }
  // BffKind of a command byte (undefined for comments / null).
  __host__ __device__ __forceinline__ static uint32_t kind(uint32_t c) {
    uint32_t h;
    uint64_t table;
    if constexpr (MAP == kMapAscii) {
      h = ((c * 7) >> 2) & 15;
      table = 0x0357296080000104ULL;
    } else if constexpr (MAP == kMap4bit) {
      h = c & 15;
      table = 0x9876543210000000ULL;
    } else if constexpr (MAP == kMapPerm) {
      h = c & 15;
      table = 0x0000010543298760ULL;
    } else {
      h = c & 15;
      table = 0x0000000001053267ULL;
    }
    return (uint32_t)(table >> (4 * h)) & 15;
  }
};

template <int MAP, bool HEADS, int LOGIC>
struct BffLang {
  template <int LANES> using Tape = LinTape<LANES>;
  template <int LANES> using Aux = LdsTape<LANES>;  // unused
  using C = BffClass<MAP>;
  // Per-kind fields, 2 bits each, indexed by BffKind: head-0 delta + 1,
  // head-1 delta + 1 (bff_selfmove's `,` also advances head 1), byte delta + 1.
  static constexpr uint32_t kTH0 = 0x58555;
  static constexpr uint32_t kTH1 = LOGIC == kLogicSelfmove ? 0x55955 : 0x85555;
  static constexpr uint32_t kTDelta = 0x55525;
  // Loop acceleration needs the standard `[` / `]` semantics.
  static constexpr bool kAccel = LOGIC == kLogicStd;
  // Static-tape cycle skip where it measured faster (bff_selfmove: ~2.6x
  // per epoch with deferral; bff8's long runs mostly cycle with a changing
  // tape, and the check only costs there).
  static constexpr bool kStaticCycle = LOGIC == kLogicSelfmove;
  // Full-state cycle skip at command boundaries for the bff8 jump logic:
  // its step-capped runs are all periodic (periods mostly 8 steps, up to
  // ~1K; lead-ins ~50 steps, up to ~2K), many with a changing tape.  From
  // kFullStart steps on, the state at a command boundary (pc, heads, tape)
  // is compared with an anchor whose tape is kept in the aux region, after
  // a weighted sum of the writes since the anchor (zero when the tape is
  // back to the anchor's) filters candidates; anchors move after windows of
  // kFullStart, 2 kFullStart, ... steps (Brent).
  static constexpr bool kFullCycle = LOGIC == kLogicBff8;
  static constexpr uint32_t kFullStart = 64;

  // Evaluate (bff.inc.h:356-389) with EvaluateOne (bff.inc.h:289-349,
  // bff8.cu:23-65, bff_selfmove.cu:44-98).
  //
  // The loop keeps pc on a command byte.  Runs of comment / null bytes
  // change nothing but the step counters, so after each command the lane
  // moves straight to the next command, found in a 128-bit mask of command
  // positions, and counts the skipped bytes as no-op steps (the step cap
  // still binds inside a run).  Each command is straight-line code: table
  // decode, head deltas, one (possibly dummy) byte store.  Bracket scans
  // (bff.inc.h:317-344) walk 128-bit masks of `[` and `]` positions; the
  // last forward and last backward jump are cached as (source, position
  // after the target, next command from there) until a write changes any
  // byte's class.  Masks always equal the tape: every write that changes a
  // byte's class flips the affected bits.
  // Full-state cycle check (standard jump logic): from kCycleStart steps
  // on, the state right after a taken backward jump is compared with an
  // anchor whose tape is kept in the aux region.  The first anchor stays for
  // kCycleFirstWin steps: the longest runs are mostly cycles that start
  // within ~150 steps with periods of 1.5K-4K steps (a head sweeping the
  // whole tape in a loop), and half the 8192-step budget is the longest
  // period that can still be skipped.  Later anchors move after windows of
  // kCycleWin, 2 kCycleWin, ... steps (Brent), so a cycle that starts late
  // and is shorter than the remaining budget is found too.  A match (pc,
  // heads, then every tape byte) means the run repeats with that period.
  static constexpr uint32_t kCycleStart = 128;
  static constexpr uint32_t kCycleFirstWin = 4096;
  static constexpr uint32_t kCycleWin = 1024;
  __device__ __forceinline__ static uint32_t cycle_weight(int p) {
    return 2654435761u * (2u * (uint32_t)p + 1u);
  }

  template <bool kBudget = false, class T, class A>
  __device__ static uint32_t evaluate(const T &t, const A &aux,
                                      uint32_t stepcount, uint32_t *steps,
                                      uint32_t *work = nullptr, uint32_t budget = 0) {
    Mask128 mc{0, 0}, ml{0, 0}, me{0, 0};
    for (int w = 0; w < kTape2 / 8; w++) {
      uint64_t bc = 0, bl = 0, be = 0;
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const uint32_t v = t.get(w * 8 + b);
        bc |= (uint64_t)C::cmd(v) << b;
        bl |= (uint64_t)C::ls(v) << b;
        be |= (uint64_t)C::le(v) << b;
      }
      const int sh = (w * 8) & 63;
      if (w < 8) {
        mc.lo |= bc << sh; ml.lo |= bl << sh; me.lo |= be << sh;
      } else {
        mc.hi |= bc << sh; ml.hi |= bl << sh; me.hi |= be << sh;
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
    uint32_t i, nskip, iters = 0;  // iters: commands actually interpreted
    {
      const int nxt = mc.next_from(pc);
      i = nskip = min((uint32_t)(nxt - pc), stepcount);
      pc = nxt;
    }
    if (i < stepcount && pc < kTape2) {
      // Jump caches: source pc, position after the jump, next command.
      int bsrc = -1, bafter = 0, bnext = 0;
      int fsrc = -1, fafter = 0, fnext = 0;
      // Loop acceleration (standard jump logic only).  The anchor is the
      // state right after the last taken backward jump: its source aq, the
      // heads, step and no-op counts, and the bytes under the heads.  Flags
      // record what the commands since then did.
      int aq = -1, ah0 = 0, ah1 = 0;
      uint32_t ai = 0, askip = 0, aw0 = 0, aw1 = 0;
      bool wrote = false, moved = false, branched = false;
      // Cycle-check anchor: jump source (-1 none yet, -2 done), heads, step
      // and no-op counts, window, and a hash of the tape changes since.
      int cq = -1, ch0 = 0, ch1 = 0;
      uint32_t ci = 0, cskip = 0, cwin = kCycleFirstWin, cdelta = 0;
      bool cfirst = true;
      // Static-tape cycle skip at command boundaries (kStaticCycle), as for
      // SUBLEQ, on the state (pc, heads, tape): anchor, its step and no-op
      // counts, window, and whether a write changed a byte since.
      int spc = -1, sh0 = 0, sh1 = 0;
      uint32_t si = 0, sskip = 0, swin = 1;
      bool schg = false, sdone = false;
      // Full-state cycle skip (kFullCycle): anchor pc (-1 none yet), heads,
      // step and no-op counts, window, weighted sum of changes since.
      int fpc = -1, fh0 = 0, fh1 = 0;
      uint32_t fi = 0, fskip = 0, fwin = kFullStart, fdelta = 0;
      bool fdone = false;
      for (;;) {
        iters++;
        if constexpr (kBudget) {
          if (iters > budget) {
            *steps = kOverBudget;
            return 0;
          }
        }
        const uint32_t c = t.get(pc);
        const uint32_t v0 = t.get(h0);
        const uint32_t v1 = t.get(h1);
        // Keep the three loads together: one LDS round trip per command.
        __builtin_amdgcn_sched_barrier(0);
        int fall = mc.next_from(pc + 1);
        const uint32_t k = C::kind(c);
        if constexpr (kAccel) {
          wrote |= (k - kPlus) < 4u;
          moved |= (k - kDec0) < 4u;
          branched |= k == kLoopStart || (k == kLoopEnd && pc != aq);
        }
        const bool c01 = k == kCopy01;
        const int waddr = c01 ? h1 : h0;
        const uint32_t wold = c01 ? v1 : v0;
        const uint32_t wval =
            ((k == kCopy10 ? v1 : v0) + ((kTDelta >> (2 * k)) & 3) - 1) & 0xff;
        const bool changed = (k - kPlus) < 4u && wval != wold;
        t.set(changed ? waddr : T::kDummy, (uint8_t)wval);
        if constexpr (kAccel && !kBudget) {
          if (changed) cdelta += (wval - wold) * cycle_weight(waddr);
        }
        if constexpr (kFullCycle) {
          if (changed) fdelta += (wval - wold) * cycle_weight(waddr);
        }
        h0 = (h0 + (int)((kTH0 >> (2 * k)) & 3) - 1) & (kTape2 - 1);
        h1 = (h1 + (int)((kTH1 >> (2 * k)) & 3) - 1) & (kTape2 - 1);
        if (changed) {
          const bool fc = C::cmd(wval) != C::cmd(wold);
          const bool fl = C::ls(wval) != C::ls(wold);
          const bool fe = C::le(wval) != C::le(wold);
          if (fc | fl | fe) {
            mc.flip(waddr, fc);
            ml.flip(waddr, fl);
            me.flip(waddr, fe);
            bsrc = fsrc = -1;
            fall = mc.next_from(pc + 1);
          }
        }
        int after, nextc;
        if constexpr (LOGIC == kLogicBff8) {
          // bff8.cu:51-60: [ skips 8 when the byte is zero, ] goes back 8.
          const bool tf = k == kLoopStart && v0 == 0;
          const bool tb = k == kLoopEnd && v0 != 0;
          after = pc + 1 + (tf ? 8 : 0) - (tb ? 8 : 0);
          if (tb && pc < 8) after = -1;  // pc - 8 < 0 stops the run
          after = min(after, kTape2);
          nextc = after < 0 ? -1 : ((tf | tb) ? mc.next_from(after) : fall);
        } else {
          const bool z = LOGIC == kLogicSelfmove ? v0 == 0 : C::null(v0);
          const bool tf = k == kLoopStart && z;
          const bool tb = k == kLoopEnd && !z;
          if ((tb && pc != bsrc) || (tf && pc != fsrc)) {
            // Bracket scan to the matching `[` (backward, unmatched -> stop)
            // or `]` (forward, unmatched -> 128), then cached.
            int depth = 1, p = tb ? pc - 1 : pc + 1, q;
            for (;;) {
              if (tb) {
                const int ql = ml.prev_upto(p), qe = me.prev_upto(p);
                q = max(ql, qe);
                if (q < 0) break;
                depth += (q == qe) ? 1 : -1;
                p = q - 1;
              } else {
                const int ql = ml.next_from(p), qe = me.next_from(p);
                q = min(ql, qe);
                if (q >= kTape2) break;
                depth += (q == ql) ? 1 : -1;
                p = q + 1;
              }
              if (depth == 0) break;
            }
            if (tb) {
              bsrc = pc;
              bafter = q + 1;
              bnext = q < 0 ? -1 : mc.next_from(q + 1);
            } else {
              fsrc = pc;
              fafter = min(q + 1, kTape2);
              fnext = mc.next_from(fafter);
            }
          }
          after = tb ? bafter : (tf ? fafter : pc + 1);
          nextc = tb ? bnext : (tf ? fnext : fall);
          if constexpr (kAccel) {
            if (tb && nextc >= 0) {
              // A second taken jump from aq with no other bracket command in
              // between closes one iteration of a straight-line loop body of
              // L steps.  Its next iterations repeat it exactly when
              //   (a) it wrote nothing: the tape is constant, the heads
              //       advance by fixed deltas, and the loop runs until the
              //       byte under head 0 at the `]` is null; or
              //   (b) it moved no head and left the two bytes under the
              //       heads (the only ones it can write) as they were: the
              //       whole state repeats.
              // Whole iterations are then counted without being run, up to
              // the step cap and the exiting iteration, which runs normally.
              const uint32_t i1 = i + 1;
              if (pc == aq && !branched &&
                  (!wrote || (!moved && v0 == aw0 && v1 == aw1))) {
                const uint32_t len = i1 - ai;
                const uint32_t skips = nskip - askip;
                uint32_t reps = (stepcount - i1) / len;
                const int d0 = (h0 - ah0) & (kTape2 - 1);
                const int d1 = (h1 - ah1) & (kTape2 - 1);
                if (d0 != 0) {
                  // First later iteration whose `]` sees a null byte.
                  for (uint32_t j = 1; j <= (uint32_t)kTape2 && j <= reps; j++) {
                    if (C::null(t.get((h0 + j * d0) & (kTape2 - 1)))) {
                      reps = j - 1;
                      break;
                    }
                  }
                }
                i += reps * len;
                nskip += reps * skips;
                h0 = (h0 + (int)(reps * d0)) & (kTape2 - 1);
                h1 = (h1 + (int)(reps * d1)) & (kTape2 - 1);
                aq = -1;
              } else {
                aq = pc;
                ah0 = h0;
                ah1 = h1;
                ai = i1;
                askip = nskip;
                aw0 = v0;
                aw1 = v1;
                wrote = moved = branched = false;
                // Not under a command budget: cycles that start after step
                // kCycleStart have periods of ~1.5K steps and more, and a
                // budgeted run is abandoned well before it could skip one;
                // the run that completes it checks instead.
                if (!kBudget && i1 >= kCycleStart && cq != -2) {
                  bool same = pc == cq && h0 == ch0 && h1 == ch1 && cdelta == 0;
                  for (int w = 0; same && w < kTape2 / 4; w++) same = t.word(w) == aux.word(w);
                  if (same) {
                    const uint32_t len = i1 - ci;
                    const uint32_t reps = (stepcount - i1) / len;
                    i += reps * len;
                    nskip += reps * (nskip - cskip);
                    cq = -2;  // fewer than one period of steps is left
                    aq = -1;  // the loop anchor's counts predate the skip
                  } else if (cq < 0 || i1 - ci >= cwin) {
                    if (cq >= 0) {
                      cwin = cfirst ? kCycleWin : min(cwin * 2, 1u << 30);
                      cfirst = false;
                    }
                    for (int w = 0; w < kTape2 / 4; w++) aux.set_word(w, t.word(w));
                    cq = pc;
                    ch0 = h0;
                    ch1 = h1;
                    ci = i1;
                    cskip = nskip;
                    cdelta = 0;
                  }
                }
              }
            }
          }
        }
        i++;
        if (nextc < 0) break;
        // Comment run [after, nextc): each byte is one counted no-op step.
        const uint32_t rem = stepcount - i;
        const uint32_t adv = min((uint32_t)(nextc - after), rem);
        i += adv;
        nskip += adv;
        if (adv == rem || nextc >= kTape2) break;
        pc = nextc;
        if constexpr (kFullCycle) {
          if (!fdone && i >= kFullStart) {
            bool same = pc == fpc && h0 == fh0 && h1 == fh1 && fdelta == 0;
            for (int w = 0; same && w < kTape2 / 4; w++) same = t.word(w) == aux.word(w);
            if (same) {
              // The state is the anchor's: skip whole periods.
              fdone = true;
              const uint32_t len = i - fi;
              const uint32_t reps = (stepcount - i) / len;
              i += reps * len;
              nskip += reps * (nskip - fskip);
              if (i >= stepcount) break;
            } else if (fpc < 0 || i - fi >= fwin) {
              if (fpc >= 0) fwin = min(fwin * 2, 1u << 30);
              for (int w = 0; w < kTape2 / 4; w++) aux.set_word(w, t.word(w));
              fpc = pc;
              fh0 = h0;
              fh1 = h1;
              fi = i;
              fskip = nskip;
              fdelta = 0;
            }
          }
        }
        if constexpr (kStaticCycle) {
          schg |= changed;
          if (!sdone) {
            if (!schg && pc == spc && h0 == sh0 && h1 == sh1) {
              // The state is the anchor's: skip whole periods.
              sdone = true;
              const uint32_t len = i - si;
              const uint32_t reps = (stepcount - i) / len;
              i += reps * len;
              nskip += reps * (nskip - sskip);
              if (i >= stepcount) break;
            } else if (schg || i - si >= swin) {
              swin = schg ? 1 : min(swin * 2, 1u << 30);
              spc = pc;
              sh0 = h0;
              sh1 = h1;
              si = i;
              sskip = nskip;
              schg = false;
            }
          }
        }
      }
    }
    *steps = i;
    if (work) *work = iters;
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

// Table-driven step of the Forth family.  Every op below 16 is described
// by nibbles of five 64-bit constants (one nibble per opcode); bytes 16..63
// share opcode 15's descriptor (a no-op in every variant), bytes 64..127 are
// constants and 128..255 jumps.  A step is then one straight-line sequence
// for every lane, whatever its opcode:
//   x, y   = the top two stack bytes (0 where the stack is shorter)
//   m      = tape[((rsrc ? y : x) & rmask) + rbase]
//   pop npop, push v0 (and v1), write wv = (wvsrc ? m : x) to
//            tape[((wsrc ? y : x) & wmask) + wbase]
// with the reference's Stack semantics: Pop on an empty stack yields 0 and
// leaves it empty, Push on a full stack sets the overflow flag and stores
// nothing (forth.inc.h:236-249).
struct ForthOpDesc {
  int npop, npush;              // values popped / pushed
  int rsrc, rmask127, rbase64;  // read address m: source (0 x, 1 y), mask, base
  int write, wsrc, wmask127, wbase64, wvsrc;  // tape write: address, value (0 x, 1 m)
  int v0;                       // 0 m, 1 x, 2 x+1, 3 x-1, 4 x+y, 5 x-y, 6 x^64
  int v1y;                      // second push: 0 x, 1 y
  int if0, noop, reset;         // skip next if x != 0; counts as no-op; reset check
};

constexpr ForthOpDesc forth_desc(int lang, int c) {
  // forth.inc.h:251-373 (and forthtrivial_reset.cu:123-252 for `reset`).
  constexpr ForthOpDesc kRead0{1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  constexpr ForthOpDesc kRead1{1, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  constexpr ForthOpDesc kRead{1, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  constexpr ForthOpDesc kWrite{2, 0, 0, 0, 0, 1, 1, 1, 0, 0, 1, 0, 0, 0, 0};
  constexpr ForthOpDesc kWrite0{2, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 0};
  constexpr ForthOpDesc kWrite1{2, 0, 0, 0, 0, 1, 1, 0, 1, 0, 1, 0, 0, 0, 1};
  constexpr ForthOpDesc kDup{1, 2, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0};
  constexpr ForthOpDesc kDrop{1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0};
  constexpr ForthOpDesc kSwap{2, 2, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0};
  constexpr ForthOpDesc kIf0{1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 1, 0, 0};
  constexpr ForthOpDesc kInc{1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 2, 0, 0, 0, 0};
  constexpr ForthOpDesc kDec{1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 3, 0, 0, 0, 0};
  constexpr ForthOpDesc kAdd{2, 1, 0, 0, 0, 0, 0, 0, 0, 0, 4, 0, 0, 0, 0};
  constexpr ForthOpDesc kSub{2, 1, 0, 0, 0, 0, 0, 0, 0, 0, 5, 0, 0, 0, 0};
  constexpr ForthOpDesc kCopy0{1, 0, 0, 0, 0, 1, 0, 0, 1, 1, 1, 0, 0, 0, 1};
  constexpr ForthOpDesc kCopy1{1, 0, 0, 0, 1, 1, 0, 0, 0, 1, 1, 0, 0, 0, 0};
  constexpr ForthOpDesc kCopy{2, 0, 1, 1, 0, 1, 0, 1, 0, 1, 1, 0, 0, 0, 0};
  constexpr ForthOpDesc kXor{1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 6, 0, 0, 0, 0};
  constexpr ForthOpDesc kNop{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0};
  switch (c) {
    case 0: return lang == kForthcopy ? kRead : kRead0;
    case 1: return lang == kForthcopy ? kWrite : kRead1;
    case 2: return lang == kForthcopy ? kCopy : kWrite0;
    case 3: return lang == kForthcopy ? kXor : kWrite1;
    case 4: return kDup;
    case 5: return kDrop;
    case 6: return kSwap;
    case 7: return kIf0;
    case 8: return kInc;
    case 9: return kDec;
    case 10: return kAdd;
    case 11: return kSub;
    case 12: return (lang == kForthtrivial || lang == kForthtrivialReset) ? kCopy0 : kNop;
    case 13: return (lang == kForthtrivial || lang == kForthtrivialReset) ? kCopy1 : kNop;
    default: return kNop;
  }
}

// Nibble table: nibble c of the result is f(forth_desc(lang, c)).
template <class F>
constexpr uint64_t forth_nibbles(int lang, F f) {
  uint64_t r = 0;
  for (int c = 0; c < 16; c++) r |= (uint64_t)(f(forth_desc(lang, c)) & 15) << (4 * c);
  return r;
}

template <int LANG>
struct ForthLang {
  template <int LANES> using Tape = LinTape<LANES>;
  template <int LANES> using Aux = LinTape<LANES>;  // the stack; row 128 is a dummy
  static constexpr bool kReset = LANG == kForthtrivialReset;
  static constexpr uint64_t kNibPop = forth_nibbles(
      LANG, [](ForthOpDesc d) { return d.npop | d.npush << 2; });
  static constexpr uint64_t kNibRead = forth_nibbles(
      LANG, [](ForthOpDesc d) { return d.rsrc | d.rmask127 << 1 | d.rbase64 << 2 | d.wvsrc << 3; });
  static constexpr uint64_t kNibWrite = forth_nibbles(
      LANG, [](ForthOpDesc d) { return d.write | d.wsrc << 1 | d.wmask127 << 2 | d.wbase64 << 3; });
  static constexpr uint64_t kNibVal = forth_nibbles(
      LANG, [](ForthOpDesc d) { return d.v0 | d.v1y << 3; });
  static constexpr uint64_t kNibFlag = forth_nibbles(
      LANG, [](ForthOpDesc d) { return d.if0 | d.noop << 1 | d.reset << 2; });

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

  // Cycle check: the state of a run is (position, live stack [0, sp),
  // tape); stack entries at or above sp are never read before they are
  // written.  The anchor is the state right after a backward jump, with the
  // top kSnap stack entries kept in a register.  Since the anchor, the first
  // two tape bytes that writes changed are logged with their anchor values.
  // If a later backward jump lands on the anchor's position with the same
  // sp, no third tape byte changed and the logged ones hold their anchor
  // values again, pushes stayed within the top kSnap slots of the anchor's
  // stack and those slots hold the anchor's values again, the whole state
  // equals the anchor's: the run repeats with that period and whole periods
  // are skipped.  Anchors move to later jumps after windows of kCycleWin,
  // 8 kCycleWin, 64 kCycleWin, ... steps (Brent's scheme with factor 8:
  // cycles here start within a few dozen steps, so long windows early
  // catch long periods sooner), so every cycle of this kind shorter than the
  // remaining budget is found.  Other runs are not skipped (still exact).
  static constexpr uint32_t kCycleWin = 16;
  static constexpr uint32_t kCycleGrow = 8;
  static constexpr int kSnap = 4;

  // Evaluate (forth.inc.h:379-397) with EvaluateOne (forth.inc.h:251-373,
  // forthcopy.cu, forthtrivial.cu, forthtrivial_reset.cu:123-252).
  template <bool kBudget = false, class T, class A>
  __device__ static uint32_t evaluate(const T &t, const A &stk,
                                      uint32_t stepcount, uint32_t *steps,
                                      uint32_t *work = nullptr, uint32_t budget = 0) {
    constexpr int kDummy = kTape2;
    int sp = 0;
    int pos = 0;
    uint32_t i = 0, nops = 0, skipped = 0;
    bool overflow = false;
    // Anchor: position (-1 none yet, -2 cycle already skipped), sp, step
    // and no-op counts, window; since the anchor: tape changed, lowest
    // stack slot written.
    int apos = -1, asp = 0, minsp = 0;
    uint32_t ai = 0, anops = 0, awin = kCycleWin, snap = 0;
    bool tchg = false;
    int la0 = -1, la1 = -1;  // logged tape addresses (-1 free) and values
    uint32_t lv0 = 0, lv1 = 0;
    while (i < stepcount) {
      i++;
      if constexpr (kBudget) {
        if (i - skipped > budget) {
          *steps = kOverBudget;
          return 0;
        }
      }
      const uint32_t c = t.get(pos);
      const uint32_t xs = stk.get(sp > 0 ? sp - 1 : kDummy);
      const uint32_t ys = stk.get(sp > 1 ? sp - 2 : kDummy);
      const uint32_t x = sp > 0 ? xs : 0u, y = sp > 1 ? ys : 0u;
      const uint32_t sh = 4 * (c < 16 ? c : 15u);
      const bool is_const = c - 64u < 64u, is_jmp = c >= 128;
      const uint32_t np = (uint32_t)(kNibPop >> sh) & 15;
      const uint32_t rd = (uint32_t)(kNibRead >> sh) & 15;
      const uint32_t wr = (uint32_t)(kNibWrite >> sh) & 15;
      const uint32_t vl = (uint32_t)(kNibVal >> sh) & 15;
      const uint32_t fl = (uint32_t)(kNibFlag >> sh) & 15;
      const uint32_t raddr = ((rd & 1 ? y : x) & (rd & 2 ? 127u : 63u)) + (rd & 4 ? 64u : 0u);
      const uint32_t waddr = ((wr & 2 ? y : x) & (wr & 4 ? 127u : 63u)) + (wr & 8 ? 64u : 0u);
      const uint32_t m = t.get(raddr);
      const uint32_t wold = t.get(waddr);
      // Stack effect (constants push c & 63, jumps touch nothing).
      const int npop = (is_const || is_jmp) ? 0 : (int)(np & 3);
      const int npush = is_jmp ? 0 : (is_const ? 1 : (int)(np >> 2));
      // First pushed value: a chain of selects (no divergent branches).
      const uint32_t sel = is_const ? 7u : vl & 7;
      uint32_t v0 = m;
      v0 = sel == 1 ? x : v0;
      v0 = sel == 2 ? x + 1 : v0;
      v0 = sel == 3 ? x - 1 : v0;
      v0 = sel == 4 ? x + y : v0;
      v0 = sel == 5 ? x - y : v0;
      v0 = sel == 6 ? x ^ 64u : v0;
      v0 = sel == 7 ? c & 63u : v0;
      const uint32_t v1 = vl & 8 ? y : x;
      const int sp1 = sp > npop ? sp - npop : 0;
      const bool ok0 = npush >= 1 && sp1 < 128;
      const bool ok1 = npush >= 2 && ok0 && sp1 + 1 < 128;
      overflow = (npush >= 1 && !ok0) || (npush >= 2 && !ok1);
      stk.set(ok0 ? sp1 : kDummy, (uint8_t)v0);
      stk.set(ok1 ? sp1 + 1 : kDummy, (uint8_t)v1);
      sp = sp1 + (int)ok0 + (int)ok1;
      // Tape write.
      const bool wr_on = (wr & 1) && !is_const && !is_jmp;
      const uint32_t wv = rd & 8 ? m : x;
      bool do_reset = false;
      if constexpr (kReset) do_reset = wr_on && (fl & 4) && wold == 0x13;
      if (do_reset) {
        reset(t, pos);
      } else {
        t.set(wr_on ? waddr : (uint32_t)kDummy, (uint8_t)wv);
      }
      if (wr_on && (wv & 0xff) != wold && (int)waddr != la0 && (int)waddr != la1) {
        if (la0 < 0) {
          la0 = (int)waddr;
          lv0 = wold;
        } else if (la1 < 0) {
          la1 = (int)waddr;
          lv1 = wold;
        } else {
          tchg = true;
        }
      }
      tchg |= do_reset;
      minsp = min(minsp, sp1);
      nops += (fl & 2) && !is_const && !is_jmp;
      // Next position.
      int delta = 1;
      if (is_jmp) delta = (c & 64) ? -(int)(c & 63) - 1 : (int)(c & 63) + 1;
      else if ((fl & 1) && !is_const && x != 0) delta = 2;
      pos += delta;
      if (pos >= kTape2 || pos < 0 || overflow) break;
      if (is_jmp && (c & 64) && apos != -2) {
        bool same = pos == apos && sp == asp && !tchg && asp - minsp <= kSnap;
        if (same && la0 >= 0) same = t.get(la0) == lv0;
        if (same && la1 >= 0) same = t.get(la1) == lv1;
        for (int q = minsp; same && q < asp; q++)
          same = stk.get(q) == ((snap >> (8 * (asp - 1 - q))) & 0xff);
        if (same) {
          const uint32_t len = i - ai;
          const uint32_t reps = (stepcount - i) / len;
          i += reps * len;
          skipped += reps * len;
          nops += reps * (nops - anops);
          apos = -2;  // fewer than one period of steps is left
        } else if (apos < 0 || i - ai >= awin) {
          if (apos >= 0) awin = min(awin * kCycleGrow, 1u << 30);
          apos = pos;
          asp = sp;
          ai = i;
          anops = nops;
          tchg = false;
          la0 = la1 = -1;
          minsp = sp;
          snap = 0;
          for (int q = 0; q < kSnap && q < sp; q++) snap |= stk.get(sp - 1 - q) << (8 * q);
        }
      }
    }
    *steps = i;
    if (work) *work = i - skipped;
    return i - nops;
  }
};

// ---------------------------------------------------------------------------
// SUBLEQ (subleq.cu:109-133) and RSUBLEQ4 (rsubleq4.cu:116-141).  Every
// iteration counts as an op.

// Static-tape cycle skip (exact), used by SUBLEQ.  The state of a run
// is (pos, tape).  An anchor holds a position and its step count, and a
// flag says whether a write changed a byte since.  When pos comes back to
// the anchor's with no byte changed, the state is the anchor's again: the
// run repeats with period L = i - anchor step, and whole periods up to the
// step cap are counted without being run (every step of these languages
// counts as an op).  The anchor moves to the current step after any
// changing write and after windows of 1, 2, 4, ... steps (Brent), so a
// cycle on a settled tape is found within about two periods plus its lead
// in.  Runs that keep changing the tape are not skipped (still exact).
struct StaticCycle {
  int apos = -1;
  uint32_t ai = 0, win = 1;
  bool chg = false, done = false;
  // After a step that left the run at `pos` with `i` steps taken; returns
  // the steps to skip.
  __device__ __forceinline__ uint32_t after_step(int pos, uint32_t i, uint32_t stepcount,
                                                 bool changed) {
    chg |= changed;
    if (done) return 0;
    if (!chg && pos == apos) {
      done = true;
      const uint32_t len = i - ai;
      return (stepcount - i) / len * len;
    }
    if (chg || i - ai >= win) {
      win = chg ? 1 : min(win * 2, 1u << 30);
      apos = pos;
      ai = i;
      chg = false;
    }
    return 0;
  }
};

struct SubleqLang {
  template <int LANES> using Tape = LdsTape<LANES>;
  template <int LANES> using Aux = LdsTape<LANES>;  // unused
  template <bool kBudget = false, class T, class A>
  __device__ static uint32_t evaluate(const T &t, const A &, uint32_t stepcount,
                                      uint32_t *steps, uint32_t *work = nullptr,
                                      uint32_t budget = 0) {
    int pos = 0;
    uint32_t i = 0, run = 0;  // steps taken, steps executed
    StaticCycle cyc;
    while (i < stepcount) {
      i++;
      run++;
      if constexpr (kBudget) {
        if (run > budget) {
          *steps = kOverBudget;
          return 0;
        }
      }
      uint32_t sa = t.get(pos) % kTape2;
      uint32_t sb = t.get(pos + 1) % kTape2;
      const uint32_t vb = t.get(sb);
      uint8_t r = (uint8_t)(t.get(sa) - vb);
      t.set(sa, r);
      if ((r & 0x80) || r == 0) pos = t.get(pos + 2);  // read after the write
      else pos += 3;
      if (pos + 3 > kTape2 || pos < 0) break;
      i += cyc.after_step(pos, i, stepcount, vb != 0);  // changed iff b != 0
    }
    *steps = i;
    if (work) *work = run;
    return i;
  }
};
struct Rsubleq4Lang {
  template <int LANES> using Tape = LdsTape<LANES>;
  template <int LANES> using Aux = LdsTape<LANES>;  // cycle-check anchor tape
  // Full-state cycle skip (exact): the step-capped runs are periodic with
  // short periods (up to ~20 steps after lead-ins of ~20), usually with a
  // changing tape.  From kFullStart steps on, the state (pos, tape) is
  // compared with an anchor whose tape is kept in the aux region, after a
  // weighted sum of the changes since the anchor (zero when the tape is back
  // to the anchor's) filters candidates; anchors move after windows of
  // kFullStart, 2 kFullStart, ... steps (Brent).  A match means the run
  // repeats with that period: whole periods up to the step cap are counted
  // without being run.
  static constexpr uint32_t kFullStart = 8;
  __device__ __forceinline__ static uint32_t weight(uint32_t p) {
    return 2654435761u * (2u * p + 1u);
  }
  template <bool kBudget = false, class T, class A>
  __device__ static uint32_t evaluate(const T &t, const A &aux, uint32_t stepcount,
                                      uint32_t *steps, uint32_t *work = nullptr,
                                      uint32_t budget = 0) {
    int pos = 0;
    uint32_t i = 0, run = 0;  // steps taken, steps executed
    int fpos = -1;
    uint32_t fi = 0, fwin = kFullStart, fdelta = 0;
    bool fdone = false;
    while (i < stepcount) {
      i++;
      run++;
      if constexpr (kBudget) {
        if (run > budget) {
          *steps = kOverBudget;
          return 0;
        }
      }
      uint32_t sa = (uint32_t)(pos + t.get(pos)) % kTape2;
      uint32_t sb = (uint32_t)(pos + t.get(pos + 1)) % kTape2;
      uint32_t sc = (uint32_t)(pos + t.get(pos + 2)) % kTape2;
      const uint32_t old = t.get(sa);
      uint8_t r = (uint8_t)(t.get(sb) - t.get(sc));
      t.set(sa, r);
      fdelta += (r - old) * weight(sa);
      if ((r & 0x80) || r == 0) pos += (int8_t)t.get(pos + 3);  // after write
      else pos += 4;
      if (pos + 4 > kTape2 || pos < 0) break;
      if (!fdone && i >= kFullStart) {
        bool same = pos == fpos && fdelta == 0;
        for (int w = 0; same && w < kTape2 / 4; w++) same = t.word(w) == aux.word(w);
        if (same) {
          fdone = true;
          const uint32_t len = i - fi;
          i += (stepcount - i) / len * len;
        } else if (fpos < 0 || i - fi >= fwin) {
          if (fpos >= 0) fwin = min(fwin * 2, 1u << 30);
          for (int w = 0; w < kTape2 / 4; w++) aux.set_word(w, t.word(w));
          fpos = pos;
          fi = i;
          fdelta = 0;
        }
      }
    }
    *steps = i;
    if (work) *work = run;
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
