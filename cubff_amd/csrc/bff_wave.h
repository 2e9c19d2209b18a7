// Wave-wide interpreter of one BFF interaction (the standard jump logic of
// bff, bff_noheads, bff_noheads_4bit and bff_perm), for the resolver, which
// runs the long interactions one per wavefront.
//
// Evaluate (bff.inc.h:356-389) one step at a time, as the reference does,
// with no per-lane memory traffic at all: the 128-byte tape is spread over
// the wavefront's registers, bytes 4j..4j+3 in lane j (j < 32), and read
// and written at a uniform position (v_readlane, and a compare and select
// on the lane id); pc, heads and counts are uniform (scalar registers), and a command byte is
// decoded with scalar bit operations (BffClass).  A lone lane working on a
// table in LDS waits on a dependent LDS access several times per command;
// here a step is a few register moves and scalar instructions.
//
// Bracket matches (the scans of bff.inc.h:319-328 and :333-342) are cached
// per position in a second register, one byte per position: kUnknown until
// a jump from that position scans, then the target or kNone.  A write that
// makes or unmakes a bracket byte clears the cache; other writes cannot
// change a scan's result, which depends only on which bytes are brackets.
#pragma once

#include <stdint.h>

#include "languages.h"

namespace cubff {

#if defined(__HIPCC__)
// One 32-bit word per lane; word j is lane j's.  Every lane of the
// wavefront must be active wherever these are called.
struct WaveWords {
  uint32_t v;
  __device__ __forceinline__ uint32_t word(uint32_t j) const {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)j);
  }
  // Lane j takes x (a compare with the lane id and a select).
  __device__ __forceinline__ void set_word(uint32_t j, uint32_t x) {
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    v = lane == j ? x : v;
  }
  __device__ __forceinline__ void fill(uint32_t x) { v = x; }
};
#endif

template <int MAP, bool HEADS>
struct BffWave {
  using C = BffClass<MAP>;
  static constexpr uint32_t kUnknown = 0xff;
  static constexpr uint32_t kNone = 0xfe;

  struct State {
    int pc, h0, h1;
    uint32_t i, ops;
  };

  template <class R>
  __host__ __device__ __forceinline__ static uint32_t byte_at(const R &t, int p) {
    return (t.word((uint32_t)p >> 2) >> (8 * (p & 3))) & 0xff;
  }
  template <class R>
  __host__ __device__ __forceinline__ static void set_byte(R &t, int p, uint32_t b) {
    const uint32_t sh = 8 * (p & 3), w = t.word((uint32_t)p >> 2);
    t.set_word((uint32_t)p >> 2, (w & ~(0xffu << sh)) | (b << sh));
  }
  __host__ __device__ __forceinline__ static bool bracket(uint32_t b) {
    return C::ls(b) || C::le(b);
  }

  // InitialState (bff.inc.h:269-283) of a fresh tape.
  template <class R>
  __host__ __device__ __forceinline__ static State start(const R &t) {
    State s;
    if constexpr (HEADS) {
      const uint32_t w = t.word(0);
      s.h0 = (int)(w & (kTape2 - 1));
      s.h1 = (int)((w >> 8) & (kTape2 - 1));
      s.pc = 2;
    } else {
      s.h0 = s.h1 = 0;
      s.pc = 0;
    }
    s.i = s.ops = 0;
    return s;
  }

  // The scans of bff.inc.h:319-328 (forward from a '[') and :333-342
  // (backward from a ']'): the matching position, or kNone.
  template <class R>
  __host__ __device__ static uint32_t scan(const R &t, int pc, bool forward) {
    int depth = 1;
    if (forward) {
      for (int p = pc + 1; p < kTape2; p++) {
        const uint32_t b = byte_at(t, p);
        depth += C::ls(b) ? 1 : (C::le(b) ? -1 : 0);
        if (depth == 0) return (uint32_t)p;
      }
    } else {
      for (int p = pc - 1; p >= 0; p--) {
        const uint32_t b = byte_at(t, p);
        depth += C::le(b) ? 1 : (C::ls(b) ? -1 : 0);
        if (depth == 0) return (uint32_t)p;
      }
    }
    return kNone;
  }

  // Runs until the step limit, pc leaving the tape, or an unmatched jump.
  // match: the cache (every byte kUnknown for a tape it has not seen).
  template <class R>
  __host__ __device__ static void run(R &t, R &match, uint32_t stepcount, State &s) {
    int pc = s.pc, h0 = s.h0, h1 = s.h1;
    uint32_t i = s.i, ops = s.ops;
    while (i < stepcount) {
      h0 &= kTape2 - 1;
      h1 &= kTape2 - 1;
      const uint32_t b = byte_at(t, pc);
      i++;
      if (C::cmd(b)) {
        ops++;
        const uint32_t k = C::kind(b);
        if (k == kLoopStart || k == kLoopEnd) {
          const bool fwd = k == kLoopStart;
          if (C::null(byte_at(t, h0)) == fwd) {
            uint32_t m = byte_at(match, pc);
            if (m == kUnknown) {
              m = scan(t, pc, fwd);
              set_byte(match, pc, m);
            }
            // No match: forward leaves the tape past its end, backward
            // before its start (bff.inc.h:329, :343).
            pc = m != kNone ? (int)m : (fwd ? kTape2 : -1);
          }
        } else if (k >= kDec0) {
          h0 += k == kInc0 ? 1 : (k == kDec0 ? -1 : 0);
          h1 += k == kInc1 ? 1 : (k == kDec1 ? -1 : 0);
        } else {
          // + - . , : the byte written, where, and the byte it replaces.
          const int dst = k == kCopy01 ? h1 : h0;
          const uint32_t old = byte_at(t, dst);
          const uint32_t nb = k == kPlus    ? ((old + 1) & 0xff)
                              : k == kMinus ? ((old - 1) & 0xff)
                                            : byte_at(t, k == kCopy01 ? h0 : h1);
          set_byte(t, dst, nb);
          if (bracket(old) || bracket(nb)) match.fill(0xffffffffu);
        }
      }
      if (pc < 0) break;
      pc++;
      if (pc >= kTape2) break;
    }
    s.pc = pc;
    s.h0 = h0;
    s.h1 = h1;
    s.i = i;
    s.ops = ops;
  }
};

template <int LANG> struct WaveOf { static constexpr bool kHas = false; };
template <> struct WaveOf<kBff> { static constexpr bool kHas = true; using T = BffWave<kMapAscii, true>; };
template <> struct WaveOf<kBffNoheads> { static constexpr bool kHas = true; using T = BffWave<kMapAscii, false>; };
template <> struct WaveOf<kBffNoheads4bit> { static constexpr bool kHas = true; using T = BffWave<kMap4bit, false>; };
template <> struct WaveOf<kBffPerm> { static constexpr bool kHas = true; using T = BffWave<kMapPerm, true>; };

}  // namespace cubff
