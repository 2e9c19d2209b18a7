// BFF interpreter of the standard jump logic (bff, bff_noheads,
// bff_noheads_4bit, bff_perm) laid out for the shortest per-command chain
// of one lane.  An epoch's budgeted interaction lasts as long as its lane's
// commands take, so what counts is the instructions one command issues,
// not throughput.
//
// Per lane and tape position p there is one 32-bit LDS word:
//   byte 0  the tape byte,
//   byte 1  its code: class in bits 4-6 (none, head move, write, '[', ']');
//           in bits 0-3 the BffKind of a command byte, and for any other
//           byte a lower bound (1..15) on how many increments or
//           decrements it is from a command byte,
//   byte 2  the next command position after p (128 if none), kept at
//           command positions,
//   byte 3  the matching bracket of a bracket at p (0xff if unmatched).
// A command is then three word loads issued together (at pc and under the
// two heads): opcode, next command and match come with the byte at pc, the
// class of the byte a write replaces comes with the bytes under the heads.
// The run of comment bytes after a command (no-op steps, bff.inc.h:346-347)
// is skipped through byte 2.  Every command stores the low half-word under
// the written head (the byte it writes and its code; commands that write
// nothing store head 0's back), so no branch picks a store.
//
// Everything that is not the common case goes through one branch:
//   * a write that changes a byte's class, or an increment / decrement of
//     a command byte or of a byte whose bound says a command may be one
//     step away (codes are exact there; elsewhere the common path lowers
//     the bound by one): class masks, next-command bytes and matches
//     follow, and the cached jumps whose range the byte lies in are dropped;
//   * a taken jump other than the cached backward and forward one: its
//     target comes from the match byte (the scan of bff.inc.h:317-344
//     done when the table is built; after a write that moves a bracket,
//     the scan runs on 128-bit masks of '[' and ']') and is cached;
//   * a taken backward jump of a loop whose body allows loop acceleration.
//
// Loop acceleration (exact, as in the budgeted interpreter of bff_fast.h):
// a taken backward jump from the same `]` exactly one body iteration after
// the last one, whose body between the brackets holds no bracket, repeats
// the last iteration when (a) the body writes nothing -- the tape is
// constant, the heads advance by fixed deltas and the loop runs until the
// byte under head 0 at the `]` is null, found by walking head 0's orbit (at
// most 128 reads) -- or (b) the body moves no head and the bytes under the
// heads are the last iteration's: the whole state repeats.  Whole
// iterations are counted (steps and ops) without being run, up to the step
// cap and the exiting iteration, which runs normally.  "Exactly one
// iteration" is checked by the command count: a straight body runs its
// commands and the `]` between two taken visits of the `]`, and any other
// path between them runs more.  A loop whose state kept changing over
// three iterations is not looked at again until its jump is cached anew.
//
// Steps add the skipped comment bytes to the executed commands, capped at
// stepcount; ops (the value Evaluate returns, bff.inc.h:356-389) are the
// executed commands plus those of skipped iterations.
#pragma once

#include "bff_fast.h"

namespace cubff {

typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) uint16_t lds_u16;

// The per-lane words in LDS for LANES lanes: word (p, lane) at
// (p * (LANES + 1) + lane) * 4, so lanes at one position use consecutive
// words and any word is one multiply-add away.  Position 128 is the word a
// lane whose run has stopped reads.
template <int LANES>
struct LeanLds {
  static constexpr int kStride = LANES + 1;
  static constexpr int kRegionBytes = (kTape2 + 1) * kStride * 4;
  lds_u32 *p;
  __device__ __forceinline__ static LeanLds at(uint8_t *region, int lane) {
    return LeanLds{(lds_u32 *)region + lane};
  }
  __device__ __forceinline__ uint32_t word(uint32_t pos) const { return p[pos * kStride]; }
  __device__ __forceinline__ void set_word(uint32_t pos, uint32_t w) const { p[pos * kStride] = w; }
  __device__ __forceinline__ void set_low(uint32_t pos, uint32_t v) const {
    *(lds_u16 *)(p + pos * kStride) = (uint16_t)v;
  }
  __device__ __forceinline__ void set_byte(uint32_t pos, int b, uint32_t v) const {
    ((lds_u8 *)(p + pos * kStride))[b] = (uint8_t)v;
  }
};

// 32 words per lane in LDS (a lane's second tape copy), strided like
// LeanLds.
template <int LANES>
struct LeanAnchorLds {
  static constexpr int kStride = LANES + 1;
  static constexpr int kRegionBytes = (kTape2 / 4) * kStride * 4;
  lds_u32 *p;
  __device__ __forceinline__ static LeanAnchorLds at(uint8_t *region, int lane) {
    return LeanAnchorLds{(lds_u32 *)region + lane};
  }
  __device__ __forceinline__ uint32_t word(uint32_t j) const { return p[j * kStride]; }
  __device__ __forceinline__ void set_word(uint32_t j, uint32_t v) const { p[j * kStride] = v; }
};

template <int MAP, bool HEADS>
struct BffLean {
  using C = BffClass<MAP>;
  // Per-kind fields, 2 bits each, indexed by BffKind: head-0 delta + 1,
  // head-1 delta + 1, byte delta + 1 (kNull and kNoop change nothing).
  static constexpr uint32_t kTH0 = 0x558555;
  static constexpr uint32_t kTH1 = 0x585555;
  static constexpr uint32_t kTDelta = 0x555525;
  static constexpr uint32_t kNoMatch = 0xff;

  __host__ __device__ __forceinline__ static uint32_t class_of_kind(uint32_t k) {
    return k == kLoopStart ? kClsOpen
                           : (k == kLoopEnd ? kClsClose
                                            : ((k - kPlus) < 4u ? kClsWrite
                                                                : ((k - kDec0) < 4u ? kClsMove : kClsNone)));
  }
  // Byte 1 of a position's word for tape byte v: kind and class of a
  // command byte; for any other byte the number of increments or
  // decrements to the nearest command byte, at most 15.
  __host__ __device__ static uint32_t code_of(uint32_t v) {
    if (C::cmd(v)) {
      const uint32_t k = C::kind(v);
      return k | (class_of_kind(k) << 4);
    }
    uint32_t d = 1;
    while (d < 15 && !C::cmd((v + d) & 0xff) && !C::cmd((v - d) & 0xff)) d++;
    return d;
  }

  struct Run {
    BffMasks m;                          // class masks of the tape
    int pc, h0, h1;
    uint32_t i, iters, extra;
    bool done;
    // Jump caches: the last taken backward / forward jump's source, next
    // command and steps (the jump itself and the comment bytes after its
    // target); for the backward one, whether its body allows acceleration,
    // holds a write, and the commands of one iteration.
    int bsrc, bq, bnext, binc, fsrc, fq, fnext, finc;
    bool bacc, bwrite;
    uint32_t bcmds;
    // The backward one's body is straight and short enough to be run
    // iteration by iteration without decoding (complete runs only).
    bool bcomp;
    // Whether the match bytes are the tape's (until a bracket is written).
    uint32_t mvalid;
    // Loop anchor: its `]`, heads, steps and commands after it, the bytes
    // under the heads; failed comparisons in a row.
    int aq, ah0, ah1;
    uint32_t ai, aiters, aw0, aw1, afail;
    // Cycle anchor (complete runs): its `]` (-1 none yet, -2 no more
    // checks), heads, steps and ops after it, window, and the weighted sum
    // of the tape changes since.
    int cq, ch0, ch1;
    uint32_t ci, cops, cwin, cdelta, cfirst;
  };

  // The match of the bracket at pc from the masks: backward from a ']'
  // (tb), forward from a '['; kNoMatch if there is none.
  __host__ __device__ static uint32_t scan_match(const BffMasks &m, int pc, bool tb) {
    int depth = 1, p = tb ? pc - 1 : pc + 1, q;
    for (;;) {
      if (tb) {
        const int ql = m.open.prev_upto(p), qe = m.close.prev_upto(p);
        q = fast_imax(ql, qe);
        if (q < 0) break;
        depth += (q == qe) ? 1 : -1;
        p = q - 1;
      } else {
        const int ql = m.open.next_from(p), qe = m.close.next_from(p);
        q = fast_imin(ql, qe);
        if (q >= kTape2) break;
        depth += (q == ql) ? 1 : -1;
        p = q + 1;
      }
      if (depth == 0) return (uint32_t)q;
    }
    return kNoMatch;
  }

  // Matches of every bracket (byte 3), from the masks.
  template <class W>
  __host__ __device__ static void match_all(const W &w, const BffMasks &m) {
    Mask128 br{m.open.lo | m.close.lo, m.open.hi | m.close.hi};
    Mask128 un{0, 0};  // '[' not matched yet
    while (br.lo | br.hi) {
      const int p = br.lo ? __builtin_ctzll(br.lo) : 64 + __builtin_ctzll(br.hi);
      br.flip(p, true);
      const bool open = p < 64 ? ((m.open.lo >> p) & 1) : ((m.open.hi >> (p - 64)) & 1);
      if (open) {
        un.flip(p, true);
        w.set_byte(p, 3, kNoMatch);
      } else {
        const int q = un.prev_upto(p - 1);
        if (q >= 0) {
          un.flip(q, true);
          w.set_byte(q, 3, (uint32_t)p);
        }
        w.set_byte(p, 3, q >= 0 ? (uint32_t)q : kNoMatch);
      }
    }
  }

  // Words, masks and matches of a tape whose raw bytes are in the words
  // of positions 0..31 (bytes 4j..4j+3 of the tape are the four bytes of
  // word j), and the initial state.  lut(v) is code_of(v) from a table.
  // Groups of four positions are built from the top down: raw word j is
  // read before positions 4j..4j+3 are written, and the raw words those
  // positions held (4j..4j+3 > j for j > 0) were read earlier.
  template <class W, class L>
  __host__ __device__ __forceinline__ static void start_raw(Run &r, const W &w, const L &lut,
                                                            uint32_t stepcount) {
    BffMasks m{};
    uint32_t next = kTape2;
    uint32_t first_word = 0;
#pragma unroll 2
    for (int j = kTape2 / 4 - 1; j >= 0; j--) {
      const uint32_t raw = w.word(j);
      if (j == 0) first_word = raw;
      uint32_t code[4];
#pragma unroll
      for (int b = 0; b < 4; b++) code[b] = lut((raw >> (8 * b)) & 0xff);
      uint32_t nib[5] = {0, 0, 0, 0, 0};  // per class: the four positions' bits
#pragma unroll
      for (int b = 3; b >= 0; b--) {
        const int p = 4 * j + b;
        const uint32_t cls = code[b] >> 4;
        w.set_word(p, ((raw >> (8 * b)) & 0xff) | (code[b] << 8) | (next << 16) | (kNoMatch << 24));
        next = cls != kClsNone ? (uint32_t)p : next;
        nib[kClsMove] |= (uint32_t)(cls == kClsMove) << b;
        nib[kClsWrite] |= (uint32_t)(cls == kClsWrite) << b;
        nib[kClsOpen] |= (uint32_t)(cls == kClsOpen) << b;
        nib[kClsClose] |= (uint32_t)(cls == kClsClose) << b;
      }
      // Positions 4j..4j+3 enter each mask at the bottom.
      const auto push = [](Mask128 &x, uint32_t n) {
        x.hi = (x.hi << 4) | (x.lo >> 60);
        x.lo = (x.lo << 4) | n;
      };
      push(m.move, nib[kClsMove]);
      push(m.write, nib[kClsWrite]);
      push(m.open, nib[kClsOpen]);
      push(m.close, nib[kClsClose]);
    }
    m.cmd = Mask128{m.move.lo | m.write.lo | m.open.lo | m.close.lo,
                    m.move.hi | m.write.hi | m.open.hi | m.close.hi};
    // Position 128: where a lane that has stopped parks.
    w.set_word(kTape2, (kNoop << 8) | ((uint32_t)kTape2 << 16) | (kNoMatch << 24));
    r.m = m;
    if ((m.open.lo | m.open.hi | m.close.lo | m.close.hi) != 0) match_all(w, m);
    init(r, first_word & 0xff, (first_word >> 8) & 0xff, stepcount);
  }

  __host__ __device__ __forceinline__ static void init(Run &r, uint32_t b0, uint32_t b1,
                                                       uint32_t stepcount) {
    int pc;
    if constexpr (HEADS) {  // InitialState, bff.inc.h:269-275
      r.h0 = b0 & (kTape2 - 1);
      r.h1 = b1 & (kTape2 - 1);
      pc = 2;
    } else {  // bff.inc.h:277-282
      r.h0 = 0;
      r.h1 = 0;
      pc = 0;
    }
    r.iters = 0;
    r.extra = 0;
    const int first = r.m.cmd.next_from(pc);
    r.i = fast_umin((uint32_t)(first - pc), stepcount);
    r.pc = first;
    r.done = !(r.i < stepcount && first < kTape2);
    r.bsrc = -1; r.bq = -1; r.bnext = 0; r.binc = 0;
    r.fsrc = -1; r.fq = kTape2; r.fnext = 0; r.finc = 0;
    r.bacc = r.bwrite = r.bcomp = false;
    r.bcmds = 0;
    r.mvalid = 1;
    r.aq = -1; r.ah0 = 0; r.ah1 = 0;
    r.ai = 0; r.aiters = 0; r.aw0 = 0; r.aw1 = 0; r.afail = 0;
    r.cq = -1; r.ch0 = 0; r.ch1 = 0;
    r.ci = 0; r.cops = 0; r.cwin = kCycleFirstWin; r.cdelta = 0; r.cfirst = 1;
  }

  // Full-state cycle skip of complete runs (exact, as in bff_fast.h): from
  // step kCycleStart on, the state right after a taken backward jump is
  // compared with an anchor whose tape is kept in a second per-lane region:
  // pc, heads, a weighted sum of the tape changes since the anchor (zero
  // when the tape is back to the anchor's; the weights only filter) and,
  // when those match, every tape byte.  A match means the run repeats with
  // that period: whole periods are counted without being run.  The first
  // anchor stays for kCycleFirstWin steps, later ones follow Brent's scheme
  // with windows kCycleWin, 2 kCycleWin, ...
  static constexpr uint32_t kCycleStart = 128;
  static constexpr uint32_t kCycleFirstWin = 4096;
  static constexpr uint32_t kCycleWin = 1024;

  // A second region of 32 words per lane: the anchor's tape.
  struct NoAnchor {
    __host__ __device__ uint32_t word(uint32_t) const { return 0; }
    __host__ __device__ void set_word(uint32_t, uint32_t) const {}
  };
  template <class W>
  __host__ __device__ __forceinline__ static uint32_t tape_word(const W &w, uint32_t j) {
    return (w.word(4 * j) & 0xff) | ((w.word(4 * j + 1) & 0xff) << 8) |
           ((w.word(4 * j + 2) & 0xff) << 16) | (w.word(4 * j + 3) << 24);
  }

  // Whether `c` holds in any lane of the wavefront (the lane itself on the
  // host, where runs go one at a time).
  __host__ __device__ __forceinline__ static bool any_lane(bool c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(c) != 0;
#else
    return c;
#endif
  }

  // Straight loop bodies of up to kStraightMax commands run without
  // decoding in complete runs.
  static constexpr uint32_t kStraightMax = 8;

  // Whole iterations of the straight loop (no bracket in its body) whose
  // `]` at `close` (matching the `[` at q) was just taken: the nc body
  // commands from pos0 run from a list of their kinds read once, without
  // the per-command decode, next-command and jump work.  On entry I and IT
  // are the steps and commands after the `]` (the next command is pos0);
  // each iteration adds len = close - pos0 + binc steps and nc + 1
  // commands.  It stops, with the state exact, where the general loop must
  // take over: before an iteration that would reach the step cap or pass
  // the command limit lim1 (at pos0), before a write into the loop's own
  // code or one that changes a byte's class (at that command), and at the
  // `]` when the byte under head 0 is null (the exit) or the cycle skip's
  // check is due there (the anchor state cq/ch0/ch1 with no change since,
  // no anchor yet, or the anchor's window passed).  Returns the position
  // of the next command; heads, I, IT and cdelta (the cycle skip's
  // weighted change sum) follow.
  template <class W, class L>
  __host__ __device__ static int run_straight(const W &w, const L &lut, int close, int q, int pos0,
                                              uint32_t binc, uint32_t nc, uint32_t stepcount,
                                              uint32_t lim1, int cq, int ch0, int ch1,
                                              uint32_t ci, uint32_t cwin, int &h0, int &h1,
                                              uint32_t &I, uint32_t &IT, uint32_t &cdelta) {
    uint32_t kinds = 0, plo = 0, phi = 0;
    int p = pos0;
    for (uint32_t j = 0; j < nc; j++) {
      const uint32_t wp = w.word(p);
      kinds |= ((wp >> 8) & 15) << (4 * j);
      if (j < 4) plo |= (uint32_t)p << (7 * j);
      else phi |= (uint32_t)p << (7 * (j - 4));
      p = (int)((wp >> 16) & 0xff);
    }
    if (p != close) return pos0;
    const uint32_t len = (uint32_t)(close - pos0) + binc;
    for (;;) {
      if (I + len >= stepcount || (int)(lim1 - (IT + nc + 1)) < 0) return pos0;
      for (uint32_t j = 0; j < nc; j++) {
        const uint32_t k = (kinds >> (4 * j)) & 15;
        if ((k - kPlus) < 4u) {
          const bool c01 = k == kCopy01;
          const int a = c01 ? h1 : h0;
          const int pj = (int)((j < 4 ? plo >> (7 * j) : phi >> (7 * (j - 4))) & 127);
          const uint32_t wold = w.word(a);
          uint32_t low;
          bool stop = a >= q && a <= close;
          if (k == kPlus || k == kMinus) {
            // A comment byte whose bound is 1 or less gets its exact code
            // (still a comment: the class stays); a command byte stops.
            const uint32_t vnew = ((wold & 0xff) + (k == kPlus ? 1u : 0xffu)) & 0xff;
            const uint32_t code = (wold >> 8) & 0x7f;
            uint32_t cnew = code - 1;
            if (code < 2) {
              cnew = lut(vnew);
              stop = stop || (cnew >> 4) != kClsNone;
            }
            stop = stop || code >= 16;
            low = vnew | (cnew << 8);
          } else {
            const uint32_t wsrc = w.word(c01 ? h0 : h1);
            stop = stop || ((wsrc >> 12) & 7) != ((wold >> 12) & 7);
            low = wsrc & 0xffff;
          }
          if (stop) {
            I += (uint32_t)(pj - pos0);
            IT += j;
            return pj;
          }
          w.set_low(a, low);
          const int d = (int)(low & 0xff) - (int)(wold & 0xff);
          cdelta += (uint32_t)(d * (int)((2u * (uint32_t)a + 1u) * 16411u));
        } else {
          h0 = (h0 + (int)((kTH0 >> (2 * k)) & 3) - 1) & (kTape2 - 1);
          h1 = (h1 + (int)((kTH1 >> (2 * k)) & 3) - 1) & (kTape2 - 1);
        }
      }
      const uint32_t ib = I + (uint32_t)(close - pos0) + 1;  // steps after the `]`
      const bool due = cq != -2 && ib >= kCycleStart &&
                       ((close == cq && h0 == ch0 && h1 == ch1 && cdelta == 0) || cq < 0 ||
                        ib - ci >= cwin);
      if (due || C::null(w.word(h0) & 0xff)) {
        I += (uint32_t)(close - pos0);
        IT += nc;
        return close;
      }
      I += len;
      IT += nc + 1;
    }
  }

  // Runs commands until the run ends (r.done) or `budget` (>= 1) has been
  // spent in this call: commands plus the tape reads of loop
  // acceleration's orbit walks.  With RESUME a later call goes on from the
  // next command; without, a run stopped by the budget is only reported
  // (r.done false).
  //
  // The wavefront stays in the loop while any lane runs: a lane that has
  // stopped parks on position 128, whose word is a comment with no next
  // command, so its loads, store and updates change nothing.  The common
  // path is straight-line code (selects, no branch, no lane leaving the
  // loop: a per-lane loop exit measured ~1.6x slower per command); the
  // rare cases are one branch, which a wavefront skips when none of its
  // lanes needs it.
  // On the device, a run also stops (as if its budget were spent) once the
  // wall clock passes `deadline`, checked every 32 commands: the wavefront
  // leaves the loop before its next command, where every lane's state is
  // whole.  The check is laid out off the loop's straight path; in line it
  // cost ~10% per command (a taken branch every command), off it ~5%.
  template <bool RESUME, bool CYCLE = false, class W, class L, class A = NoAnchor>
  __host__ __device__ __forceinline__ static void advance(Run &r, const W &w, const L &lut,
                                                          uint32_t stepcount, uint32_t budget,
                                                          const A &anchor = A{},
                                                          uint64_t deadline = ~0ull) {
    BffMasks m = r.m;
    int pc = r.done ? kTape2 : r.pc, h0 = r.h0, h1 = r.h1;
    int rpc = r.done ? kTape2 : r.pc;
    uint32_t i = r.i, iters = r.iters, extra = r.extra;
    int bsrc = r.bsrc, bq = r.bq, bnext = r.bnext, binc = r.binc;
    int fsrc = r.fsrc, fq = r.fq, fnext = r.fnext, finc = r.finc;
    uint32_t bacc = r.bacc ? 1u : 0u;
    uint32_t bcomp = r.bcomp ? 1u : 0u;
    uint32_t bwrite = r.bwrite ? 1u : 0u;
    uint32_t bcmds = r.bcmds, mvalid = r.mvalid;
    int aq = r.aq, ah0 = r.ah0, ah1 = r.ah1;
    uint32_t ai = r.ai, aiters = r.aiters, aw0 = r.aw0, aw1 = r.aw1, afail = r.afail;
    int cq = r.cq, ch0 = r.ch0, ch1 = r.ch1;
    uint32_t ci = r.ci, cops = r.cops, cwin = r.cwin, cdelta = r.cdelta, cfirst = r.cfirst;
    // i >= stepcount and iters >= lim as sign bits (all stay below 2^31);
    // the tape reads of orbit walks lower lim.
    const uint32_t sc1 = stepcount - 1;
    uint32_t lim1 = iters + fast_umin(budget, 1u << 30) - 1;
    int next = rpc;
    uint32_t tick = 0;
    while (any_lane(pc < kTape2)) {
#if defined(__HIP_DEVICE_COMPILE__)
      if (__builtin_expect(((++tick) & 31) == 0, 0) && wall_clock64() > deadline) break;
#else
      (void)tick;
      (void)deadline;
#endif
      const uint32_t wp = w.word(pc);
      const uint32_t w0 = w.word(h0);
      const uint32_t w1 = w.word(h1);
      keep_loads_together();
      const uint32_t k = (wp >> 8) & 15;
      const uint32_t k2 = 2 * k;
      const uint32_t nx = (wp >> 16) & 0xff;
      const bool c01 = k == kCopy01, c10 = k == kCopy10;
      const uint32_t wsrc = c10 ? w1 : w0;  // the word whose byte is written
      const uint32_t wold = c01 ? w1 : w0;  // the word written over
      const int waddr = c01 ? h1 : h0;
      const bool pm = (k - kPlus) < 2u;
      const uint32_t vnew = ((wsrc & 0xff) + ((kTDelta >> k2) & 3) - 1) & 0xff;
      // The code moves with a copied byte; an increment or decrement
      // lowers a non-command byte's bound by one.
      w.set_low(waddr, vnew | ((wsrc & 0xff00) - (pm ? 0x100u : 0u)));
      if constexpr (CYCLE) {
        // Weight (2p + 1) * 16411 < 2^23 of the written position p: the
        // product with a byte change stays a 24-bit multiply.
        const int d = (int)vnew - (int)(wold & 0xff);
        cdelta += (uint32_t)(d * (int)((2u * (uint32_t)waddr + 1u) * 16411u));
      }
      const uint32_t cold = (wold >> 12) & 7, csrc = (wsrc >> 12) & 7;
      // An increment or decrement is rare when its byte is a command or
      // may be one step from one (code < 2 or >= 16).
      const bool wrare = (csrc != cold) | (pm & (((wold >> 8) & 0x7f) - 2 >= 14u));
      const uint32_t v0 = w0 & 0xff, v1 = w1 & 0xff;
      h0 = (h0 + (int)((kTH0 >> k2) & 3) - 1) & (kTape2 - 1);
      h1 = (h1 + (int)((kTH1 >> k2) & 3) - 1) & (kTape2 - 1);
      // t: 1 for a '[' over a null byte, 2 for a ']' over a non-null one.
      const uint32_t t = k2 + (C::null(v0) ? 1u : 0u);
      const bool tb = t == 2;
      const bool taken = (t - 1) < 2u;
      const int jsrc = tb ? bsrc : fsrc;
      next = taken ? (tb ? bnext : fnext) : (int)nx;
      uint32_t inc = taken ? (uint32_t)(tb ? binc : finc) : nx - (uint32_t)pc;
      const uint32_t miss_b = taken ? (uint32_t)(pc ^ jsrc) : 0u;
      bool rare = wrare | (miss_b != 0) | (tb & (bacc != 0));
      if constexpr (CYCLE) rare = rare | (tb & (bcomp != 0) & (bacc == 0));
      if constexpr (CYCLE) {
        // A cycle check is due after a taken backward jump once the anchor
        // state may have recurred or its window has passed.
        const bool due = ((pc == cq) & (h0 == ch0) & (h1 == ch1) & (cdelta == 0)) |
                         ((cq == -1) | (i + 1 - ci >= cwin));
        rare = rare | (tb & (i + 1 >= kCycleStart) & (cq != -2) & due);
      }
      {
        if (__builtin_expect(rare, 0)) {
          if (wrare) {
            // A write only (no jump).  Its code, then what the class change
            // moves: masks, next-command bytes, matches, the cached jumps
            // it falls in.
            const uint32_t code = pm ? lut(vnew) : ((wsrc >> 8) & 0xff);
            const uint32_t cnew = code >> 4;
            if (pm) w.set_byte(waddr, 1, code);
            if (cnew != cold) {
              const bool was = cold != kClsNone, is = cnew != kClsNone;
              m.cmd.flip(waddr, was != is);
              m.open.flip(waddr, (cold == kClsOpen) != (cnew == kClsOpen));
              m.close.flip(waddr, (cold == kClsClose) != (cnew == kClsClose));
              m.write.flip(waddr, (cold == kClsWrite) != (cnew == kClsWrite));
              m.move.flip(waddr, (cold == kClsMove) != (cnew == kClsMove));
              if (was != is) {
                const int prev = m.cmd.prev_upto(waddr - 1);
                const int after = m.cmd.next_from(waddr + 1);
                if (prev >= 0) w.set_byte(prev, 2, is ? (uint32_t)waddr : (uint32_t)after);
                if (is) w.set_byte(waddr, 2, (uint32_t)after);
              }
              if (cold >= kClsOpen || cnew >= kClsOpen) {
                // The brackets moved: every match may have changed.  From
                // here on jumps find their match in the masks.
                mvalid = 0;
                bsrc = -1;
                fsrc = -1;
              }
              // A cached jump stands unless the byte lies in the body of
              // the backward one (its classes and command count) or between
              // the forward one's match and its next command.
              if (waddr > bq && waddr < bsrc) bsrc = -1;
              if (waddr > fq && waddr <= fnext) fsrc = -1;
              if (bsrc < 0) aq = -1;
              next = m.cmd.next_from(pc + 1);
              inc = (uint32_t)(next - pc);
            }
          } else {
            if (miss_b != 0) {
              // The jump's target: the next command after its match (none:
              // past the tape), or for an unmatched bracket the end of the
              // run after this step.  The match is the match byte while
              // that is valid, else the bracket scan of bff.inc.h:317-344
              // done on the masks.
              const uint32_t q = mvalid ? (wp >> 24) : scan_match(m, pc, tb);
              if (q == kNoMatch) {
                next = kTape2;
                inc = 1;
              } else {
                next = (int)((w.word(q) >> 16) & 0xff);
                inc = (uint32_t)next - q;
              }
              if (tb) {
                bsrc = pc;
                bq = q == kNoMatch ? -1 : (int)q;
                bnext = next;
                binc = (int)inc;
                bacc = 0;
                bcomp = 0;
                if (q != kNoMatch) {
                  const Mask128 body = mask_between((int)q, pc);
                  const bool straight = !mask_any(m.open, body) && !mask_any(m.close, body);
                  const bool bw = mask_any(m.write, body), bmove = mask_any(m.move, body);
                  bwrite = bw ? 1u : 0u;
                  bacc = straight && !(bw && bmove) ? 1u : 0u;
                  const Mask128 cm{m.cmd.lo & body.lo, m.cmd.hi & body.hi};
                  bcmds = (uint32_t)(__builtin_popcountll(cm.lo) + __builtin_popcountll(cm.hi)) + 1;
                  bcomp = straight && bcmds >= 2 && bcmds <= kStraightMax + 1 ? 1u : 0u;
                }
                afail = 0;
              } else {
                fsrc = pc;
                fq = q == kNoMatch ? kTape2 : (int)q;
                fnext = next;
                finc = (int)inc;
              }
            }
            bool skipped = false;
            if (tb && bacc) {
              const uint32_t ib = i + 1;  // steps after this `]`
              if (pc == aq && iters + 1 == aiters + bcmds &&
                  (!bwrite || (v0 == aw0 && v1 == aw1))) {
                // One more iteration with the last one's state advanced by
                // fixed deltas: skip whole iterations.
                const uint32_t len = ib - ai;
                uint32_t reps = (stepcount - ib) / len;
                const int d0 = (h0 - ah0) & (kTape2 - 1);
                const int d1 = (h1 - ah1) & (kTape2 - 1);
                if (d0 != 0) {
                  // First later iteration whose `]` sees a null byte (the
                  // tape is constant: no write in the body), 8 at a time.
                  const uint32_t lim = fast_umin(reps, (uint32_t)kTape2);
                  uint32_t j = 1;
                  bool found = false;
                  for (; j <= lim; j += 8) {
                    uint32_t hit = 0;
#pragma unroll
                    for (int u = 0; u < 8; u++)
                      hit |= (uint32_t)C::null(w.word((h0 + (j + u) * d0) & (kTape2 - 1)) & 0xff)
                             << u;
                    if (lim - j < 7) hit &= (1u << (lim - j + 1)) - 1;
                    if (hit) {
                      j += __builtin_ctz(hit);
                      found = true;
                      break;
                    }
                  }
                  if (found) reps = j - 1;
                  lim1 -= found ? j : lim + 1;
                }
                i += reps * len;
                extra += reps * bcmds;
                h0 = (h0 + (int)(reps * d0)) & (kTape2 - 1);
                h1 = (h1 + (int)(reps * d1)) & (kTape2 - 1);
                aq = -1;
                skipped = true;
              } else {
                // A new anchor; a loop whose state changed over three
                // iterations in a row is left alone.
                afail = pc == aq ? afail + 1 : 0;
                if (afail >= 3) bacc = 0;
                aq = pc;
                ah0 = h0;
                ah1 = h1;
                ai = ib;
                aiters = iters + 1;
                aw0 = v0;
                aw1 = v1;
              }
            }
            if constexpr (CYCLE) {
              const uint32_t ib = i + 1;  // steps after this `]`
              if (tb && !skipped && next < kTape2 && ib >= kCycleStart && cq != -2) {
                bool same = pc == cq && h0 == ch0 && h1 == ch1 && cdelta == 0;
                for (uint32_t j = 0; same && j < kTape2 / 4; j++)
                  same = tape_word(w, j) == anchor.word(j);
                if (same) {
                  const uint32_t len = ib - ci;
                  const uint32_t reps = (stepcount - ib) / len;
                  i += reps * len;
                  extra += reps * (iters + 1 + extra - cops);
                  cq = -2;  // fewer than one period of steps is left
                  aq = -1;  // the loop anchor's counts predate the skip
                  skipped = true;
                } else if (cq < 0 || ib - ci >= cwin) {
                  if (cq >= 0) {
                    cwin = cfirst ? kCycleWin : (cwin < (1u << 29) ? cwin * 2 : cwin);
                    cfirst = 0;
                  }
                  for (uint32_t j = 0; j < kTape2 / 4; j++) anchor.set_word(j, tape_word(w, j));
                  cq = pc;
                  ch0 = h0;
                  ch1 = h1;
                  ci = ib;
                  cops = iters + 1 + extra;
                  cdelta = 0;
                }
              }
              // A straight loop that acceleration does not take: its next
              // iterations run here without decoding.
              if (tb && !skipped && bcomp && !bacc && bsrc == pc && next == bnext) {
                uint32_t I = i + inc, IT = iters + 1;
                const int out = run_straight(w, lut, pc, bq, bnext, (uint32_t)binc, bcmds - 1, stepcount,
                                             lim1, cq, ch0, ch1, ci, cwin, h0, h1, I, IT, cdelta);
                next = out;
                inc = I - i;
                iters = IT - 1;
                aq = -1;
              }
            }
          }
        }
      }
      i += inc;
      const bool act = pc < kTape2;
      iters += act ? 1u : 0u;
      rpc = act ? next : rpc;
      const uint32_t stop = ((uint32_t)next >> 7) | ((sc1 - i) >> 31) | ((lim1 - iters) >> 31);
      pc = (int)fast_umin((uint32_t)next + (stop << 8), (uint32_t)kTape2);
    }
    next = rpc;
    r.m = m;
    r.h0 = h0; r.h1 = h1; r.pc = next;
    r.i = i; r.iters = iters; r.extra = extra;
    r.done = next >= kTape2 || i >= stepcount;
    r.bsrc = bsrc; r.bq = bq; r.bnext = bnext; r.binc = binc;
    r.fsrc = fsrc; r.fq = fq; r.fnext = fnext; r.finc = finc;
    r.bacc = bacc != 0; r.bwrite = bwrite != 0; r.bcmds = bcmds; r.mvalid = mvalid;
    r.bcomp = bcomp != 0;
    r.aq = aq; r.ah0 = ah0; r.ah1 = ah1;
    r.ai = ai; r.aiters = aiters; r.aw0 = aw0; r.aw1 = aw1; r.afail = afail;
    r.cq = cq; r.ch0 = ch0; r.ch1 = ch1;
    r.ci = ci; r.cops = cops; r.cwin = cwin; r.cdelta = cdelta; r.cfirst = cfirst;
  }

  // Results of an ended run: ops (the reference's count of executed
  // non-no-op steps), *steps and *work (commands executed).
  __host__ __device__ __forceinline__ static uint32_t finish(const Run &r, uint32_t stepcount,
                                                             uint32_t *steps, uint32_t *work) {
    *steps = fast_umin(r.i, stepcount);
    if (work) *work = r.iters;
    return r.iters + r.extra;
  }

  // A complete run (no budget) with the full-state cycle skip; anchor is a
  // second region of 32 words per lane.
  template <class W, class L, class A>
  __host__ __device__ static uint32_t evaluate_full(const W &w, const A &anchor, const L &lut,
                                                    uint32_t stepcount, uint32_t *steps,
                                                    uint32_t *work) {
    Run r;
    start_raw(r, w, lut, stepcount);
    advance<false, true>(r, w, lut, stepcount, 0xffffffffu, anchor);
    return finish(r, stepcount, steps, work);
  }

  // What a run stopped by its budget needs, besides its tape bytes, to go
  // on from a fresh table: the next command's position and the heads
  // (pos = pc | h0 << 8 | h1 << 16), steps, commands and accelerated ones.
  struct Resume {
    uint32_t pos, i, iters, extra;
  };
  __host__ __device__ __forceinline__ static Resume saved(const Run &r) {
    return Resume{(uint32_t)r.pc | ((uint32_t)r.h0 << 8) | ((uint32_t)r.h1 << 16), r.i, r.iters,
                  r.extra};
  }
  // A run that goes on from `s` over the tape whose raw bytes are in the
  // words of positions 0..31: the table, masks and matches are built from
  // the bytes (as start_raw does), the jump caches and accelerations start
  // empty (they only save work), and the next command, heads and counts
  // are the stopped run's.
  template <class W, class L>
  __host__ __device__ __forceinline__ static void start_resume(Run &r, const W &w, const L &lut,
                                                               uint32_t stepcount,
                                                               const Resume &s) {
    start_raw(r, w, lut, stepcount);
    resume_from(r, s, stepcount);
  }
  // The second half of start_resume, on a run start_raw began.
  __host__ __device__ __forceinline__ static void resume_from(Run &r, const Resume &s,
                                                              uint32_t stepcount) {
    r.pc = (int)(s.pos & 0xff);
    r.h0 = (int)((s.pos >> 8) & 0xff);
    r.h1 = (int)((s.pos >> 16) & 0xff);
    r.i = s.i;
    r.iters = s.iters;
    r.extra = s.extra;
    r.done = !(r.i < stepcount && r.pc < kTape2);
  }

  // The complete run (with the full-state cycle skip) of a run stopped by
  // its budget, from `s` and its tape's raw words.
  template <class W, class L, class A>
  __host__ __device__ static uint32_t evaluate_resumed(const W &w, const A &anchor, const L &lut,
                                                       uint32_t stepcount, const Resume &s,
                                                       uint32_t *steps, uint32_t *work) {
    Run r;
    start_resume(r, w, lut, stepcount, s);
    advance<false, true>(r, w, lut, stepcount, 0xffffffffu, anchor);
    return finish(r, stepcount, steps, work);
  }

  // Evaluate (bff.inc.h:356-389) of the tape whose raw bytes are in the
  // words of positions 0..31 (see start_raw), budgeted: a run that would
  // spend more than `budget` stops and reports *steps = kOverBudget.  The
  // final tape is byte 0 of the words.
  // With `stopped`, a run stopped by the budget (or the deadline) also
  // leaves there what start_resume needs to go on from it.
  template <class W, class L>
  __host__ __device__ static uint32_t evaluate(const W &w, const L &lut, uint32_t stepcount,
                                               uint32_t *steps, uint32_t *work, uint32_t budget,
                                               Resume *stopped = nullptr,
                                               uint64_t deadline = ~0ull) {
    Run r;
    start_raw(r, w, lut, stepcount);
    advance<false>(r, w, lut, stepcount, budget, NoAnchor{}, deadline);
    if (!r.done) {
      if (stopped) *stopped = saved(r);
      *steps = kOverBudget;
      return 0;
    }
    return finish(r, stepcount, steps, work);
  }
};

}  // namespace cubff

namespace cubff {

// The languages BffLean runs (the standard jump logic), with their map.
template <int LANG> struct LeanOf { static constexpr bool kHas = false; };
template <> struct LeanOf<kBff> { static constexpr bool kHas = true; using T = BffLean<kMapAscii, true>; };
template <> struct LeanOf<kBffNoheads> { static constexpr bool kHas = true; using T = BffLean<kMapAscii, false>; };
template <> struct LeanOf<kBffNoheads4bit> { static constexpr bool kHas = true; using T = BffLean<kMap4bit, false>; };
template <> struct LeanOf<kBffPerm> { static constexpr bool kHas = true; using T = BffLean<kMapPerm, true>; };

#if defined(__HIPCC__)
// A workgroup's code table (BffLean::code_of for every byte value); every
// thread of the block calls it before a barrier.
template <int LANG>
__device__ __forceinline__ void fill_lean_lut(uint8_t *lut) {
  if constexpr (LeanOf<LANG>::kHas)
    for (uint32_t v = threadIdx.x; v < 256; v += blockDim.x)
      lut[v] = (uint8_t)LeanOf<LANG>::T::code_of(v);
}

// A soup row into half h of a lane's raw tape (words 16h .. 16h + 15 of
// positions 0..31, see BffLean::start_raw), with 16-byte loads, or with
// 8-byte device-coherent loads.
template <int LANES>
__device__ __forceinline__ void load_raw_row(const LeanLds<LANES> &w, int half, const uint8_t *src) {
  const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint4 v = s4[q];
    w.set_word(16 * half + 4 * q, v.x);
    w.set_word(16 * half + 4 * q + 1, v.y);
    w.set_word(16 * half + 4 * q + 2, v.z);
    w.set_word(16 * half + 4 * q + 3, v.w);
  }
}
template <int LANES>
__device__ __forceinline__ void load_raw_row_coherent(const LeanLds<LANES> &w, int half,
                                                      const uint8_t *src) {
  uint64_t *s = reinterpret_cast<uint64_t *>(const_cast<uint8_t *>(src));
#pragma unroll
  for (int q = 0; q < kTape / 8; q++) {
    const uint64_t v = __hip_atomic_load(s + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    w.set_word(16 * half + 2 * q, (uint32_t)v);
    w.set_word(16 * half + 2 * q + 1, (uint32_t)(v >> 32));
  }
}
// Mutation of common_language.h:172-180 on bytes [b0, b1) of a raw tape:
// byte b of slot k is replaced by r & 0xFF when ((r >> 8) & (2^30 - 1)) <
// P, r = SM((N*S(e) + k)*128 + b).
template <int LANES>
__device__ __forceinline__ void mutate_raw(const LeanLds<LANES> &w, uint64_t n, uint64_t seed_value,
                                           uint64_t k, uint32_t prob, int b0 = 0,
                                           int b1 = kTape2) {
  const uint64_t base = (n * seed_value + k) * (uint64_t)kTape2;
  for (int b = b0; b < b1; b++) {
    const uint64_t r = splitmix64(base + b);
    if (((uint32_t)(r >> 8) & 0x3fffffffu) < prob) w.set_byte(b >> 2, b & 3, (uint32_t)r & 0xff);
  }
}
// The raw tape as the final one, for an interaction that is not
// evaluated: byte 0 of position p's word is tape byte p.
template <int LANES>
__device__ __forceinline__ void raw_as_final(const LeanLds<LANES> &w) {
  uint32_t raw[kTape2 / 4];
#pragma unroll
  for (int j = 0; j < kTape2 / 4; j++) raw[j] = w.word(j);
#pragma unroll
  for (int p = 0; p < kTape2; p++) w.set_word(p, (raw[p >> 2] >> (8 * (p & 3))) & 0xff);
}
// Half h of a lane's final tape (byte 0 of its words) written to a soup row
// with 16-byte stores, or with 8-byte device-coherent stores.
template <int LANES>
__device__ __forceinline__ void store_lean_row(const LeanLds<LANES> &w, int half, uint8_t *dst) {
  uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    uint32_t v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int p = kTape * half + 16 * q + 4 * j;
      v[j] = (w.word(p) & 0xff) | ((w.word(p + 1) & 0xff) << 8) | ((w.word(p + 2) & 0xff) << 16) |
             (w.word(p + 3) << 24);
    }
    d4[q] = make_uint4(v[0], v[1], v[2], v[3]);
  }
}
template <int LANES>
__device__ __forceinline__ void store_lean_row_coherent(const LeanLds<LANES> &w, int half,
                                                        uint8_t *dst) {
  uint64_t *d = reinterpret_cast<uint64_t *>(dst);
#pragma unroll
  for (int q = 0; q < kTape / 8; q++) {
    uint64_t v = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) v |= (uint64_t)(w.word(kTape * half + 8 * q + j) & 0xff) << (8 * j);
    __hip_atomic_store(d + q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
#endif

}  // namespace cubff
