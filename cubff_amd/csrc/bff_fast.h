// BFF interpreter of the standard jump logic (bff, bff_noheads,
// bff_noheads_4bit, bff_perm) built for the shortest chain of instructions
// per executed command.  It serves the budgeted runs of the deferring epoch
// kernel, where an epoch lasts as long as its longest budgeted interaction,
// and so costs what one lane's command costs.
//
// Per lane it keeps, beside the 128-byte tape, a 128-byte table `nx`: at
// every command position, the next command position after it (128 if
// none).  pc is always on a command, so a command reads its byte and
// nx[pc] in the same LDS round trip as the bytes under the two heads, and
// the run of comment bytes after it (no-op steps, bff.inc.h:346-347) is
// skipped without any mask search.  Every command stores one byte (the byte
// it writes, or for commands that write nothing the byte under head 0 back
// in place), so no select picks a store address and no branch skips a
// store.  A 256-byte table of byte classes (comment, head move, write, '[',
// ']') shared by the workgroup tells whether a write changed a byte's
// class; only then are the class masks, two nx entries and the jump caches
// brought up to date (rare).  Bracket scans (bff.inc.h:317-344) walk the
// masks and are cached per jump source, with the classes of the loop body
// they enclose, until a class changes.
//
// Loop acceleration (exact): a taken backward jump from the same `]` as the
// last one, with no class changed since, closes one iteration of the body
// between the brackets.  If that body holds no bracket it runs straight
// through, and its next iterations repeat the last one exactly when (a) it
// holds no write -- the tape is constant, the heads advance by fixed
// deltas, and the loop runs until the byte under head 0 at the `]` is null,
// found by walking head 0's orbit (at most 128 reads) -- or (b) it holds no
// head move and left the two bytes under the heads as they were: the whole
// state repeats.  Whole iterations are then counted (steps and ops) without
// being run, up to the step cap and the exiting iteration, which runs
// normally.
//
// Steps add the skipped comment bytes to the executed commands, capped at
// stepcount; ops (the value Evaluate returns, bff.inc.h:356-389) are the
// executed commands plus those of skipped iterations.
#pragma once

#include "languages.h"

namespace cubff {

// Class of a byte under a BFF map: 0 comment (or null), 1 head move
// (< > { }), 2 write (+ - . ,), 3 '[', 4 ']'.
enum : uint32_t { kClsNone = 0, kClsMove, kClsWrite, kClsOpen, kClsClose };
template <int MAP>
__host__ __device__ __forceinline__ uint32_t bff_byte_class(uint32_t v) {
  using C = BffClass<MAP>;
  if (!C::cmd(v)) return kClsNone;
  const uint32_t k = C::kind(v);
  return k == kLoopStart ? kClsOpen
                         : (k == kLoopEnd ? kClsClose : ((k - kPlus) < 4u ? kClsWrite : kClsMove));
}

// Positions strictly between a and b (-1 <= a < b <= 128) as a mask.
__host__ __device__ __forceinline__ Mask128 mask_between(int a, int b) {
  const auto upto = [](int n) {  // positions [0, n), 0 <= n <= 128
    Mask128 m;
    m.lo = n >= 64 ? ~0ULL : ((1ULL << (n & 63)) - 1);
    m.hi = n >= 128 ? ~0ULL : (n <= 64 ? 0ULL : ((1ULL << (n & 63)) - 1));
    return m;
  };
  const Mask128 hi = upto(b), lo = upto(a + 1);
  return Mask128{hi.lo & ~lo.lo, hi.hi & ~lo.hi};
}
__host__ __device__ __forceinline__ bool mask_any(const Mask128 &m, const Mask128 &r) {
  return ((m.lo & r.lo) | (m.hi & r.hi)) != 0;
}

// Keeps the loads issued before it together (device pass only).
__host__ __device__ __forceinline__ void keep_loads_together() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_sched_barrier(0);
#endif
}

__host__ __device__ __forceinline__ uint32_t fast_umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ int fast_imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ int fast_imax(int a, int b) { return a > b ? a : b; }

// Class masks of a tape: every command, '[', ']', writes, head moves.
struct BffMasks {
  Mask128 cmd, open, close, write, move;
};

template <int MAP, bool HEADS, bool ACCEL = true>
struct BffFast {
  using C = BffClass<MAP>;
  // Per-kind 2-bit fields indexed by BffKind: head-0 delta + 1, head-1
  // delta + 1, byte delta + 1 (the same tables as BffLang).
  static constexpr uint32_t kTH0 = 0x58555;
  static constexpr uint32_t kTH1 = 0x85555;
  static constexpr uint32_t kTDelta = 0x55525;

  // Class masks and nx from the tape bytes.
  template <class T, class N, class L>
  __host__ __device__ __forceinline__ static void build(const T &t, const N &nx, const L &lut,
                                                        BffMasks &m) {
    m = BffMasks{};
    for (int w = 0; w < kTape2 / 8; w++) {
      uint64_t bc = 0, bo = 0, bx = 0, bw = 0, bm = 0;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const uint32_t c = lut(t.get(w * 8 + j));
        bc |= (uint64_t)(c != kClsNone) << j;
        bo |= (uint64_t)(c == kClsOpen) << j;
        bx |= (uint64_t)(c == kClsClose) << j;
        bw |= (uint64_t)(c == kClsWrite) << j;
        bm |= (uint64_t)(c == kClsMove) << j;
      }
      const int sh = (w * 8) & 63;
      if (w < 8) {
        m.cmd.lo |= bc << sh; m.open.lo |= bo << sh; m.close.lo |= bx << sh;
        m.write.lo |= bw << sh; m.move.lo |= bm << sh;
      } else {
        m.cmd.hi |= bc << sh; m.open.hi |= bo << sh; m.close.hi |= bx << sh;
        m.write.hi |= bw << sh; m.move.hi |= bm << sh;
      }
    }
    fill_nx(nx, m);
  }

  // nx[p] = smallest command position > p, from the top down (read at
  // command positions only).
  template <class N>
  __host__ __device__ __forceinline__ static void fill_nx(const N &nx, const BffMasks &m) {
    uint32_t next = kTape2;
#pragma unroll 4
    for (int p = kTape2 - 1; p >= 0; p--) {
      nx.set(p, (uint8_t)next);
      const uint64_t mm = p < 64 ? m.cmd.lo : m.cmd.hi;
      next = ((mm >> (p & 63)) & 1) ? (uint32_t)p : next;
    }
  }

  // Byte w changed class from cold to cnew: masks and nx follow (nx is
  // kept at command positions: the command before w, and w itself).
  // Returns the command before w (-1 if none) and in *next_of_prev its new
  // next command.
  template <class N>
  __host__ __device__ static int class_change(const N &nx, BffMasks &m, int w, uint32_t cold,
                                              uint32_t cnew, int *next_of_prev) {
    const bool was = cold != kClsNone, is = cnew != kClsNone;
    m.cmd.flip(w, was != is);
    m.open.flip(w, (cold == kClsOpen) != (cnew == kClsOpen));
    m.close.flip(w, (cold == kClsClose) != (cnew == kClsClose));
    m.write.flip(w, (cold == kClsWrite) != (cnew == kClsWrite));
    m.move.flip(w, (cold == kClsMove) != (cnew == kClsMove));
    const int prev = m.cmd.prev_upto(w - 1);
    const int after = m.cmd.next_from(w + 1);
    *next_of_prev = is ? w : after;
    if (was != is) {
      if (prev >= 0) nx.set(prev, (uint8_t)*next_of_prev);
      if (is) nx.set(w, (uint8_t)after);
    }
    return prev;
  }

  // Body classes of the loop closed by the `]` at src with its `[` at q.
  __host__ __device__ __forceinline__ static void body_classes(const BffMasks &m, int q, int src,
                                                               bool *straight, bool *write,
                                                               bool *move) {
    const Mask128 body = mask_between(q, src);
    *straight = !mask_any(m.open, body) && !mask_any(m.close, body);
    *write = mask_any(m.write, body);
    *move = mask_any(m.move, body);
  }

  // Full-state cycle skip of complete runs (exact): from step kCycleStart
  // on, the state right after a taken backward jump is compared with an
  // anchor whose tape is kept in a second per-lane region: pc, heads, a
  // weighted sum of the tape changes since the anchor (zero when the tape
  // is back to the anchor's) and, only when those match, every tape word.
  // A match means the run repeats with that period: whole periods are
  // counted without being run.  The first anchor stays for kCycleFirstWin
  // steps (the long runs are mostly cycles that start within ~150 steps
  // with periods of 1.5K-4K steps, a head sweeping the tape), later anchors
  // follow Brent's scheme with windows kCycleWin, 2 kCycleWin, ...
  static constexpr uint32_t kCycleStart = 128;
  static constexpr uint32_t kCycleFirstWin = 4096;
  static constexpr uint32_t kCycleWin = 1024;

  // Tape accessor without a second region (runs without the cycle skip).
  struct NoAux {
    __host__ __device__ uint32_t word(uint32_t) const { return 0; }
    __host__ __device__ void set_word(uint32_t, uint32_t) const {}
  };

  // Evaluate (bff.inc.h:356-389) with EvaluateOne (:289-349), for a
  // budgeted run.  With a budget (>= 1) below 2^32 - 1, a run whose
  // executed commands plus orbit-walk reads would exceed `budget` stops and
  // reports *steps = kOverBudget (its tape is then partly run).  lut(v) is
  // bff_byte_class<MAP>(v).  *work receives the commands executed.
  template <class T, class N, class L>
  __host__ __device__ static uint32_t evaluate(const T &t, const N &nx, const L &lut,
                                               uint32_t stepcount, uint32_t *steps,
                                               uint32_t *work, uint32_t budget) {
    return run<false>(t, nx, NoAux{}, lut, stepcount, steps, work, budget);
  }
  // A complete run with the full-state cycle skip; aux is a second
  // 128-byte per-lane region (the anchor tape).
  template <class T, class N, class A, class L>
  __host__ __device__ static uint32_t evaluate_full(const T &t, const N &nx, const A &aux,
                                                    const L &lut, uint32_t stepcount,
                                                    uint32_t *steps, uint32_t *work) {
    return run<true>(t, nx, aux, lut, stepcount, steps, work, 0xffffffffu);
  }

  // State of one run between calls of advance(): masks, pc and heads,
  // steps, commands executed, ops of skipped iterations, whether it has
  // ended, and the jump, loop and cycle anchors described in advance().
  struct Run {
    BffMasks m;
    int h0, h1, pc;
    uint32_t i, iters, extra;
    bool done;
    int bsrc, bq, bnext, bgap, fsrc, fq, fnext, fgap;
    bool bstraight, bwrite, bmove;
    int aq, ah0, ah1;
    uint32_t ai, aiters, aw0, aw1;
    int cq, ch0, ch1;
    uint32_t ci, cops, cwin, cdelta;
    bool cfirst;
  };

  // Class masks and nx from the tape's 32 words (byte 4w + j is byte j of
  // words[w]), held in registers: no tape reads.
  template <class N, class L>
  __host__ __device__ __forceinline__ static void build_words(const uint32_t *words, const N &nx,
                                                              const L &lut, BffMasks &m) {
    m = BffMasks{};
#pragma unroll
    for (int w = 0; w < kTape2 / 4; w++) {
      uint64_t bc = 0, bo = 0, bx = 0, bw = 0, bm = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t c = lut((words[w] >> (8 * j)) & 0xff);
        bc |= (uint64_t)(c != kClsNone) << j;
        bo |= (uint64_t)(c == kClsOpen) << j;
        bx |= (uint64_t)(c == kClsClose) << j;
        bw |= (uint64_t)(c == kClsWrite) << j;
        bm |= (uint64_t)(c == kClsMove) << j;
      }
      const int sh = (w * 4) & 63;
      if (w < 16) {
        m.cmd.lo |= bc << sh; m.open.lo |= bo << sh; m.close.lo |= bx << sh;
        m.write.lo |= bw << sh; m.move.lo |= bm << sh;
      } else {
        m.cmd.hi |= bc << sh; m.open.hi |= bo << sh; m.close.hi |= bx << sh;
        m.write.hi |= bw << sh; m.move.hi |= bm << sh;
      }
    }
    fill_nx(nx, m);
  }

  // Start of a run on tape t: masks and nx, initial state, first command.
  template <class T, class N, class L>
  __host__ __device__ __forceinline__ static void start(Run &r, const T &t, const N &nx,
                                                        const L &lut, uint32_t stepcount) {
    build(t, nx, lut, r.m);
    if constexpr (HEADS)
      init(r, t.get(0), t.get(1), stepcount);
    else
      init(r, 0, 0, stepcount);
  }
  // The same for a tape whose 32 words are also in registers.
  template <class N, class L>
  __host__ __device__ __forceinline__ static void start_words(Run &r, const uint32_t *words,
                                                              const N &nx, const L &lut,
                                                              uint32_t stepcount) {
    build_words(words, nx, lut, r.m);
    init(r, words[0] & 0xff, (words[0] >> 8) & 0xff, stepcount);
  }

  // Initial state from the masks (b0, b1: the first two tape bytes).
  __host__ __device__ __forceinline__ static void init(Run &r, uint32_t b0, uint32_t b1,
                                                       uint32_t stepcount) {
    int pc;
    if constexpr (HEADS) {  // InitialState, bff.inc.h:269-275
      r.h0 = b0 & (kTape2 - 1);
      r.h1 = b1 & (kTape2 - 1);
      pc = 2;
    } else {  // bff.inc.h:277-282; 128 masks to 0
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
    // Jump caches: source pc and matching bracket of the last backward
    // jump (bq -1: unmatched) and forward jump (fq 128: unmatched), the
    // next command after the target (nx at the matching bracket, kept up
    // to date through class changes) and the comment bytes before it.
    // For the backward jump, whether the body between its brackets is
    // straight (no bracket), holds a write, holds a head move.
    r.bsrc = -1; r.bq = -1; r.bnext = 0; r.bgap = 0;
    r.fsrc = -1; r.fq = kTape2; r.fnext = 0; r.fgap = 0;
    r.bstraight = r.bwrite = r.bmove = false;
    // Loop anchor: the last taken backward jump's source (-1 none), the
    // heads, steps and commands after it, the bytes under the heads.
    r.aq = -1; r.ah0 = 0; r.ah1 = 0;
    r.ai = 0; r.aiters = 0; r.aw0 = 0; r.aw1 = 0;
    // Cycle anchor: jump source (-1 none yet, -2 done), heads, steps and
    // ops after it, window, weighted sum of the tape changes since.
    r.cq = -1; r.ch0 = 0; r.ch1 = 0;
    r.ci = 0; r.cops = 0; r.cwin = kCycleFirstWin; r.cdelta = 0;
    r.cfirst = true;
  }

  // Runs commands until the run ends (r.done) or `budget` (>= 1) has been
  // spent in this call: commands plus the tape reads of loop acceleration's
  // orbit walks.  A later call goes on from the next command.
  template <bool CYCLE, class T, class N, class A, class L>
  __host__ __device__ __forceinline__ static void advance(Run &r, const T &t, const N &nx,
                                                          const A &aux, const L &lut,
                                                          uint32_t stepcount, uint32_t budget) {
    if (r.done) return;
    BffMasks m = r.m;
    int h0 = r.h0, h1 = r.h1, pc = r.pc;
    uint32_t i = r.i, iters = r.iters, extra = r.extra, spent = 0;
    bool done = false;
    int bsrc = r.bsrc, bq = r.bq, bnext = r.bnext, bgap = r.bgap;
    int fsrc = r.fsrc, fq = r.fq, fnext = r.fnext, fgap = r.fgap;
    bool bstraight = r.bstraight, bwrite = r.bwrite, bmove = r.bmove;
    int aq = r.aq, ah0 = r.ah0, ah1 = r.ah1;
    uint32_t ai = r.ai, aiters = r.aiters, aw0 = r.aw0, aw1 = r.aw1;
    int cq = r.cq, ch0 = r.ch0, ch1 = r.ch1;
    uint32_t ci = r.ci, cops = r.cops, cwin = r.cwin, cdelta = r.cdelta;
    bool cfirst = r.cfirst;
    for (;;) {
      // pc < 128 here; the mask lets the address be one 24-bit multiply-add.
      const int pcm = pc & (kTape2 - 1);
      const uint32_t c = t.get(pcm);
      const uint32_t fall = nx.get(pcm);
      const uint32_t v0 = t.get(h0);
      const uint32_t v1 = t.get(h1);
      keep_loads_together();
      const uint32_t k = C::kind(c);
      const uint32_t k2 = 2 * k;
      const bool c01 = k == kCopy01, c10 = k == kCopy10;
      const uint32_t wold = c01 ? v1 : v0;
      const uint32_t vnew = ((c10 ? v1 : v0) + ((kTDelta >> k2) & 3) - 1) & 0xff;
      const int waddr = c01 ? h1 : h0;
      t.set(waddr, (uint8_t)vnew);
      if constexpr (CYCLE) cdelta += (vnew - wold) * (2654435761u * (2u * (uint32_t)waddr + 1u));
      const uint32_t cold = lut(wold), cnew = lut(vnew);
      h0 = (h0 + (int)((kTH0 >> k2) & 3) - 1) & (kTape2 - 1);
      h1 = (h1 + (int)((kTH1 >> k2) & 3) - 1) & (kTape2 - 1);
      const bool z = C::null(v0);
      const bool tb = k == kLoopEnd && !z;
      const bool tf = k == kLoopStart && z;
      int nextc, gap;
      if ((tb && pc != bsrc) || (tf && pc != fsrc)) {
        // Matching bracket: backward to the `[` (unmatched: the run
        // stops), forward to the `]` (unmatched: past the tape).
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
          if (depth == 0) break;
        }
        if (tb) {
          bsrc = pc;
          bq = q;
          bnext = nextc = q < 0 ? -1 : m.cmd.next_from(q + 1);
          bgap = gap = q < 0 ? 0 : nextc - (q + 1);
          if constexpr (ACCEL) {
            body_classes(m, q, pc, &bstraight, &bwrite, &bmove);
            bstraight = bstraight && q >= 0;
          }
        } else {
          fsrc = pc;
          fq = fast_imin(q, kTape2);
          const int after = fast_imin(q + 1, kTape2);
          fnext = nextc = m.cmd.next_from(after);
          fgap = gap = nextc - after;
        }
      } else {
        nextc = tb ? bnext : (tf ? fnext : (int)fall);
        gap = tb ? bgap : (tf ? fgap : (int)fall - pc - 1);
      }
      if constexpr (ACCEL) {
        // Any other bracket command between two visits of the anchor `]`
        // means the path between them was not one body iteration.
        if (k <= kLoopEnd && pc != aq) aq = -1;
        bool skipped = false;
        if (tb && bstraight) {
          const uint32_t ib = i + 1;  // steps after this `]`
          if (pc == aq &&
              (!bwrite || (!bmove && h0 == ah0 && h1 == ah1 && v0 == aw0 && v1 == aw1))) {
            // One more iteration of the loop closed with the state of the
            // last one advanced by fixed deltas: skip whole iterations.
            const uint32_t len = ib - ai;
            uint32_t reps = (stepcount - ib) / len;
            const int d0 = (h0 - ah0) & (kTape2 - 1);
            const int d1 = (h1 - ah1) & (kTape2 - 1);
            if (d0 != 0) {
              // First later iteration whose `]` sees a null byte (the
              // tape is constant here: no write in the body), looked for
              // 8 iterations at a time: the 8 reads of a group do not wait
              // for each other.  Past 128 iterations the positions repeat.
              const uint32_t lim = fast_umin(reps, (uint32_t)kTape2);
              uint32_t j = 1;
              bool found = false;
              for (; j <= lim; j += 8) {
                uint32_t hit = 0;
#pragma unroll
                for (int u = 0; u < 8; u++)
                  hit |= (uint32_t)C::null(t.get((h0 + (j + u) * d0) & (kTape2 - 1))) << u;
                if (lim - j < 7) hit &= (1u << (lim - j + 1)) - 1;  // iterations past lim
                if (hit) {
                  j += __builtin_ctz(hit);
                  found = true;
                  break;
                }
              }
              if (found) reps = j - 1;
              spent += found ? j : lim + 1;
            }
            i += reps * len;
            extra += reps * (iters + 1 - aiters);
            h0 = (h0 + (int)(reps * d0)) & (kTape2 - 1);
            h1 = (h1 + (int)(reps * d1)) & (kTape2 - 1);
            aq = -1;
            skipped = true;
          } else {
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
          if (tb && !skipped && nextc >= 0 && ib >= kCycleStart && cq != -2) {
            bool same = pc == cq && h0 == ch0 && h1 == ch1 && cdelta == 0;
            for (int w = 0; same && w < kTape2 / 4; w++) same = t.word(w) == aux.word(w);
            if (same) {
              const uint32_t len = ib - ci;
              const uint32_t reps = (stepcount - ib) / len;
              i += reps * len;
              extra += reps * (iters + 1 + extra - cops);
              cq = -2;  // fewer than one period of steps is left
              aq = -1;  // the loop anchor's counts predate the skip
            } else if (cq < 0 || ib - ci >= cwin) {
              if (cq >= 0) {
                cwin = cfirst ? kCycleWin : (cwin < (1u << 29) ? cwin * 2 : cwin);
                cfirst = false;
              }
              for (int w = 0; w < kTape2 / 4; w++) aux.set_word(w, t.word(w));
              cq = pc;
              ch0 = h0;
              ch1 = h1;
              ci = ib;
              cops = iters + 1 + extra;
              cdelta = 0;
            }
          }
        }
      }
      if (cold != cnew) {
        // Only a write changes a class, and a write is no jump.  A change
        // that involves a bracket drops the jump caches; otherwise they
        // stay (their next commands come from nx), and a change inside
        // the cached loop's body updates its classes and drops the
        // anchor.
        int np;
        const int prev = class_change(nx, m, waddr, cold, cnew, &np);
        if (cold >= kClsOpen || cnew >= kClsOpen) {
          // A bracket appeared or went: a cached match stands unless the
          // change lies between the jump and its match (or past the jump
          // where it found none).
          if (waddr >= bq && waddr <= bsrc) bsrc = -1;
          if (waddr >= fsrc && waddr <= fq) fsrc = -1;
        }
        if constexpr (ACCEL) {
          if (bsrc >= 0 && waddr > bq && waddr < bsrc) {
            body_classes(m, bq, bsrc, &bstraight, &bwrite, &bmove);
            aq = -1;
          }
          if (bsrc < 0) aq = -1;
        }
        if (prev == pc) {
          nextc = np;
          gap = np - pc - 1;
        }
        // The next commands after the cached matches are nx there.
        if (bq >= 0 && prev == bq) {
          bnext = np;
          bgap = np - bq - 1;
        }
        if (prev == fq) {
          fnext = np;
          fgap = np - fq - 1;
        }
      }
      i += 1 + (uint32_t)gap;
      iters++;
      spent++;
      if ((uint32_t)nextc >= (uint32_t)kTape2 || i >= stepcount) {
        done = true;
        break;
      }
      pc = nextc;
      if (spent >= budget) break;
    }
    r.m = m;
    r.h0 = h0; r.h1 = h1; r.pc = pc;
    r.i = i; r.iters = iters; r.extra = extra; r.done = done;
    r.bsrc = bsrc; r.bq = bq; r.bnext = bnext; r.bgap = bgap;
    r.fsrc = fsrc; r.fq = fq; r.fnext = fnext; r.fgap = fgap;
    r.bstraight = bstraight; r.bwrite = bwrite; r.bmove = bmove;
    r.aq = aq; r.ah0 = ah0; r.ah1 = ah1;
    r.ai = ai; r.aiters = aiters; r.aw0 = aw0; r.aw1 = aw1;
    r.cq = cq; r.ch0 = ch0; r.ch1 = ch1;
    r.ci = ci; r.cops = cops; r.cwin = cwin; r.cdelta = cdelta;
    r.cfirst = cfirst;
  }

  // Results of an ended run: ops (the reference's count of executed
  // non-no-op steps), *steps and *work (commands executed).
  __host__ __device__ __forceinline__ static uint32_t finish(const Run &r, uint32_t stepcount,
                                                             uint32_t *steps, uint32_t *work) {
    *steps = fast_umin(r.i, stepcount);
    if (work) *work = r.iters;
    return r.iters + r.extra;
  }

  template <bool CYCLE, class T, class N, class A, class L>
  __host__ __device__ static uint32_t run(const T &t, const N &nx, const A &aux, const L &lut,
                                          uint32_t stepcount, uint32_t *steps, uint32_t *work,
                                          uint32_t budget) {
    Run r;
    start(r, t, nx, lut, stepcount);
    advance<CYCLE>(r, t, nx, aux, lut, stepcount, budget);
    if (!r.done) {
      *steps = kOverBudget;
      return 0;
    }
    return finish(r, stepcount, steps, work);
  }
};

// Device accessors: the workgroup's byte-class table in LDS.
struct LdsByteLut {
  const lds_u8 *p;
  __device__ __forceinline__ uint32_t operator()(uint32_t v) const { return p[v]; }
};

// Fills a workgroup's byte-class table (every thread of the block calls it
// before a __syncthreads()).
template <int MAP>
__device__ __forceinline__ void fill_byte_lut(uint8_t *lut) {
  for (uint32_t v = threadIdx.x; v < 256; v += blockDim.x) lut[v] = (uint8_t)bff_byte_class<MAP>(v);
}

}  // namespace cubff

namespace cubff {

// The languages that BffFast interprets (the standard jump logic), with
// the LDS it needs beside the tape: nx (a LinTape-shaped region) and the
// workgroup's byte-class table.
template <int LANG> struct FastOf { static constexpr bool kHas = false; };
template <> struct FastOf<kBff> { static constexpr bool kHas = true; using T = BffFast<kMapAscii, true>; static constexpr int kMap = kMapAscii; };
template <> struct FastOf<kBffNoheads> { static constexpr bool kHas = true; using T = BffFast<kMapAscii, false>; static constexpr int kMap = kMapAscii; };
template <> struct FastOf<kBffNoheads4bit> { static constexpr bool kHas = true; using T = BffFast<kMap4bit, false>; static constexpr int kMap = kMap4bit; };
template <> struct FastOf<kBffPerm> { static constexpr bool kHas = true; using T = BffFast<kMapPerm, true>; static constexpr int kMap = kMapPerm; };

// Second per-lane LDS region of a kernel over LANES lanes: the language's
// aux region, or nx where BffFast runs (whichever is larger).
template <int LANG, int LANES>
constexpr int second_region_bytes() {
  using L = typename LangOf<LANG>::T;
  constexpr int aux = lang_needs_aux(LANG) ? L::template Aux<LANES>::kRegionBytes : 0;
  constexpr int nx = FastOf<LANG>::kHas ? LinTape<LANES>::kRegionBytes : 0;
  return aux > nx ? aux : nx;
}

// Fills the byte-class table of a workgroup running BffFast for LANG (a
// no-op for other languages); every thread calls it before a barrier.
template <int LANG>
__device__ __forceinline__ void fill_fast_lut(uint8_t *lut) {
  if constexpr (FastOf<LANG>::kHas) fill_byte_lut<FastOf<LANG>::kMap>(lut);
}

}  // namespace cubff
