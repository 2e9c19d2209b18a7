// CPU check of the budgeted BFF interpreter (cubff_amd/csrc/bff_fast.h)
// against the oracle's Evaluate: the header's template compiled for the
// host, run on random tapes at random step limits and command budgets, and
// runs it stopped handed to the word-table interpreter (bff_lean.h).
// Test infrastructure: built and run by tests/test_cpu_interp.py.
//
//   bff_fast_check <tapes per language> <seed>
// prints one line per language and exits non-zero on the first mismatch.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../cubff_amd/csrc/bff_fast.h"
#include "../../cubff_amd/csrc/bff_lean.h"
#include "../../oracle/cubff_oracle.h"

using namespace cubff;

namespace {

struct HostBytes {
  uint8_t *p;
  uint32_t get(uint32_t i) const { return p[i]; }
  void set(uint32_t i, uint8_t v) const { p[i] = v; }
  uint32_t word(uint32_t w) const {
    return (uint32_t)p[4 * w] | ((uint32_t)p[4 * w + 1] << 8) | ((uint32_t)p[4 * w + 2] << 16) |
           ((uint32_t)p[4 * w + 3] << 24);
  }
  void set_word(uint32_t w, uint32_t v) const {
    for (int b = 0; b < 4; b++) p[4 * w + b] = (uint8_t)(v >> (8 * b));
  }
};

// The word-table interpreter's per-position words and code table, in host
// memory.
struct LeanWords {
  uint32_t *p;
  uint32_t word(uint32_t pos) const { return p[pos]; }
  void set_word(uint32_t pos, uint32_t w) const { p[pos] = w; }
  void set_low(uint32_t pos, uint32_t v) const { p[pos] = (p[pos] & 0xffff0000u) | (v & 0xffff); }
  void set_byte(uint32_t pos, int b, uint32_t v) const {
    p[pos] = (p[pos] & ~(0xffu << (8 * b))) | ((v & 0xff) << (8 * b));
  }
};
template <int MAP, bool HEADS>
struct LeanCodes {
  uint8_t t[256];
  LeanCodes() {
    for (int v = 0; v < 256; v++) t[v] = (uint8_t)BffLean<MAP, HEADS>::code_of(v);
  }
  uint32_t operator()(uint32_t v) const { return t[v & 255]; }
};

template <int MAP>
struct HostLut {
  uint8_t t[256];
  HostLut() {
    for (int v = 0; v < 256; v++) t[v] = (uint8_t)bff_byte_class<MAP>(v);
  }
  uint32_t operator()(uint32_t v) const { return t[v & 255]; }
};

uint64_t sm(uint64_t x) { return orc_splitmix64(x); }

// Tapes with many commands, loops and heads near the code: byte j of tape
// r is a command character with probability 5/8, else a random byte (for
// the permuted and 4-bit maps, small values or random low nibbles).
void make_tape(int map, uint64_t seed, uint8_t *t) {
  static const char kCmds[] = "[]+-.,<>{}";
  for (int j = 0; j < 128; j++) {
    const uint64_t r = sm(seed * 131 + j);
    uint8_t v;
    if (map == kMapPerm) {
      v = (r & 7) < 6 ? (uint8_t)((r >> 8) % 12) : (uint8_t)(r >> 16);
    } else if (map == kMap4bit) {
      v = (uint8_t)(r >> 8);
      if ((r & 7) < 5) v = (uint8_t)((v & 0xf0) | (6 + (r >> 20) % 10));
    } else {
      v = (r & 7) < 5 ? (uint8_t)kCmds[(r >> 8) % 10] : (uint8_t)(r >> 16);
      if ((r & 63) == 7) v = 0;
    }
    t[j] = v;
  }
}

template <int MAP, bool HEADS, bool ACCEL>
int check(const char *name, int orc_lang, int count, uint64_t seed) {
  HostLut<MAP> lut;
  uint8_t tape[129], ref[128], nxb[129];
  uint64_t cap_runs = 0, over = 0, ops_sum = 0, full_runs = 0, resumed = 0, handed = 0;
  for (int r = 0; r < count; r++) {
    make_tape(MAP, seed * 1000003 + r, tape);
    const uint64_t x = sm(seed ^ (0x9e37ull * (r + 1)));
    const uint32_t caps[4] = {8192, (uint32_t)(x % 300), 1 + (uint32_t)(x >> 20) % 3000, 0};
    const uint32_t stepcount = caps[(x >> 40) & 3];
    const uint32_t budget = ((x >> 44) & 3) == 0 ? 0xffffffffu : 1 + (uint32_t)((x >> 46) % 400);
    memcpy(ref, tape, 128);
    uint64_t ref_steps = 0;
    const uint64_t ref_ops = orc_evaluate(orc_lang, ref, stepcount, &ref_steps);
    uint32_t steps = 0, work = 0;
    // Complete runs (no budget) also go through the full-state cycle skip
    // (with loop acceleration), every other one.
    const bool full = ACCEL && budget == 0xffffffffu && (r & 1);
    uint8_t anchor[128];
    const uint32_t ops =
        full ? BffFast<MAP, HEADS, ACCEL>::evaluate_full(HostBytes{tape}, HostBytes{nxb},
                                                        HostBytes{anchor}, lut, stepcount, &steps,
                                                        &work)
             : BffFast<MAP, HEADS, ACCEL>::evaluate(HostBytes{tape}, HostBytes{nxb}, lut,
                                                   stepcount, &steps, &work, budget);
    full_runs += full;
    if (steps == kOverBudget) {
      over++;
      // Without acceleration the budget counts exactly the run's commands.
      if (!ACCEL && ref_ops <= budget) {
        printf("%s tape %d: over budget %u but the run has %llu commands\n", name, r, budget,
               (unsigned long long)ref_ops);
        return 1;
      }
      continue;
    }
    if (!ACCEL && budget != 0xffffffffu && ref_ops > budget) {
      printf("%s tape %d: %llu commands under budget %u not reported\n", name, r,
             (unsigned long long)ref_ops, budget);
      return 1;
    }
    if (ops != ref_ops || steps != ref_steps || work > ops || (!ACCEL && work != ops) ||
        (budget != 0xffffffffu && work > budget) || memcmp(tape, ref, 128) != 0) {
      printf("%s tape %d cap %u: ops %u/%llu steps %u/%llu tape %s\n", name, r, stepcount, ops,
             (unsigned long long)ref_ops, steps, (unsigned long long)ref_steps,
             memcmp(tape, ref, 128) ? "differs" : "same");
      return 1;
    }
    cap_runs += steps == stepcount;
    ops_sum += ops;
    // The same run resumed: start (or start_words, from the tape's words),
    // then advance() in slices of 1..64 commands until it ends (with the
    // cycle skip every other tape).
    {
      using F = BffFast<MAP, HEADS, ACCEL>;
      uint8_t t2[129], nx2[129], anchor2[128];
      make_tape(MAP, seed * 1000003 + r, t2);
      typename F::Run run;
      if (r & 2) {
        uint32_t words[32];
        for (int w = 0; w < 32; w++) words[w] = HostBytes{t2}.word(w);
        F::start_words(run, words, HostBytes{nx2}, lut, stepcount);
      } else {
        F::start(run, HostBytes{t2}, HostBytes{nx2}, lut, stepcount);
      }
      uint64_t calls = 0;
      while (!run.done) {
        const uint32_t slice = 1 + (uint32_t)(sm(x + 77 * calls) % 64);
        if (full)
          F::template advance<true>(run, HostBytes{t2}, HostBytes{nx2}, HostBytes{anchor2}, lut,
                                    stepcount, slice);
        else
          F::template advance<false>(run, HostBytes{t2}, HostBytes{nx2},
                                     typename F::NoAux{}, lut, stepcount, slice);
        if (++calls > 100000) {
          printf("%s tape %d: resumed run does not end\n", name, r);
          return 1;
        }
      }
      uint32_t steps2 = 0, work2 = 0;
      const uint32_t ops2 = F::finish(run, stepcount, &steps2, &work2);
      if (ops2 != ref_ops || steps2 != ref_steps || memcmp(t2, ref, 128) != 0) {
        printf("%s tape %d cap %u: resumed ops %u/%llu steps %u/%llu tape %s\n", name, r,
               stepcount, ops2, (unsigned long long)ref_ops, steps2,
               (unsigned long long)ref_steps, memcmp(t2, ref, 128) ? "differs" : "same");
        return 1;
      }
      resumed++;
    }
    // A run stopped after a command budget in slices of 8 (as the batched
    // refill kernel runs it) goes on to the end in the word-table
    // interpreter from its tape bytes and state (as the resolver does).
    if (ACCEL) {
      using F = BffFast<MAP, HEADS, ACCEL>;
      using G = BffLean<MAP, HEADS>;
      uint8_t t3[129], nx3[129];
      make_tape(MAP, seed * 1000003 + r, t3);
      typename F::Run run;
      F::start(run, HostBytes{t3}, HostBytes{nx3}, lut, stepcount);
      const uint32_t cut = 1 + (uint32_t)(sm(x + 5) % 300);
      while (!run.done && run.iters < cut)
        F::template advance<false>(run, HostBytes{t3}, HostBytes{nx3}, typename F::NoAux{}, lut,
                                   stepcount, 8);
      if (!run.done) {
        const typename G::Resume sv{(uint32_t)run.pc | ((uint32_t)run.h0 << 8) |
                                        ((uint32_t)run.h1 << 16),
                                    run.i, run.iters, run.extra};
        uint32_t wr[129];
        for (int w = 0; w < 32; w++) wr[w] = HostBytes{t3}.word(w);
        static const LeanCodes<MAP, HEADS> codes;
        typename G::Run gr;
        G::start_resume(gr, LeanWords{wr}, codes, stepcount, sv);
        G::template advance<false>(gr, LeanWords{wr}, codes, stepcount, 0xffffffffu);
        uint32_t steps3 = 0;
        const uint32_t ops3 = G::finish(gr, stepcount, &steps3, nullptr);
        uint8_t out3[128];
        for (int q = 0; q < 128; q++) out3[q] = (uint8_t)wr[q];
        if (ops3 != ref_ops || steps3 != ref_steps || memcmp(out3, ref, 128) != 0) {
          printf("%s tape %d cap %u: handed to the word table after %u commands: ops %u/%llu "
                 "steps %u/%llu tape %s\n", name, r, stepcount, cut, ops3,
                 (unsigned long long)ref_ops, steps3, (unsigned long long)ref_steps,
                 memcmp(out3, ref, 128) ? "differs" : "same");
          return 1;
        }
        handed++;
      }
    }
  }
  printf("%s: %d tapes ok (%llu at the step limit, %llu over budget, %llu with the cycle skip, "
         "%llu resumed in slices, %llu handed to the word table, %llu ops)\n", name, count,
         (unsigned long long)cap_runs, (unsigned long long)over, (unsigned long long)full_runs,
         (unsigned long long)resumed, (unsigned long long)handed, (unsigned long long)ops_sum);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const int count = argc > 1 ? atoi(argv[1]) : 20000;
  const uint64_t seed = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;
  int bad = 0;
  for (int accel = 0; accel < 2; accel++) {
    printf("loop acceleration %s\n", accel ? "on" : "off");
    if (accel) {
      bad |= check<kMapAscii, true, true>("bff", ORC_BFF, count, seed);
      bad |= check<kMapAscii, false, true>("bff_noheads", ORC_BFF_NOHEADS, count, seed + 1);
      bad |= check<kMap4bit, false, true>("bff_noheads_4bit", ORC_BFF_NOHEADS_4BIT, count, seed + 2);
      bad |= check<kMapPerm, true, true>("bff_perm", ORC_BFF_PERM, count, seed + 3);
    } else {
      bad |= check<kMapAscii, true, false>("bff", ORC_BFF, count, seed);
      bad |= check<kMapAscii, false, false>("bff_noheads", ORC_BFF_NOHEADS, count, seed + 1);
      bad |= check<kMap4bit, false, false>("bff_noheads_4bit", ORC_BFF_NOHEADS_4BIT, count, seed + 2);
      bad |= check<kMapPerm, true, false>("bff_perm", ORC_BFF_PERM, count, seed + 3);
    }
  }
  return bad;
}
