// CPU check of the wave-wide BFF interpreter (cubff_amd/csrc/bff_wave.h)
// against the oracle's Evaluate: the header's template compiled for the
// host (the wavefront's registers as an array of 64 words), on random
// tapes at random step limits, from the start and from the state of a
// word-table run (bff_lean.h) stopped by a command budget, as the resolver
// takes it over.  Test infrastructure: built and run by
// tests/test_cpu_interp.py.
//
//   bff_wave_check <tapes per language> <seed>
// prints one line per language and exits non-zero on the first mismatch.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../cubff_amd/csrc/bff_lean.h"
#include "../../cubff_amd/csrc/bff_wave.h"
#include "../../oracle/cubff_oracle.h"

using namespace cubff;

namespace {

// The wavefront's 64 registers of one word each.
struct HostRegs {
  uint32_t w[64];
  uint32_t word(uint32_t j) const { return w[j]; }
  void set_word(uint32_t j, uint32_t x) { w[j] = x; }
  void fill(uint32_t x) {
    for (int j = 0; j < 64; j++) w[j] = x;
  }
};

// The word table of bff_lean.h in host memory, and its code table.
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

uint64_t sm(uint64_t x) { return orc_splitmix64(x); }

// Tapes with many commands, loops and heads near the code.
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

void to_regs(const uint8_t *t, HostRegs &regs) {
  regs.fill(0);
  for (int w = 0; w < 32; w++)
    regs.w[w] = (uint32_t)t[4 * w] | ((uint32_t)t[4 * w + 1] << 8) |
                ((uint32_t)t[4 * w + 2] << 16) | ((uint32_t)t[4 * w + 3] << 24);
}

void from_regs(const HostRegs &regs, uint8_t *t) {
  for (int p = 0; p < 128; p++) t[p] = (uint8_t)(regs.w[p >> 2] >> (8 * (p & 3)));
}

template <int MAP, bool HEADS>
int check_tape(const char *name, int orc_lang, const uint8_t *tape, uint32_t stepcount,
               uint32_t budget, int r, uint64_t *handed, uint64_t *cap_runs) {
  using V = BffWave<MAP, HEADS>;
  using F = BffLean<MAP, HEADS>;
  static const LeanCodes<MAP, HEADS> lut;
  uint8_t ref[128];
  memcpy(ref, tape, 128);
  uint64_t ref_steps = 0;
  const uint64_t ref_ops = orc_evaluate(orc_lang, ref, stepcount, &ref_steps);
  // From the start.
  HostRegs t, match;
  to_regs(tape, t);
  match.fill(0xffffffffu);
  typename V::State s = V::start(t);
  V::run(t, match, stepcount, s);
  uint8_t out[128];
  from_regs(t, out);
  const uint32_t steps = s.i < stepcount ? s.i : stepcount;
  if (s.ops != ref_ops || steps != ref_steps || memcmp(out, ref, 128) != 0) {
    printf("%s tape %d cap %u: ops %u/%llu steps %u/%llu tape %s\n", name, r, stepcount, s.ops,
           (unsigned long long)ref_ops, steps, (unsigned long long)ref_steps,
           memcmp(out, ref, 128) ? "differs" : "same");
    return 1;
  }
  for (int j = 32; j < 64; j++)
    if (t.w[j] != 0) {
      printf("%s tape %d: register of lane %d written\n", name, r, j);
      return 1;
    }
  *cap_runs += steps == stepcount;
  // From a word-table run stopped by the budget: its tape bytes and saved
  // state, as the deferring epoch kernel leaves them.
  uint32_t wd[129];
  HostRegs raw;
  to_regs(tape, raw);
  for (int j = 0; j < 32; j++) wd[j] = raw.w[j];
  typename F::Run run;
  F::start_raw(run, LeanWords{wd}, lut, stepcount);
  F::template advance<false>(run, LeanWords{wd}, lut, stepcount, budget);
  if (run.done) return 0;
  const typename F::Resume sv = F::saved(run);
  uint8_t mid[128];
  for (int p = 0; p < 128; p++) mid[p] = (uint8_t)wd[p];
  HostRegs t2, m2;
  to_regs(mid, t2);
  m2.fill(0xffffffffu);
  typename V::State s2;
  s2.pc = (int)(sv.pos & 0xff);
  s2.h0 = (int)((sv.pos >> 8) & 0xff);
  s2.h1 = (int)((sv.pos >> 16) & 0xff);
  s2.i = sv.i;
  s2.ops = sv.iters + sv.extra;
  V::run(t2, m2, stepcount, s2);
  from_regs(t2, out);
  const uint32_t steps2 = s2.i < stepcount ? s2.i : stepcount;
  if (s2.ops != ref_ops || steps2 != ref_steps || memcmp(out, ref, 128) != 0) {
    printf("%s tape %d cap %u budget %u: handed over, ops %u/%llu steps %u/%llu tape %s\n", name,
           r, stepcount, budget, s2.ops, (unsigned long long)ref_ops, steps2,
           (unsigned long long)ref_steps, memcmp(out, ref, 128) ? "differs" : "same");
    return 1;
  }
  (*handed)++;
  return 0;
}

template <int MAP, bool HEADS>
int check_lang(const char *name, int orc_lang, int count, uint64_t seed) {
  uint64_t handed = 0, cap_runs = 0;
  for (int r = 0; r < count; r++) {
    uint8_t tape[128];
    make_tape(MAP, seed * 1000003 + r, tape);
    const uint64_t x = sm(seed ^ (0x51ed27ull * r));
    // Step limits from tiny to the reference's 8192 (most runs), budgets
    // from 1 command to a few hundred.
    const uint32_t stepcount = (x & 3) ? 8192 : 1 + (uint32_t)((x >> 8) % 3000);
    const uint32_t budget = 1 + (uint32_t)((x >> 24) % 400);
    if (check_tape<MAP, HEADS>(name, orc_lang, tape, stepcount, budget, r, &handed, &cap_runs))
      return 1;
  }
  printf("%s: %d tapes ok (%llu at the step limit, %llu handed over)\n", name, count,
         (unsigned long long)cap_runs, (unsigned long long)handed);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const int count = argc > 1 ? atoi(argv[1]) : 20000;
  const uint64_t seed = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;
  int bad = 0;
  bad |= check_lang<kMapAscii, true>("bff", orc_lang_id("bff"), count, seed);
  bad |= check_lang<kMapAscii, false>("bff_noheads", orc_lang_id("bff_noheads"), count, seed);
  bad |= check_lang<kMap4bit, false>("bff_noheads_4bit", orc_lang_id("bff_noheads_4bit"), count,
                                     seed);
  bad |= check_lang<kMapPerm, true>("bff_perm", orc_lang_id("bff_perm"), count, seed);
  return bad;
}
