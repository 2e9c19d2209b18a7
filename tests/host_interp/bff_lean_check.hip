// CPU check of the word-table BFF interpreter (cubff_amd/csrc/bff_lean.h)
// against the oracle's Evaluate: the header's template compiled for the
// host, run on random tapes at random step limits and command budgets,
// whole and resumed in slices.  Test infrastructure: built and run by
// tests/test_cpu_interp.py.
//
//   bff_lean_check <tapes per language> <seed> [tape file]
// prints one line per language and exits non-zero on the first mismatch.
// With a tape file (128-byte bff_noheads tapes back to back) its tapes are
// checked too, unbudgeted and at a budget of 768.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../cubff_amd/csrc/bff_lean.h"
#include "../../oracle/cubff_oracle.h"

using namespace cubff;

namespace {

// The per-position words of one lane, in host memory.
struct HostWords {
  uint32_t *p;
  uint32_t word(uint32_t pos) const { return p[pos]; }
  void set_word(uint32_t pos, uint32_t w) const { p[pos] = w; }
  void set_low(uint32_t pos, uint32_t v) const { p[pos] = (p[pos] & 0xffff0000u) | (v & 0xffff); }
  void set_byte(uint32_t pos, int b, uint32_t v) const {
    p[pos] = (p[pos] & ~(0xffu << (8 * b))) | ((v & 0xff) << (8 * b));
  }
};

// code_of as a table, as the kernels keep it in LDS.
template <int MAP, bool HEADS>
struct HostCodes {
  uint8_t t[256];
  HostCodes() {
    for (int v = 0; v < 256; v++) t[v] = (uint8_t)BffLean<MAP, HEADS>::code_of(v);
  }
  uint32_t operator()(uint32_t v) const { return t[v & 255]; }
};

struct HostAnchor {
  uint32_t *p;
  uint32_t word(uint32_t j) const { return p[j]; }
  void set_word(uint32_t j, uint32_t v) const { p[j] = v; }
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

void to_words(const uint8_t *t, uint32_t *words) {
  for (int w = 0; w < 32; w++)
    words[w] = (uint32_t)t[4 * w] | ((uint32_t)t[4 * w + 1] << 8) |
               ((uint32_t)t[4 * w + 2] << 16) | ((uint32_t)t[4 * w + 3] << 24);
}

// Every word's bytes 1 and 2 agree with its tape byte (codes, next commands
// at command positions): the tables the run keeps up to date equal those a
// fresh build gives, except that the bound of a non-command byte may be
// lower (but at least 1).  Match bytes are only read until a bracket is
// written (then the masks serve), so they are not compared.
template <class F, class L>
bool tables_consistent(const uint32_t *wd, const L &lut) {
  uint32_t fresh[129], words[32];
  uint8_t t[128];
  for (int p = 0; p < 128; p++) t[p] = (uint8_t)wd[p];
  to_words(t, words);
  typename F::Run r;
  for (int j = 0; j < 32; j++) fresh[j] = words[j];
  F::start_raw(r, HostWords{fresh}, lut, 8192);
  for (int p = 0; p < 128; p++) {
    const uint32_t code = (fresh[p] >> 8) & 0xff, have = (wd[p] >> 8) & 0xff;
    if ((code >> 4) == kClsNone) {
      if (have < 1 || have > code) return false;
    } else if (have != code) {
      return false;
    }
    if ((code >> 4) != kClsNone && ((wd[p] >> 16) & 0xff) != ((fresh[p] >> 16) & 0xff)) return false;

  }
  return true;
}

template <int MAP, bool HEADS>
int check_tape(const char *name, int orc_lang, const uint8_t *tape, uint32_t stepcount,
               uint32_t budget, uint64_t slice_seed, int r, uint64_t *over, uint64_t *cap_runs,
               uint64_t *resumed) {
  using F = BffLean<MAP, HEADS>;
  static const HostCodes<MAP, HEADS> lut;
  uint8_t ref[128];
  memcpy(ref, tape, 128);
  uint64_t ref_steps = 0;
  const uint64_t ref_ops = orc_evaluate(orc_lang, ref, stepcount, &ref_steps);
  uint32_t words[32], wd[129];
  to_words(tape, words);
  uint32_t steps = 0, work = 0;
  for (int j = 0; j < 32; j++) wd[j] = words[j];
  // Complete runs go through the full-state cycle skip every other time.
  uint32_t anchor[32];
  const bool full = budget == 0xffffffffu && (r & 1);
  const uint32_t ops = full ? F::evaluate_full(HostWords{wd}, HostAnchor{anchor}, lut, stepcount,
                                               &steps, &work)
                            : F::evaluate(HostWords{wd}, lut, stepcount, &steps, &work, budget);
  if (steps == kOverBudget) {
    (*over)++;
    // A run stopped by its budget goes on to the end from its tape bytes
    // and saved state on a table built afresh (as a deferred interaction
    // does): the oracle's results, with and without the cycle skip.
    typename F::Run run;
    uint32_t ws[129];
    for (int j = 0; j < 32; j++) ws[j] = words[j];
    F::start_raw(run, HostWords{ws}, lut, stepcount);
    F::template advance<false>(run, HostWords{ws}, lut, stepcount, budget);
    if (run.done) {
      printf("%s tape %d: over budget once, not the second time\n", name, r);
      return 1;
    }
    const typename F::Resume sv = F::saved(run);
    uint8_t mid[128];
    for (int p = 0; p < 128; p++) mid[p] = (uint8_t)ws[p];
    uint32_t wr[129];
    to_words(mid, wr);
    uint32_t steps3 = 0, ops3;
    uint32_t anchor3[32];
    if (r & 1) {
      ops3 = F::evaluate_resumed(HostWords{wr}, HostAnchor{anchor3}, lut, stepcount, sv, &steps3,
                                 nullptr);
    } else {
      typename F::Run r3;
      F::start_resume(r3, HostWords{wr}, lut, stepcount, sv);
      F::template advance<false>(r3, HostWords{wr}, lut, stepcount, 0xffffffffu);
      ops3 = F::finish(r3, stepcount, &steps3, nullptr);
    }
    uint8_t out3[128];
    for (int p = 0; p < 128; p++) out3[p] = (uint8_t)wr[p];
    if (ops3 != ref_ops || steps3 != ref_steps || memcmp(out3, ref, 128) != 0) {
      printf("%s tape %d cap %u budget %u: resumed on a new table ops %u/%llu steps %u/%llu "
             "tape %s\n", name, r, stepcount, budget, ops3, (unsigned long long)ref_ops, steps3,
             (unsigned long long)ref_steps, memcmp(out3, ref, 128) ? "differs" : "same");
      return 1;
    }
    (*resumed)++;
    return 0;
  }
  uint8_t out[128];
  for (int p = 0; p < 128; p++) out[p] = (uint8_t)wd[p];
  if (ops != ref_ops || steps != ref_steps || work > ops ||
      (budget != 0xffffffffu && work > budget) || memcmp(out, ref, 128) != 0) {
    printf("%s tape %d cap %u budget %u: ops %u/%llu steps %u/%llu tape %s\n", name, r, stepcount,
           budget, ops, (unsigned long long)ref_ops, steps, (unsigned long long)ref_steps,
           memcmp(out, ref, 128) ? "differs" : "same");
    return 1;
  }
  if (!tables_consistent<F>(wd, lut)) {
    printf("%s tape %d: tables out of step with the tape after the run\n", name, r);
    return 1;
  }
  *cap_runs += steps == stepcount;
  // The same run resumed in slices of 1..64 commands.
  typename F::Run run;
  uint32_t wd2[129];
  for (int j = 0; j < 32; j++) wd2[j] = words[j];
  F::start_raw(run, HostWords{wd2}, lut, stepcount);
  uint64_t calls = 0;
  uint32_t anchor2[32];
  while (!run.done) {
    const uint32_t slice = 1 + (uint32_t)(sm(slice_seed + 77 * calls) % 64);
    if (r & 2)
      F::template advance<true, true>(run, HostWords{wd2}, lut, stepcount, slice,
                                      HostAnchor{anchor2});
    else
      F::template advance<true>(run, HostWords{wd2}, lut, stepcount, slice);
    if (++calls > 100000) {
      printf("%s tape %d: resumed run does not end\n", name, r);
      return 1;
    }
  }
  uint32_t steps2 = 0, work2 = 0;
  const uint32_t ops2 = F::finish(run, stepcount, &steps2, &work2);
  for (int p = 0; p < 128; p++) out[p] = (uint8_t)wd2[p];
  if (ops2 != ref_ops || steps2 != ref_steps || memcmp(out, ref, 128) != 0) {
    printf("%s tape %d cap %u: resumed ops %u/%llu steps %u/%llu tape %s\n", name, r, stepcount,
           ops2, (unsigned long long)ref_ops, steps2, (unsigned long long)ref_steps,
           memcmp(out, ref, 128) ? "differs" : "same");
    return 1;
  }
  return 0;
}

template <int MAP, bool HEADS>
int check(const char *name, int orc_lang, int count, uint64_t seed) {
  uint8_t tape[128];
  uint64_t over = 0, cap_runs = 0, resumed = 0;
  for (int r = 0; r < count; r++) {
    make_tape(MAP, seed * 1000003 + r, tape);
    const uint64_t x = sm(seed ^ (0x9e37ull * (r + 1)));
    const uint32_t caps[4] = {8192, (uint32_t)(x % 300), 1 + (uint32_t)(x >> 20) % 3000, 0};
    const uint32_t stepcount = caps[(x >> 40) & 3];
    const uint32_t budget = ((x >> 44) & 3) == 0 ? 0xffffffffu : 1 + (uint32_t)((x >> 46) % 400);
    if (check_tape<MAP, HEADS>(name, orc_lang, tape, stepcount, budget, x, r, &over, &cap_runs,
                               &resumed))
      return 1;
  }
  printf("%s: %d tapes ok (%llu at the step limit, %llu over budget and resumed on a new "
         "table; complete runs with the cycle skip every other tape, resumed runs every other "
         "pair)\n", name, count, (unsigned long long)cap_runs, (unsigned long long)over);
  if (over != resumed) return 1;
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const int count = argc > 1 ? atoi(argv[1]) : 20000;
  const uint64_t seed = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;
  int bad = 0;
  bad |= check<kMapAscii, true>("bff", ORC_BFF, count, seed);
  bad |= check<kMapAscii, false>("bff_noheads", ORC_BFF_NOHEADS, count, seed + 1);
  bad |= check<kMap4bit, false>("bff_noheads_4bit", ORC_BFF_NOHEADS_4BIT, count, seed + 2);
  bad |= check<kMapPerm, true>("bff_perm", ORC_BFF_PERM, count, seed + 3);
  if (argc > 3) {
    FILE *f = fopen(argv[3], "rb");
    if (!f) {
      perror(argv[3]);
      return 2;
    }
    std::vector<uint8_t> all;
    uint8_t buf[128];
    while (fread(buf, 1, 128, f) == 128) all.insert(all.end(), buf, buf + 128);
    fclose(f);
    uint64_t over = 0, cap_runs = 0, resumed = 0;
    const int n = (int)(all.size() / 128);
    for (int r = 0; r < n && !bad; r++)
      for (uint32_t budget : {0xffffffffu, 768u})
        bad |= check_tape<kMapAscii, false>("file", ORC_BFF_NOHEADS, &all[128 * (size_t)r], 8192,
                                            budget, sm(r), r, &over, &cap_runs, &resumed);
    printf("file: %d tapes ok (%llu at the step limit, %llu over budget)\n", n,
           (unsigned long long)cap_runs, (unsigned long long)over);
  }
  return bad;
}
