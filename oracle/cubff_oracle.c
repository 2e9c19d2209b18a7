/*
 * cubff CPU oracle -- TEST INFRASTRUCTURE ONLY (see cubff_oracle.h).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this code, and only as the checker or the timed CPU baseline; the
 * product path never calls it.
 *
 * Reference citations are file:line in paradigms-of-intelligence/cubff.
 */
#include "cubff_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TAPE 64           /* kSingleTapeSize, common.h:55 */
#define TAPE2 (2 * TAPE)
#define EVAL_STEPS 8192   /* 8 * 1024, common_language.h:184, :231, :250 */

static const char *const kLangNames[ORC_NUM_LANGS] = {
    "bff",          "bff_noheads",  "bff8",
    "bff8_noheads", "bff_noheads_4bit", "bff_perm",
    "bff_selfmove", "forth",        "forthcopy",
    "forthtrivial", "forthtrivial_reset", "subleq",
    "rsubleq4"};

int orc_num_langs(void) { return ORC_NUM_LANGS; }

const char *orc_lang_name(int lang) {
  return (lang >= 0 && lang < ORC_NUM_LANGS) ? kLangNames[lang] : NULL;
}

int orc_lang_id(const char *name) {
  for (int i = 0; i < ORC_NUM_LANGS; i++)
    if (strcmp(name, kLangNames[i]) == 0) return i;
  return -1;
}

/* SplitMix64 finaliser: common_language.h:131-136. */
uint64_t orc_splitmix64(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* The `seed` lambda of RunSimulation: common_language.h:367-369. */
uint64_t orc_epoch_seed(uint64_t params_seed, uint64_t x) {
  return orc_splitmix64(orc_splitmix64(params_seed) ^ orc_splitmix64(x));
}

/* ------------------------------------------------------------------------ */
/* BFF family: bff.inc.h:26-389 and the variant .cu files.                   */

enum { B_LS, B_LE, B_PLUS, B_MINUS, B_COPY01, B_COPY10, B_DEC0, B_INC0,
       B_DEC1, B_INC1, B_NULL, B_NOOP };

/* Default ASCII map, bff.inc.h:58-85.  The reference switches on a signed
 * char against ASCII constants, so bytes >= 128 fall to the default. */
static int bff_kind_ascii(uint8_t b) {
  switch (b) {
    case '[': return B_LS;
    case ']': return B_LE;
    case '+': return B_PLUS;
    case '-': return B_MINUS;
    case '.': return B_COPY01;
    case ',': return B_COPY10;
    case '<': return B_DEC0;
    case '>': return B_INC0;
    case '{': return B_DEC1;
    case '}': return B_INC1;
    case 0: return B_NULL;
    default: return B_NOOP;
  }
}

/* bff_noheads_4bit.cu:22-49: ((int)c + 256) % 16 == byte % 16. */
static int bff_kind_4bit(uint8_t b) {
  static const int m[16] = {B_NULL, B_NOOP, B_NOOP, B_NOOP, B_NOOP, B_NOOP,
                            B_LS, B_LE, B_PLUS, B_MINUS, B_COPY01, B_COPY10,
                            B_DEC0, B_INC0, B_DEC1, B_INC1};
  return m[b % 16];
}

/* bff_perm.cu:23-50. */
static int bff_kind_perm(uint8_t b) {
  static const int m[11] = {B_NULL, B_DEC0, B_INC0, B_DEC1, B_INC1, B_PLUS,
                            B_MINUS, B_COPY01, B_COPY10, B_LS, B_LE};
  return b < 11 ? m[b] : B_NOOP;
}

/* bff_selfmove.cu:24-42. */
static int bff_kind_selfmove(uint8_t b) {
  static const int m[7] = {B_INC0, B_DEC0, B_PLUS, B_MINUS, B_COPY10, B_LS,
                           B_LE};
  return b < 7 ? m[b] : B_NOOP;
}

enum { LOGIC_STD, LOGIC_BFF8, LOGIC_SELFMOVE };

typedef struct {
  int (*kind)(uint8_t);
  int heads;  /* BFF_HEADS */
  int logic;
} BffVariant;

static BffVariant bff_variant(int lang) {
  BffVariant v = {bff_kind_ascii, 0, LOGIC_STD};
  switch (lang) {
    case ORC_BFF: v.heads = 1; break;                          /* bff.cu */
    case ORC_BFF_NOHEADS: break;                               /* bff_noheads.cu */
    case ORC_BFF8: v.heads = 1; v.logic = LOGIC_BFF8; break;   /* bff8.cu */
    case ORC_BFF8_NOHEADS: v.logic = LOGIC_BFF8; break;        /* bff8_noheads.cu */
    case ORC_BFF_NOHEADS_4BIT: v.kind = bff_kind_4bit; break;
    case ORC_BFF_PERM: v.kind = bff_kind_perm; v.heads = 1; break;
    case ORC_BFF_SELFMOVE:
      v.kind = bff_kind_selfmove; v.heads = 1; v.logic = LOGIC_SELFMOVE;
      break;
    default: break;
  }
  return v;
}

/* Forward bracket scan of bff.inc.h:319-328 (also bff_selfmove.cu:66-75). */
static void bff_scan_forward(const BffVariant *v, const uint8_t *tape,
                             int *pc) {
  long depth = 1;
  int p = *pc + 1;
  for (; p < TAPE2 && depth > 0; p++) {
    int k = v->kind(tape[p]);
    if (k == B_LE) depth--;
    if (k == B_LS) depth++;
  }
  p--;
  *pc = depth != 0 ? TAPE2 : p;
}

/* Backward bracket scan of bff.inc.h:333-342 (also bff_selfmove.cu:79-88). */
static void bff_scan_backward(const BffVariant *v, const uint8_t *tape,
                              int *pc) {
  long depth = 1;
  int p = *pc - 1;
  for (; p >= 0 && depth > 0; p--) {
    int k = v->kind(tape[p]);
    if (k == B_LE) depth++;
    if (k == B_LS) depth--;
  }
  p++;
  *pc = depth != 0 ? -1 : p;
}

/* EvaluateOne: bff.inc.h:289-349, bff8.cu:23-65, bff_selfmove.cu:44-98.
 * Returns 1 if the op was not a comment. */
static int bff_step(const BffVariant *v, uint8_t *tape, int *h0, int *h1,
                    int *pc) {
  int k = v->kind(tape[*pc]);
  if (v->logic == LOGIC_SELFMOVE) {
    switch (k) {
      case B_DEC0: (*h0)--; return 1;
      case B_INC0: (*h0)++; return 1;
      case B_PLUS: tape[*h0]++; return 1;
      case B_MINUS: tape[*h0]--; return 1;
      case B_COPY10: tape[*h0] = tape[*h1]; (*h1)++; return 1;
      case B_LS:
        if (!tape[*h0]) bff_scan_forward(v, tape, pc);
        return 1;
      case B_LE:
        if (tape[*h0]) bff_scan_backward(v, tape, pc);
        return 1;
      default: return 0;
    }
  }
  switch (k) {
    case B_DEC0: (*h0)--; return 1;
    case B_INC0: (*h0)++; return 1;
    case B_DEC1: (*h1)--; return 1;
    case B_INC1: (*h1)++; return 1;
    case B_PLUS: tape[*h0]++; return 1;
    case B_MINUS: tape[*h0]--; return 1;
    case B_COPY01: tape[*h1] = tape[*h0]; return 1;
    case B_COPY10: tape[*h0] = tape[*h1]; return 1;
    case B_LS:
      if (v->logic == LOGIC_BFF8) {
        if (!tape[*h0]) *pc += 8;
      } else if (v->kind(tape[*h0]) == B_NULL) {
        bff_scan_forward(v, tape, pc);
      }
      return 1;
    case B_LE:
      if (v->logic == LOGIC_BFF8) {
        if (tape[*h0]) *pc -= 8;
      } else if (v->kind(tape[*h0]) != B_NULL) {
        bff_scan_backward(v, tape, pc);
      }
      return 1;
    default: return 0;
  }
}

/* Step trace of the BFF family: before every step, pc, both heads (masked)
 * and the tape, 140 bytes per state (u32 little-endian fields), the layout
 * of python/bff_interpreter.py:evaluate_and_save (:327-393) after its 12-byte
 * header.  Same loop as bff_evaluate below; returns ops, *states = records
 * written (at most cap). */
uint64_t orc_trace_bff(int lang, uint8_t *tape, uint64_t stepcount, uint8_t *out,
                       uint64_t cap, uint64_t *states, uint64_t *steps_out) {
  BffVariant v = bff_variant(lang);
  int h0, h1, pc;
  if (v.heads) {
    h0 = tape[0] % TAPE2;
    h1 = tape[1] % TAPE2;
    pc = 2;
  } else {
    h0 = TAPE2;
    h1 = TAPE2;
    pc = 0;
  }
  uint64_t i = 0, nskip = 0, rec = 0;
  for (; i < stepcount; i++) {
    h0 &= TAPE2 - 1;
    h1 &= TAPE2 - 1;
    if (rec < cap) {
      uint8_t *o = out + rec * (12 + TAPE2);
      const uint32_t f[3] = {(uint32_t)pc, (uint32_t)h0, (uint32_t)h1};
      for (int q = 0; q < 3; q++)
        for (int b = 0; b < 4; b++) o[4 * q + b] = (uint8_t)(f[q] >> (8 * b));
      memcpy(o + 12, tape, TAPE2);
      rec++;
    }
    if (!bff_step(&v, tape, &h0, &h1, &pc)) nskip++;
    if (pc < 0) { i++; break; }
    pc++;
    if (pc >= TAPE2) { i++; break; }
  }
  if (states) *states = rec;
  if (steps_out) *steps_out = i;
  return i - nskip;
}

/* Evaluate: bff.inc.h:356-389, InitialState :269-283. */
static uint64_t bff_evaluate(int lang, uint8_t *tape, uint64_t stepcount,
                             uint64_t *steps_out) {
  BffVariant v = bff_variant(lang);
  int h0, h1, pc;
  if (v.heads) {
    h0 = tape[0] % TAPE2;
    h1 = tape[1] % TAPE2;
    pc = 2;
  } else {
    h0 = TAPE2;
    h1 = TAPE2;
    pc = 0;
  }
  uint64_t i = 0, nskip = 0;
  for (; i < stepcount; i++) {
    h0 &= TAPE2 - 1;
    h1 &= TAPE2 - 1;
    if (!bff_step(&v, tape, &h0, &h1, &pc)) nskip++;
    if (pc < 0) { i++; break; }
    pc++;
    if (pc >= TAPE2) { i++; break; }
  }
  if (steps_out) *steps_out = i;
  return i - nskip;
}

/* ------------------------------------------------------------------------ */
/* Forth family: forth.inc.h:26-397 and the variant .cu files.               */

enum { F_WRITE0, F_WRITE1, F_WRITE, F_READ, F_READ0, F_READ1, F_DUP, F_DROP,
       F_SWAP, F_IF0, F_INC, F_DEC, F_ADD, F_SUB, F_COPY, F_COPY0, F_COPY1,
       F_XOR, F_CONST, F_JMP, F_NOOP };

/* forth.inc.h:54-84 (forth), forthcopy.cu:21-51, forthtrivial.cu:21-55,
 * forthtrivial_reset.cu:22-56. */
static int forth_kind(int lang, uint8_t c) {
  static const int k_forth[12] = {F_READ0, F_READ1, F_WRITE0, F_WRITE1, F_DUP,
                                  F_DROP, F_SWAP, F_IF0, F_INC, F_DEC, F_ADD,
                                  F_SUB};
  static const int k_copy[12] = {F_READ, F_WRITE, F_COPY, F_XOR, F_DUP, F_DROP,
                                 F_SWAP, F_IF0, F_INC, F_DEC, F_ADD, F_SUB};
  if (c >= 128) return F_JMP;
  if (c >= 64) return F_CONST;
  if (c < 12) return lang == ORC_FORTHCOPY ? k_copy[c] : k_forth[c];
  if (lang == ORC_FORTHTRIVIAL || lang == ORC_FORTHTRIVIAL_RESET) {
    if (c == 0xC) return F_COPY0;
    if (c == 0xD) return F_COPY1;
  }
  return F_NOOP;
}

typedef struct {
  uint8_t data[128];  /* Stack::kStackSize, forth.inc.h:237 */
  int pos;
  int overflow;
} FStack;

static void fpush(FStack *s, uint8_t v) {  /* forth.inc.h:241-247 */
  if (s->pos != 128) s->data[s->pos++] = v;
  else s->overflow = 1;
}
static uint8_t fpop(FStack *s) {            /* forth.inc.h:248 */
  return s->pos == 0 ? 0 : s->data[--s->pos];
}

/* Reset: forthtrivial_reset.cu:58-68. */
static void forth_reset(uint8_t *tape, int pos) {
  uint64_t seed = 0;
  for (int i = 0; i < TAPE2; i++) seed = orc_splitmix64(seed ^ tape[i]);
  for (int i = 0; i < TAPE; i++)
    tape[i] = (uint8_t)orc_splitmix64(seed * TAPE * TAPE +
                                      (uint64_t)(int64_t)(pos * TAPE) +
                                      (uint64_t)i);
}

/* EvaluateOne: forth.inc.h:251-373; forthtrivial_reset.cu:123-252. */
static void forth_step(int lang, uint8_t *tape, int *pos, uint64_t *nops,
                       FStack *st) {
  uint8_t cmd = tape[*pos];
  int reset = lang == ORC_FORTHTRIVIAL_RESET;
  switch (forth_kind(lang, cmd)) {
    case F_READ0: { int a = fpop(st) % TAPE; fpush(st, tape[a]); break; }
    case F_READ1: { int a = fpop(st) % TAPE; fpush(st, tape[TAPE + a]); break; }
    case F_READ: { int a = fpop(st) % TAPE2; fpush(st, tape[a]); break; }
    case F_WRITE: {
      int val = fpop(st);
      int a = fpop(st) % TAPE2;
      tape[a] = (uint8_t)val;
      break;
    }
    case F_WRITE0: {
      int val = fpop(st);
      int a = fpop(st) % TAPE;
      tape[a] = (uint8_t)val;
      break;
    }
    case F_WRITE1: {
      int val = fpop(st);
      int a = fpop(st) % TAPE;
      if (reset && tape[TAPE + a] == 0x13) forth_reset(tape, *pos);
      else tape[TAPE + a] = (uint8_t)val;
      break;
    }
    case F_DUP: { int v = fpop(st); fpush(st, v); fpush(st, v); break; }
    case F_DROP: fpop(st); break;
    case F_SWAP: {
      int a = fpop(st), b = fpop(st);
      fpush(st, a);
      fpush(st, b);
      break;
    }
    case F_IF0: {
      int v = fpop(st);
      if (v) (*pos)++;
      fpush(st, v);
      break;
    }
    case F_INC: fpush(st, (uint8_t)(fpop(st) + 1)); break;
    case F_DEC: fpush(st, (uint8_t)(fpop(st) - 1)); break;
    case F_ADD: { int a = fpop(st), b = fpop(st); fpush(st, (uint8_t)(a + b)); break; }
    case F_SUB: { int a = fpop(st), b = fpop(st); fpush(st, (uint8_t)(a - b)); break; }
    case F_CONST: fpush(st, cmd & 63); break;
    case F_COPY0: {
      int a = fpop(st) % TAPE;
      if (reset && tape[TAPE + a] == 0x13) forth_reset(tape, *pos);
      else tape[TAPE + a] = tape[a];
      break;
    }
    case F_COPY1: { int a = fpop(st) % TAPE; tape[a] = tape[TAPE + a]; break; }
    case F_COPY: {
      int to = fpop(st) % TAPE2;
      int from = fpop(st) % TAPE2;
      tape[to] = tape[from];
      break;
    }
    case F_XOR: fpush(st, fpop(st) ^ 64); break;
    case F_JMP: {
      int a = (cmd & 63) + 1;
      *pos += (cmd & 64) ? -a : a;
      (*pos)--;
      break;
    }
    default: (*nops)++; break;
  }
  (*pos)++;
}

/* Evaluate: forth.inc.h:379-397. */
static uint64_t forth_evaluate(int lang, uint8_t *tape, uint64_t stepcount,
                               uint64_t *steps_out) {
  FStack st;
  memset(&st, 0, sizeof(st));
  int pos = 0;
  uint64_t i = 0, nops = 0;
  for (; i < stepcount; i++) {
    forth_step(lang, tape, &pos, &nops, &st);
    if (pos >= TAPE2 || pos < 0 || st.overflow) { i++; break; }
  }
  if (steps_out) *steps_out = i;
  return i - nops;
}

/* ------------------------------------------------------------------------ */
/* SUBLEQ: subleq.cu:109-133.  RSUBLEQ4: rsubleq4.cu:116-141.                */

static uint64_t subleq_evaluate(uint8_t *tape, uint64_t stepcount,
                                uint64_t *steps_out) {
  int pos = 0;
  uint64_t i = 0;
  for (; i < stepcount; i++) {
    int sa = tape[pos] % TAPE2;
    int sb = tape[pos + 1] % TAPE2;
    tape[sa] -= tape[sb];
    if ((tape[sa] & 0x80) || tape[sa] == 0) pos = tape[pos + 2];
    else pos += 3;
    if (pos + 3 > TAPE2 || pos < 0) { i++; break; }
  }
  if (steps_out) *steps_out = i;
  return i;
}

static uint64_t rsubleq4_evaluate(uint8_t *tape, uint64_t stepcount,
                                  uint64_t *steps_out) {
  int pos = 0;
  uint64_t i = 0;
  for (; i < stepcount; i++) {
    int sa = (pos + tape[pos]) % TAPE2;
    int sb = (pos + tape[pos + 1]) % TAPE2;
    int sc = (pos + tape[pos + 2]) % TAPE2;
    tape[sa] = (uint8_t)(tape[sb] - tape[sc]);
    if ((tape[sa] & 0x80) || tape[sa] == 0) pos += (int8_t)tape[pos + 3];
    else pos += 4;
    if (pos + 4 > TAPE2 || pos < 0) { i++; break; }
  }
  if (steps_out) *steps_out = i;
  return i;
}

uint64_t orc_evaluate(int lang, uint8_t *tape, uint64_t stepcount,
                      uint64_t *steps_out) {
  switch (lang) {
    case ORC_BFF: case ORC_BFF_NOHEADS: case ORC_BFF8: case ORC_BFF8_NOHEADS:
    case ORC_BFF_NOHEADS_4BIT: case ORC_BFF_PERM: case ORC_BFF_SELFMOVE:
      return bff_evaluate(lang, tape, stepcount, steps_out);
    case ORC_FORTH: case ORC_FORTHCOPY: case ORC_FORTHTRIVIAL:
    case ORC_FORTHTRIVIAL_RESET:
      return forth_evaluate(lang, tape, stepcount, steps_out);
    case ORC_SUBLEQ: return subleq_evaluate(tape, stepcount, steps_out);
    case ORC_RSUBLEQ4: return rsubleq4_evaluate(tape, stepcount, steps_out);
    default: abort();
  }
}

/* ------------------------------------------------------------------------ */
/* Soup-level operations.                                                     */

/* InitPrograms: common_language.h:138-155. */
void orc_init_programs(uint8_t *soup, uint64_t n, uint64_t seed_value,
                       int zero_init) {
  for (uint64_t i = 0; i < n; i++)
    for (uint64_t j = 0; j < TAPE; j++)
      soup[i * TAPE + j] =
          zero_init ? 0
                    : (uint8_t)orc_splitmix64(TAPE * n * seed_value +
                                              TAPE * i + j);
}

/* Fisher-Yates over the identity: common_language.h:418-424, :464-467, :476.
 * `do_shuffle` ignores its base_seed argument and uses the captured epoch. */
void orc_pairing_shuffle(uint32_t *s, uint64_t n, uint64_t params_seed,
                         uint64_t epoch) {
  for (uint64_t i = 0; i < n; i++) s[i] = (uint32_t)i;
  for (uint64_t i = n; i-- > 0;) {
    uint64_t j =
        orc_splitmix64(orc_epoch_seed(params_seed, epoch * n + i)) % (i + 1);
    uint32_t t = s[i];
    s[i] = s[j];
    s[j] = t;
  }
}

/* --fixed_shuffle: common_language.h:468-474. */
void orc_pairing_fixed(uint32_t *s, uint64_t n, uint64_t epoch) {
  uint64_t flip = epoch & 1;
  uint64_t max_pow2 = 31 - __builtin_clz((unsigned)n);
  /* (1 << k) - 1 with int arithmetic; k <= 31 for any n < 2^32. */
  int k = (int)(epoch % max_pow2 + 1);
  uint64_t offset = (uint64_t)(int64_t)(int32_t)((uint32_t)1 << k) - 1;
  if (k == 31) offset = 0x7fffffffULL;
  for (uint64_t i = 0; i < n; i++) s[i] = (uint32_t)(((i * offset) % n) ^ flip);
}

/* --nopermute_programs: common_language.h:478-486. */
void orc_pairing_cyclic(uint32_t *s, uint64_t n, uint64_t epoch) {
  for (uint64_t i = 0; i < n; i++)
    s[i] = (epoch % 2 == 1) ? (uint32_t)i
                            : (uint32_t)(i == 0 ? n - 1 : i - 1);
}

/* 2-D greedy neighbour pairing: common_language.h:435-463. */
uint64_t orc_pairing_2d(uint32_t *s, uint64_t n, const uint64_t *row_off,
                        const uint32_t *nbr, uint64_t params_seed,
                        uint64_t epoch) {
  uint32_t *order = (uint32_t *)malloc(n * sizeof(uint32_t));
  char *used = (char *)calloc(n, 1);
  orc_pairing_shuffle(order, n, params_seed, epoch);
  uint64_t num = 0;
  for (uint64_t k = 0; k < n; k++) {
    uint32_t i = order[k];
    uint64_t deg = row_off[i + 1] - row_off[i];
    if (deg == 0) continue;
    uint64_t idx = orc_epoch_seed(params_seed,
                                  orc_epoch_seed(params_seed, epoch) ^
                                      orc_epoch_seed(params_seed, i)) %
                   deg;
    uint32_t nb = nbr[row_off[i] + idx];
    if (used[i] || used[nb]) continue;
    used[i] = used[nb] = 1;
    s[num * 2] = i;
    s[num * 2 + 1] = nb;
    num++;
  }
  uint64_t idx = num * 2;
  for (uint64_t i = 0; i < n && idx < n; i++)
    if (!used[i]) s[idx++] = (uint32_t)i;
  free(order);
  free(used);
  return num;
}

/* MutateAndRunPrograms: common_language.h:157-193. */
/* Programs that sit in two or more different slots of one epoch's pairing
 * s[0 .. 2*(n/2)).  Slots of one epoch run in parallel and write their
 * programs back, so such a program is written by several slots at once
 * (common_language.h:165-192 relies on s being a permutation; a 2-D pattern
 * line "a a" puts one program in both halves of a single slot, which is no
 * race: p2's half wins).  Debug builds (ORC_DEBUG_CHECKS) check every epoch
 * with it. */
uint64_t orc_pairing_repeats(const uint32_t *s, uint64_t n) {
  uint64_t slots = n / 2, repeats = 0;
  uint64_t *last = (uint64_t *)malloc(n * sizeof(uint64_t));
  if (!last) return 0;
  for (uint64_t p = 0; p < n; p++) last[p] = UINT64_MAX;
  for (uint64_t k = 0; k < slots; k++) {
    for (int h = 0; h < 2; h++) {
      const uint32_t p = s[2 * k + h];
      if (p >= n) continue;
      if (last[p] != UINT64_MAX && last[p] != k) repeats++;
      last[p] = k;
    }
  }
  free(last);
  return repeats;
}

static uint64_t g_debug_repeats = 0;
uint64_t orc_debug_repeats_seen(void) { return g_debug_repeats; }

uint64_t orc_mutate_and_run(int lang, uint8_t *soup, const uint32_t *s,
                            uint64_t n, uint64_t num_indices,
                            uint64_t seed_value, uint32_t mutation_prob,
                            uint32_t *pair_ops, uint32_t *pair_steps,
                            int nthreads) {
  uint64_t total = 0;
  long long slots = (long long)(n / 2);
#ifdef ORC_DEBUG_CHECKS
  {
    const uint64_t rep = orc_pairing_repeats(s, n);
    if (rep) {
      g_debug_repeats += rep;
      fprintf(stderr, "orc debug: the epoch's pairing puts %llu programs in more than one slot "
                      "(their slots race on the write back)\n", (unsigned long long)rep);
    }
  }
#endif
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) \
    num_threads(nthreads)
#endif
  for (long long k = 0; k < slots; k++) {
    uint8_t tape[TAPE2];
    uint32_t p1 = s[2 * k], p2 = s[2 * k + 1];
    memcpy(tape, soup + (uint64_t)p1 * TAPE, TAPE);
    memcpy(tape + TAPE, soup + (uint64_t)p2 * TAPE, TAPE);
    for (uint64_t i = 0; i < TAPE2; i++) {
      uint64_t r = orc_splitmix64((n * seed_value + (uint64_t)k) * TAPE2 + i);
      if (((r >> 8) & ((1ULL << 30) - 1)) < mutation_prob)
        tape[i] = (uint8_t)(r & 0xFF);
    }
    uint64_t ops = 0, steps = 0;
    if ((uint64_t)k < num_indices)
      ops = orc_evaluate(lang, tape, EVAL_STEPS, &steps);
    /* Byte-interleaved write back: p2's half wins if p1 == p2 (:188-191). */
    for (uint64_t i = 0; i < TAPE; i++) {
      soup[(uint64_t)p1 * TAPE + i] = tape[i];
      soup[(uint64_t)p2 * TAPE + i] = tape[i + TAPE];
    }
    if (pair_ops) pair_ops[k] = (uint32_t)ops;
    if (pair_steps) pair_steps[k] = (uint32_t)steps;
    total += ops;
  }
  (void)nthreads;
  return total;
}

/* CheckSelfRep: common_language.h:203-277. */
static uint64_t selfrep_one(int lang, const uint8_t *prog, uint64_t n,
                            uint64_t seed_value, uint64_t index) {
  enum { ITERS = 13, GENS = 4 };
  uint8_t tapes[ITERS][TAPE2];
  uint64_t local = orc_splitmix64(n * seed_value + index);
  for (uint64_t it = 0; it < ITERS; it++) {
    uint8_t noise[TAPE];
    for (int j = 0; j < TAPE; j++)
      noise[j] = (uint8_t)orc_splitmix64(
          local ^ orc_splitmix64((it + 1) * TAPE + (uint64_t)j));
    uint8_t *t = tapes[it];
    memcpy(t, prog, TAPE);
    memcpy(t + TAPE, noise, TAPE);
    orc_evaluate(lang, t, EVAL_STEPS, NULL);
    for (int g = 0; g < GENS; g++) {
      memmove(t, t + TAPE, TAPE);
      memcpy(t + TAPE, noise, TAPE);
      orc_evaluate(lang, t, EVAL_STEPS, NULL);
    }
  }
  uint64_t res[2] = {0, 0};
  for (int i = 0; i < TAPE2; i++) {
    for (int a = 0; a < ITERS; a++) {
      uint64_t count = 1;
      if (i < TAPE && tapes[a][i] != prog[i]) continue;
      for (int b = a + 1; b < ITERS; b++)
        if (tapes[a][i] == tapes[b][i]) count++;
      if (count > ITERS / 4) {
        res[i / TAPE]++;
        break;
      }
    }
  }
  return res[0] < res[1] ? res[0] : res[1];
}

void orc_check_selfrep(int lang, const uint8_t *soup, uint64_t n,
                       uint64_t count, uint64_t seed_value, uint64_t *result,
                       int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
  for (long long i = 0; i < (long long)count; i++)
    result[i] = selfrep_one(lang, soup + (uint64_t)i * TAPE, n, seed_value,
                            (uint64_t)i);
  (void)nthreads;
}
