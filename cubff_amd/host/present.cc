// Host-side presentation of the 13 languages (see present.h).  Output formats,
// glyphs and palettes follow the reference's language files so that logs,
// dashboards and Python callers see the same text:
//   BFF family   bff.inc.h:91-283 and the map overrides in the variant files
//   Forth family forth.inc.h:93-234, forth.cu / forthcopy.cu / forthtrivial.cu
//                / forthtrivial_reset.cu (InitByteColors)
//   SUBLEQ       subleq.cu:26-107, rsubleq4.cu:26-114
#include "present.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>

#include "cubff.h"

namespace cubff_host {
namespace {

// ---------------------------------------------------------------------------
// BFF family.

enum BffKind { kLS, kLE, kPlus, kMinus, kC01, kC10, kDec0, kInc0, kDec1, kInc1, kNull, kNoop };
enum BffMapId { kAscii, k4bit, kPerm, kSelfmove };

// GetOpKind of each map on a (signed) char, as the reference's switches see it.
BffKind bff_kind(int map, char c) {
  switch (map) {
    case kAscii:
      switch (c) {
        case '[': return kLS;
        case ']': return kLE;
        case '+': return kPlus;
        case '-': return kMinus;
        case '.': return kC01;
        case ',': return kC10;
        case '<': return kDec0;
        case '>': return kInc0;
        case '{': return kDec1;
        case '}': return kInc1;
        case 0: return kNull;
        default: return kNoop;
      }
    case k4bit: {
      const int m = ((int)c + 256) % 16;
      if (m == 0) return kNull;
      if (m < 6) return kNoop;
      return (BffKind)(m - 6);
    }
    case kPerm: {
      static const BffKind t[11] = {kNull, kDec0, kInc0, kDec1, kInc1, kPlus,
                                    kMinus, kC01, kC10, kLS, kLE};
      return (c >= 0 && c <= 10) ? t[(int)c] : kNoop;
    }
    default: {
      static const BffKind t[7] = {kInc0, kDec0, kPlus, kMinus, kC10, kLS, kLE};
      return (c >= 0 && c <= 6) ? t[(int)c] : kNoop;
    }
  }
}

const char *kCommandRepr = "[]+-.,<>{}";

// Glyph of a non-command byte: U+0100 + b, except 196-204 -> A-I and
// 241-243 -> J-L (bff.inc.h:93-134).
struct GlyphTable {
  char s[256][4];
  GlyphTable() {
    for (int b = 0; b < 256; b++) {
      if (b >= 196 && b <= 204) {
        s[b][0] = (char)('A' + (b - 196));
        s[b][1] = 0;
      } else if (b >= 241 && b <= 243) {
        s[b][0] = (char)('J' + (b - 241));
        s[b][1] = 0;
      } else {
        const unsigned cp = 0x100 + b;
        s[b][0] = (char)(0xC0 | (cp >> 6));
        s[b][1] = (char)(0x80 | (cp & 0x3F));
        s[b][2] = 0;
      }
    }
  }
};
const GlyphTable &glyphs() {
  static const GlyphTable g;
  return g;
}

struct BffPresenter : Presenter {
  int map;
  bool heads;
  BffPresenter(int m, bool h) : map(m), heads(h) {}

  std::vector<uint8_t> Parse(const std::string &bff) const override {
    std::vector<uint8_t> ret;
    char command_bytes[kNoop] = {};
    for (int i = 0; i < 256; i++) {
      const BffKind k = bff_kind(map, (char)i);
      if (k < kNoop) command_bytes[k] = (char)i;
    }
    for (size_t i = 0; i < bff.size();) {
      if (bff[i] == '0') {
        ret.push_back((uint8_t)command_bytes[kNull]);
        i++;
        continue;
      }
      const char *cmd = strchr(kCommandRepr, bff[i]);
      if (cmd && *cmd) {
        ret.push_back((uint8_t)command_bytes[cmd - kCommandRepr]);
        i++;
        continue;
      }
      bool found = false;
      for (int j = 0; j < 256 && !found; j++) {
        const char *s = glyphs().s[j];
        const size_t l = strlen(s);
        if (bff.compare(i, l, s) == 0) {
          ret.push_back((uint8_t)j);
          i += l;
          found = true;
        }
      }
      if (!found) {
        // The reference prints this and then never advances; here the
        // character is skipped so the call returns.
        fprintf(stderr, "Invalid BFF program, character %d not recognized: %s\n", (int)i,
                bff.c_str());
        i++;
      }
    }
    return ret;
  }

  const char *MapChar(uint8_t c, char *mem) const override {
    const BffKind k = bff_kind(map, (char)c);
    if (k < kNull) {
      mem[0] = kCommandRepr[k];
      mem[1] = 0;
      return mem;
    }
    if (k == kNull) return "0";
    return glyphs().s[c];
  }

  void PrintInternal(size_t h0, size_t h1, size_t pc, const uint8_t *mem, size_t len,
                     const size_t *seps, size_t nseps) const {
    size_t sep = 0;
    for (size_t i = 0; i < len; i++) {
      if (sep < nseps && seps[sep] == i) {
        printf("   ");
        sep++;
      }
      const BffKind k = bff_kind(map, (char)mem[i]);
      if (k == kNull) printf("\x1b[38:5:88m");
      else if (k < kNull) printf("\x1b[38:5:255m");
      else printf("\x1b[38:5:237m");
      if (i == h0) printf("\x1b[48:5:27m");
      if (i == h1) printf("\x1b[48:5:196m");
      if (i == pc) printf("\x1b[48:5:34m");
      char chmem[32] = {};
      printf("%s%s", MapChar(mem[i], chmem), ResetColors());
    }
    printf("\n");
  }

  // bff.inc.h:368-372.
  void PrintStep(const uint8_t *rec) const override {
    PrintInternal(step_field(rec, 1), step_field(rec, 2), step_field(rec, 0), rec + 16,
                  2 * kSingleTapeSize, nullptr, 0);
  }

  void PrintProgram(size_t pc_pos, const uint8_t *mem, size_t len, const size_t *seps,
                    size_t nseps) const override {
    size_t h0 = 2 * kSingleTapeSize, h1 = 2 * kSingleTapeSize;
    if (heads && len >= 2) {  // InitialState, bff.inc.h:269-275
      h0 = mem[0] % (2 * kSingleTapeSize);
      h1 = mem[1] % (2 * kSingleTapeSize);
    }
    PrintInternal(h0, h1, pc_pos, mem, len, seps, nseps);
  }

  void InitByteColors(std::array<std::array<uint8_t, 3>, 256> &colors) const override {
    for (int i = 0; i < 256; i++) {
      const uint8_t v = (uint8_t)(192 + i / 4);
      switch (bff_kind(map, (char)i)) {
        case kLS: case kLE: colors[i] = {0, 192, 0}; break;
        case kPlus: case kMinus: case kC01: case kC10: colors[i] = {200, 0, 200}; break;
        case kDec0: case kDec1: case kInc0: case kInc1: colors[i] = {200, 128, 220}; break;
        case kNull: colors[i] = {255, 0, 0}; break;
        default: colors[i] = {v, v, v};
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Hex-coded languages (Forth family, SUBLEQ family).

std::vector<uint8_t> parse_hex(const std::string &hex) {
  auto fromhex = [](char c) -> int {
    if ('0' <= c && c <= '9') return c - '0';
    if ('a' <= c && c <= 'f') return c - 'a' + 10;
    if ('A' <= c && c <= 'F') return c - 'A' + 10;
    fprintf(stderr, "Unknown input character: %c\n", c);
    exit(1);
  };
  std::vector<uint8_t> ret;
  for (size_t i = 0; i < hex.size(); i += 2) {
    const char lo = i + 1 < hex.size() ? hex[i + 1] : '\0';
    ret.push_back((uint8_t)(fromhex(hex[i]) * 16 + fromhex(lo)));
  }
  return ret;
}

const char *hex_char(uint8_t c, char *mem) {
  static const char *chars = "0123456789ABCDEF";
  mem[0] = chars[c / 16];
  mem[1] = chars[c % 16];
  mem[2] = 0;
  return mem;
}

enum ForthKind { fW0, fW1, fW, fR, fR0, fR1, fDup, fDrop, fSwap, fIf0, fInc, fDec, fAdd,
                 fSub, fCopy, fCopy0, fCopy1, fXor, fConst, fJmp, fNoop };
enum ForthVariant { kForth, kForthcopy, kForthtrivial, kForthtrivialReset };

ForthKind forth_kind(int var, uint8_t c) {
  if (c >= 128) return fJmp;
  if (c >= 64) return fConst;
  if (c < 4) {
    static const ForthKind a[4] = {fR0, fR1, fW0, fW1}, b[4] = {fR, fW, fCopy, fXor};
    return var == kForthcopy ? b[c] : a[c];
  }
  if (c < 12) {
    static const ForthKind m[8] = {fDup, fDrop, fSwap, fIf0, fInc, fDec, fAdd, fSub};
    return m[c - 4];
  }
  if (var >= kForthtrivial && c == 12) return fCopy0;
  if (var >= kForthtrivial && c == 13) return fCopy1;
  return fNoop;
}

struct ForthPresenter : Presenter {
  int var;
  explicit ForthPresenter(int v) : var(v) {}
  std::vector<uint8_t> Parse(const std::string &p) const override { return parse_hex(p); }
  const char *MapChar(uint8_t c, char *mem) const override { return hex_char(c, mem); }

  void PrintProgram(size_t pc_pos, const uint8_t *mem, size_t len, const size_t *seps,
                    size_t nseps) const override {
    PrintInternal(pc_pos, mem, len, seps, nseps, nullptr, 0);
  }

  // forth.inc.h:385-388: the tape, then the stack bottom to top.
  void PrintStep(const uint8_t *rec) const override {
    PrintInternal(step_field(rec, 0), rec + 16, 2 * kSingleTapeSize, nullptr, 0,
                  rec + 16 + 2 * kSingleTapeSize, step_field(rec, 1));
  }

  void PrintInternal(size_t pc_pos, const uint8_t *mem, size_t len, const size_t *seps,
                     size_t nseps, const uint8_t *stack, size_t stack_len) const {
    size_t sep = 0;
    for (size_t i = 0; i < len; i++) {
      if (sep < nseps && seps[sep] == i) {
        printf("   ");
        sep++;
      }
      const uint8_t c = mem[i];
      const ForthKind k = forth_kind(var, c);
      char chmem[4];
      if (i == pc_pos) printf("\x1b[48:5:22m");
      if (k == fJmp) {
        printf("\x1b[38:5:%dm\x1b]8;;%d%s%s%d\x1b\\%s\x1b]8;;\x1b\\", (c & 64) ? 221 : 117,
               (int)i, "Jmp", (c & 64) ? "-" : "+", (c & 63) + 1, hex_char(c, chmem));
      } else if (k == fConst) {
        printf("\x1b[38:5:255m\x1b]8;;%d%s%d\x1b\\%s\x1b]8;;\x1b\\", (int)i, "Const", c & 63,
               hex_char(c, chmem));
      } else {
        int color = 255;
        switch (k) {
          case fW: case fW0: case fW1: color = 207; break;
          case fR: case fR0: case fR1: color = 112; break;
          case fCopy: case fCopy0: case fCopy1: color = 141; break;
          case fXor: color = 196; break;
          case fNoop: color = 237; break;
          default: color = 255;
        }
        printf("\x1b[38:5:%dm%s", color, hex_char(c, chmem));
      }
      printf("%s", ResetColors());
    }
    if (stack) {
      printf("   ");
      for (size_t i = 0; i < stack_len; i++) {
        char chmem[4];
        printf("%s ", hex_char(stack[i], chmem));
      }
    }
    printf("\n");
  }

  void InitByteColors(std::array<std::array<uint8_t, 3>, 256> &col) const override {
    if (var == kForth || var == kForthcopy) {
      for (int i = 0; i < 256; i++) col[i] = {(uint8_t)i, 0, (uint8_t)(255 - i)};
      for (int i = 0; i < 0xC; i++) col[i][1] = 128;
      return;
    }
    // forthtrivial.cu:58-101: hue per op class, darkened along each range.
    auto scale = [](std::array<uint8_t, 3> &c, size_t offset, size_t num) {
      const float mult = 1.0f - offset * 0.2f / (num - 1);
      for (int k = 0; k < 3; k++)
        c[k] = (uint8_t)std::round(std::min(255.0f, std::max(0.0f, mult * c[k])));
    };
    for (int i : {0x00, 0x01, 0x02, 0x03, 0x0C, 0x0D}) col[i] = {0x73, 0x01, 0xce};
    for (int i : {0x04, 0x05, 0x06, 0x08, 0x09, 0x0A, 0x0B}) col[i] = {0x4e, 0x10, 0x01};
    col[0x07] = {0x00, 0x00, 0x00};
    for (size_t i = 0x80; i < 0xC0; i++) {
      col[i] = {0x94, 0xd9, 0xff};
      scale(col[i], i - 0x80, 0x40);
    }
    for (size_t i = 0xC0; i < 0x100; i++) {
      col[i] = {0xff, 0x77, 0x7d};
      scale(col[i], i - 0xC0, 0x40);
    }
    for (size_t i = 0x40; i < 0x80; i++) {
      col[i] = {0xff, 0xff, 0xff};
      scale(col[i], i - 0x40, 0x40);
    }
    for (size_t i = 0x0E; i < 0x40; i++) {
      col[i] = {0x02, 0x8a, 0x37};
      scale(col[i], i - 0x0E, 50);
    }
  }
};

struct SubleqPresenter : Presenter {
  bool rel4;  // rsubleq4
  explicit SubleqPresenter(bool r) : rel4(r) {}
  std::vector<uint8_t> Parse(const std::string &p) const override { return parse_hex(p); }
  const char *MapChar(uint8_t c, char *mem) const override { return hex_char(c, mem); }

  void PrintProgram(size_t pc, const uint8_t *mem, size_t len, const size_t *seps,
                    size_t nseps) const override {
    const size_t none = 2 * kSingleTapeSize;
    PrintInternal(none, none, none, pc, mem, len, seps, nseps);
  }

  // subleq.cu:114-118 (heads a, b), rsubleq4.cu:122-126 (heads a, b, c).
  void PrintStep(const uint8_t *rec) const override {
    PrintInternal(step_field(rec, 1), step_field(rec, 2),
                  rel4 ? step_field(rec, 3) : 2 * kSingleTapeSize, step_field(rec, 0), rec + 16,
                  2 * kSingleTapeSize, nullptr, 0);
  }

  // subleq.cu:66-96 and rsubleq4.cu:70-104: the instruction at pc, then the
  // operand heads; a later colour code overrides an earlier one.
  void PrintInternal(size_t h0, size_t h1, size_t h2, size_t pc, const uint8_t *mem,
                     size_t len, const size_t *seps, size_t nseps) const {
    size_t sep = 0;
    for (size_t i = 0; i < len; i++) {
      if (sep < nseps && seps[sep] == i) {
        printf("   ");
        sep++;
      }
      bool color = false;
      const size_t span = rel4 ? 4 : 3;
      if (i >= pc && i < pc + span) {
        printf(rel4 ? "\x1b[33;1m" : "\x1b[32;1m");
        color = true;
      }
      if (rel4) {
        if (i == h2) { printf("\x1b[34;1m"); color = true; }
        if (i == h1) { printf("\x1b[32;1m"); color = true; }
        if (i == h0) { printf("\x1b[31;1m"); color = true; }
      } else {
        if (i == h0) { printf("\x1b[34;1m"); color = true; }
        if (i == h1) { printf("\x1b[31;1m"); color = true; }
      }
      printf("%02X", mem[i]);
      if (color) printf("\x1b[;m");
    }
    printf("\n");
  }

  void InitByteColors(std::array<std::array<uint8_t, 3>, 256> &col) const override {
    for (int i = 0; i < 256; i++) col[i] = {(uint8_t)i, 0, (uint8_t)(255 - i)};
  }
};

}  // namespace

std::unique_ptr<Presenter> MakePresenter(const std::string &name) {
  if (name == "bff") return std::make_unique<BffPresenter>(kAscii, true);
  if (name == "bff_noheads") return std::make_unique<BffPresenter>(kAscii, false);
  if (name == "bff8") return std::make_unique<BffPresenter>(kAscii, true);
  if (name == "bff8_noheads") return std::make_unique<BffPresenter>(kAscii, false);
  if (name == "bff_noheads_4bit") return std::make_unique<BffPresenter>(k4bit, false);
  if (name == "bff_perm") return std::make_unique<BffPresenter>(kPerm, true);
  if (name == "bff_selfmove") return std::make_unique<BffPresenter>(kSelfmove, true);
  if (name == "forth") return std::make_unique<ForthPresenter>(kForth);
  if (name == "forthcopy") return std::make_unique<ForthPresenter>(kForthcopy);
  if (name == "forthtrivial") return std::make_unique<ForthPresenter>(kForthtrivial);
  if (name == "forthtrivial_reset") return std::make_unique<ForthPresenter>(kForthtrivialReset);
  if (name == "subleq") return std::make_unique<SubleqPresenter>(false);
  if (name == "rsubleq4") return std::make_unique<SubleqPresenter>(true);
  return nullptr;
}

}  // namespace cubff_host
