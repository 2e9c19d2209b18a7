// Host-side presentation of the 13 languages: parsing program text, byte
// glyphs, coloured printing and palette.  No compute here: evaluation always
// runs on the GPU.
#pragma once

#include <stdint.h>

#include <array>
#include <memory>
#include <string>
#include <vector>

namespace cubff_host {

struct Presenter {
  virtual ~Presenter() {}
  virtual std::vector<uint8_t> Parse(const std::string &program) const = 0;
  // Glyph of byte c (MapChar); `mem` must hold at least 32 bytes.
  virtual const char *MapChar(uint8_t c, char *mem) const = 0;
  virtual void PrintProgram(size_t pc_pos, const uint8_t *mem, size_t len,
                            const size_t *separators, size_t num_separators) const = 0;
  virtual void InitByteColors(std::array<std::array<uint8_t, 3>, 256> &colors) const = 0;
  // One `debug` line: the state before a step, from a step record of
  // cubff_trace_steps (u32 pc, r0, r1, r2, 128 tape bytes, 128 stack bytes).
  virtual void PrintStep(const uint8_t *record) const = 0;
};

// Little-endian u32 field q of a step record.
inline uint32_t step_field(const uint8_t *record, int q) {
  return (uint32_t)record[4 * q] | (uint32_t)record[4 * q + 1] << 8 |
         (uint32_t)record[4 * q + 2] << 16 | (uint32_t)record[4 * q + 3] << 24;
}

// Presenter of the language registered under `name`, or null.
std::unique_ptr<Presenter> MakePresenter(const std::string &name);

}  // namespace cubff_host
