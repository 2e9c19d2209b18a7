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
};

// Presenter of the language registered under `name`, or null.
std::unique_ptr<Presenter> MakePresenter(const std::string &name);

}  // namespace cubff_host
