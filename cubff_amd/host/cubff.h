// Host API of cubff on MI355X: the reference's plugin interface
// (common.h:55-139 of paradigms-of-intelligence/cubff) with the same type
// names, fields and virtual methods, so callers written against the
// reference compile and behave the same.  Every language is a
// LanguageInterface found by name; RunSimulation drives the HIP hot path
// through the C-ABI in include/cubff_hip.h.
#pragma once

#include <stdint.h>
#include <stdio.h>

#include <array>
#include <functional>
#include <memory>
#include <optional>
#include <string>
#include <utility>
#include <vector>

inline constexpr const char *ResetColors() { return "\x1b[0;38:5:35;48:5:232m"; }

constexpr int kSingleTapeSize = 64;   // common.h:55
constexpr int kSelfrepThreshold = 5;  // common.h:56

// common.h:58-72.
struct SimulationParams {
  size_t num_programs = 128 * 1024;
  size_t seed = 0;
  uint32_t mutation_prob = 1 << 18;  // out of 2^30
  std::optional<size_t> reset_interval = std::nullopt;
  std::optional<std::string> load_from = std::nullopt;
  std::optional<std::string> save_to = std::nullopt;
  size_t callback_interval = 128;
  size_t save_interval = 0;
  bool permute_programs = true;
  bool fixed_shuffle = false;
  bool zero_init = false;
  bool eval_selfrep = false;
  std::vector<std::vector<uint32_t>> allowed_interactions;
};

// common.h:74-91.
struct SimulationState {
  std::vector<uint8_t> soup;
  std::vector<uint32_t> shuffle_idx;
  std::array<std::array<uint8_t, 3>, 256> byte_colors;
  float elapsed_s;
  size_t total_ops;
  float mops_s;
  size_t epoch;
  float ops_per_run;
  size_t brotli_size;
  float brotli_bpb;
  float bytes_per_prog;
  float h0;
  float higher_entropy;
  std::array<std::pair<std::string, float>, 16> frequent_bytes;
  std::array<std::pair<std::string, float>, 16> uncommon_bytes;
  std::vector<size_t> replication_per_prog;
};

// common.h:93-111.
struct LanguageInterface {
  virtual void RunSingleProgram(std::string program, size_t stepcount,
                                bool debug) const = 0;
  virtual void RunSingleParsedProgram(const std::vector<uint8_t> &parsed,
                                      size_t stepcount, bool debug) const = 0;
  virtual void PrintProgram(size_t pc_pos, const uint8_t *mem, size_t len,
                            const size_t *separators,
                            size_t num_separators) const = 0;
  virtual void RunSimulation(
      const SimulationParams &params, std::optional<std::string> initial_program,
      std::function<bool(const SimulationState &)> callback) const = 0;
  virtual ~LanguageInterface() {}
  virtual size_t EvalSelfrep(std::string program, size_t epoch, size_t seed,
                             bool debug) = 0;
  virtual size_t EvalParsedSelfrep(std::vector<uint8_t> &parsed, size_t epoch,
                                   size_t seed, bool debug) = 0;
  virtual std::vector<uint8_t> Parse(const std::string &program) = 0;
};

// Language registry (common.cc:22-45, common.h:133-134).  It starts with
// the 13 device languages of the HIP library; RegisterLanguage adds another
// LanguageInterface under `lang`, or replaces the one registered there, and
// the registry owns it.  It may be called from a static constructor of
// another shared object (the reference's REGISTER does so, common.h:41-45),
// before or after the built-in languages are set up.  GetLanguage returns a
// non-owning pointer; unknown names print the available languages and
// exit(1).
void RegisterLanguage(const char *lang, std::unique_ptr<LanguageInterface> interface);
const LanguageInterface *GetLanguage(const std::string &language);
std::vector<std::string> LanguageNames();

FILE *CheckFopen(const char *f, const char *mode);  // common.h:141-150
