// A LanguageInterface that registers itself from a shared object of its
// own, the way each language file of the reference registers at load time
// (REGISTER, common.h:41-45, calling RegisterLanguage, common.cc:42-45).
// Test infrastructure: every method reports what it was called with, so a
// caller can check that GetLanguage hands back this object.
#include <stdio.h>

#include <memory>

#include "cubff.h"

namespace {

struct EchoLanguage : LanguageInterface {
  void RunSingleProgram(std::string program, size_t stepcount, bool debug) const override {
    printf("echo RunSingleProgram %s %zu %d\n", program.c_str(), stepcount, (int)debug);
    fflush(stdout);
  }
  void RunSingleParsedProgram(const std::vector<uint8_t> &parsed, size_t stepcount,
                              bool debug) const override {
    printf("echo RunSingleParsedProgram %zu %zu %d\n", parsed.size(), stepcount, (int)debug);
    fflush(stdout);
  }
  void PrintProgram(size_t pc_pos, const uint8_t *mem, size_t len, const size_t *separators,
                    size_t num_separators) const override {
    printf("echo PrintProgram %zu %zu %zu\n", pc_pos, len, num_separators);
    fflush(stdout);
  }
  void RunSimulation(const SimulationParams &params, std::optional<std::string> initial_program,
                     std::function<bool(const SimulationState &)> callback) const override {
    SimulationState state{};
    state.soup.assign(params.num_programs, 7);
    for (size_t e = 0;; e++) {
      state.epoch = e;
      state.total_ops = params.seed + e;
      if (callback(state)) break;
    }
  }
  size_t EvalSelfrep(std::string program, size_t epoch, size_t seed, bool debug) override {
    return program.size() * 1000 + epoch * 10 + seed;
  }
  size_t EvalParsedSelfrep(std::vector<uint8_t> &parsed, size_t epoch, size_t seed,
                           bool debug) override {
    return parsed.size() * 1000 + epoch * 10 + seed + 1;
  }
  std::vector<uint8_t> Parse(const std::string &program) override {
    return std::vector<uint8_t>(program.rbegin(), program.rend());
  }
};

__attribute__((constructor)) void register_echo() {
  RegisterLanguage("echo", std::make_unique<EchoLanguage>());
}

}  // namespace
