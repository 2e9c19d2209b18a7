// Test-infrastructure harness around the compiled reference (not product
// code): drives the reference's own RunSimulation (common_language.h:344-597)
// through its public LanguageInterface and hands every callback's
// SimulationState to a C callback.  Built only into oracle/_ref/.
#include <stdio.h>
#include <cstdint>
#include <optional>
#include <string>
#include <vector>

#include "common.h"

extern "C" {

struct RefState {
  const uint8_t *soup;
  unsigned long long soup_size;
  const uint32_t *shuffle_idx;
  const unsigned long long *replication_per_prog;
  unsigned long long epoch;
  unsigned long long total_ops;
  float ops_per_run;
  unsigned long long brotli_size;
  float brotli_bpb;
  float h0;
  float higher_entropy;
  // The reference's own rate over the epochs since the previous callback,
  // callback time excluded (common_language.h:508-512, :539).
  float mops_s;
};

typedef int (*RefCallback)(const RefState *state, void *user);

// pattern: CSR neighbour lists (row_off has num_programs+1 entries) or null.
int ref_run_simulation(const char *lang, unsigned long long num_programs,
                       unsigned long long seed, uint32_t mutation_prob,
                       unsigned long long callback_interval,
                       int permute_programs, int fixed_shuffle, int zero_init,
                       int eval_selfrep, long long reset_interval,
                       const char *load_from, const char *save_to,
                       unsigned long long save_interval,
                       const char *initial_program,
                       const unsigned long long *row_off, const uint32_t *nbr,
                       RefCallback cb, void *user) {
  SimulationParams params;
  params.num_programs = num_programs;
  params.seed = seed;
  params.mutation_prob = mutation_prob;
  params.callback_interval = callback_interval;
  params.permute_programs = permute_programs != 0;
  params.fixed_shuffle = fixed_shuffle != 0;
  params.zero_init = zero_init != 0;
  params.eval_selfrep = eval_selfrep != 0;
  if (reset_interval > 0) params.reset_interval = (size_t)reset_interval;
  if (load_from) params.load_from = std::string(load_from);
  if (save_to) params.save_to = std::string(save_to);
  params.save_interval = save_interval;
  if (row_off) {
    params.allowed_interactions.resize(num_programs);
    for (size_t i = 0; i < num_programs; i++)
      for (size_t e = row_off[i]; e < row_off[i + 1]; e++)
        params.allowed_interactions[i].push_back(nbr[e]);
  }
  std::optional<std::string> init;
  if (initial_program) init = std::string(initial_program);
  const LanguageInterface *l = GetLanguage(lang);
  l->RunSimulation(params, init, [&](const SimulationState &st) {
    RefState rs;
    rs.soup = st.soup.data();
    rs.soup_size = st.soup.size();
    rs.shuffle_idx = st.shuffle_idx.data();
    rs.replication_per_prog =
        reinterpret_cast<const unsigned long long *>(
            st.replication_per_prog.data());
    rs.epoch = st.epoch;
    rs.total_ops = st.total_ops;
    rs.ops_per_run = st.ops_per_run;
    rs.brotli_size = st.brotli_size;
    rs.brotli_bpb = st.brotli_bpb;
    rs.h0 = st.h0;
    rs.higher_entropy = st.higher_entropy;
    rs.mops_s = st.mops_s;
    return cb(&rs, user) != 0;
  });
  return 0;
}

// LanguageInterface::EvalParsedSelfrep (common_language.h:326-342).
unsigned long long ref_eval_parsed_selfrep(const char *lang,
                                           const uint8_t *prog,
                                           unsigned long long len,
                                           unsigned long long epoch,
                                           unsigned long long seed) {
  std::vector<uint8_t> parsed(prog, prog + len);
  LanguageInterface *l = const_cast<LanguageInterface *>(GetLanguage(lang));
  return l->EvalParsedSelfrep(parsed, epoch, seed, false);
}

// The same with debug on: CheckSelfRep prints every trial's tapes to stdout.
unsigned long long ref_eval_parsed_selfrep_debug(const char *lang,
                                                 const uint8_t *prog,
                                                 unsigned long long len,
                                                 unsigned long long epoch,
                                                 unsigned long long seed) {
  std::vector<uint8_t> parsed(prog, prog + len);
  LanguageInterface *l = const_cast<LanguageInterface *>(GetLanguage(lang));
  unsigned long long r = l->EvalParsedSelfrep(parsed, epoch, seed, true);
  fflush(stdout);
  return r;
}

// LanguageInterface::Parse; returns the parsed length (<= cap bytes copied).
unsigned long long ref_parse(const char *lang, const char *program,
                             uint8_t *out, unsigned long long cap) {
  LanguageInterface *l = const_cast<LanguageInterface *>(GetLanguage(lang));
  std::vector<uint8_t> p = l->Parse(program);
  for (size_t i = 0; i < p.size() && i < cap; i++) out[i] = p[i];
  return p.size();
}
}
