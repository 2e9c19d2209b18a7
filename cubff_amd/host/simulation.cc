// LanguageInterface implementations over the HIP C-ABI, and the registry.
//
// RunSimulation keeps the reference's host loop contract
// (common_language.h:344-597): the same pairing modes, epoch numbering,
// callback cadence and SimulationState fields, brotli entropy metric,
// checkpoint format and reset rule.  The per-epoch work -- pairing,
// mutation, interpretation, write back, self-rep -- runs on the GPU; the soup
// only comes back to the host at callback epochs.
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <thread>

#include "../../include/cubff_hip.h"
#include "brotli/encode.h"
#include "cubff.h"
#include "present.h"

using cubff_host::MakePresenter;
using cubff_host::Presenter;

FILE *CheckFopen(const char *f, const char *mode) {
  FILE *out = fopen(f, mode);
  if (out == nullptr) {
    char buf[4096];
    fprintf(stderr, "Could not open %s: %s\n", f, strerror_r(errno, buf, sizeof(buf)));
    exit(1);
  }
  return out;
}

namespace {

// Failures of the HIP library end the process, as CUCHECK does
// (common_language.h:35-42).
void hip_check(int rc, const char *what) {
  if (rc != CUBFF_OK) {
    printf("%s failed with error: %s\n", what, cubff_last_error());
    exit(1);
  }
}

// Steps --debug prints at most (272 bytes of record each on the host).
constexpr uint64_t kMaxDebugSteps = 1ull << 20;

#define CHECK(op)                   \
  if (!(op)) {                      \
    printf("%s is false\n", #op);   \
    exit(1);                        \
  }

class HipLanguage : public LanguageInterface {
 public:
  HipLanguage(std::string name, int lang, std::unique_ptr<Presenter> p)
      : name_(std::move(name)), lang_(lang), p_(std::move(p)) {}

  std::vector<uint8_t> Parse(const std::string &program) override {
    return p_->Parse(program);
  }

  void PrintProgram(size_t pc_pos, const uint8_t *mem, size_t len, const size_t *separators,
                    size_t num_separators) const override {
    p_->PrintProgram(pc_pos, mem, len, separators, num_separators);
  }

  void RunSingleProgram(std::string program, size_t stepcount, bool debug) const override {
    RunSingleParsedProgram(p_->Parse(program), stepcount, debug);
  }

  // common_language.h:279-296: zero-padded 128-byte tape, printed before and
  // after one Evaluate of up to stepcount steps on the GPU.
  void RunSingleParsedProgram(const std::vector<uint8_t> &parsed, size_t stepcount,
                              bool debug) const override {
    uint8_t tape[2 * kSingleTapeSize] = {};
    memcpy(tape, parsed.data(), std::min<size_t>(parsed.size(), sizeof(tape)));
    p_->PrintProgram(2 * kSingleTapeSize, tape, 2 * kSingleTapeSize, nullptr, 0);
    uint64_t ops = 0;
    if (debug) {
      // Every step's state, recorded on the GPU, printed in order
      // (bff.inc.h:368-372 and the other Evaluate loops).
      const uint64_t cap = std::min<uint64_t>(stepcount, kMaxDebugSteps);
      std::vector<uint8_t> rec(cap * CUBFF_STEP_RECORD_BYTES);
      uint64_t nrec = 0, run = 0;
      hip_check(cubff_trace_steps(lang_, 0, tape, stepcount, rec.data(), cap, &nrec, &ops, &run),
                "cubff_trace_steps");
      for (uint64_t i = 0; i < nrec; i++) p_->PrintStep(rec.data() + i * CUBFF_STEP_RECORD_BYTES);
      if (run > nrec)
        fprintf(stderr, "note: --debug printed the first %llu of %llu steps\n",
                (unsigned long long)nrec, (unsigned long long)run);
    } else {
      uint32_t ops32 = 0, steps = 0;
      hip_check(cubff_eval_tapes(lang_, 0, tape, 1, stepcount, &ops32, &steps),
                "cubff_eval_tapes");
      ops = ops32;
    }
    printf("%s", ResetColors());
    printf("ops: %d\n", (int)ops);
    printf("\n");
    p_->PrintProgram(2 * kSingleTapeSize, tape, 2 * kSingleTapeSize, nullptr, 0);
  }

  size_t EvalSelfrep(std::string program, size_t epoch, size_t seed, bool debug) override {
    std::vector<uint8_t> parsed = p_->Parse(program);
    return EvalParsedSelfrep(parsed, epoch, seed, debug);
  }

  // common_language.h:326-342.
  size_t EvalParsedSelfrep(std::vector<uint8_t> &parsed, size_t epoch, size_t seed,
                           bool debug) override {
    uint64_t res = 0;
    if (!debug) {
      hip_check(cubff_eval_selfrep(lang_, 0, parsed.data(), parsed.size(), epoch, seed, &res),
                "cubff_eval_selfrep");
      return res;
    }
    // common_language.h:225-255: each trial's tapes before and after each
    // generation, recorded on the GPU and printed trial by trial.
    std::vector<uint8_t> tapes(CUBFF_SELFREP_TRACE_BYTES);
    hip_check(cubff_eval_selfrep_trace(lang_, 0, parsed.data(), parsed.size(), epoch, seed, &res,
                                       tapes.data()),
              "cubff_eval_selfrep_trace");
    const size_t seps[1] = {kSingleTapeSize};
    for (size_t i = 0; i < 13; i++) {
      for (size_t g = 0; g < 5; g++) {
        const uint8_t *t = tapes.data() + (i * 5 + g) * 2 * 2 * kSingleTapeSize;
        for (int after = 0; after < 2; after++) {
          if (g == 0)
            printf("Iteration %lu: %s first step\n", (unsigned long)i, after ? "after" : "before");
          else
            printf("Iteration %lu: %s step %lu\n", (unsigned long)i, after ? "after" : "before",
                   (unsigned long)(g + 1));
          p_->PrintProgram(2 * kSingleTapeSize, t + after * 2 * kSingleTapeSize,
                           2 * kSingleTapeSize, seps, 1);
        }
      }
    }
    return res;
  }

  void RunSimulation(const SimulationParams &params, std::optional<std::string> initial_program,
                     std::function<bool(const SimulationState &)> callback) const override;

 private:
  std::string name_;
  int lang_;
  std::unique_ptr<Presenter> p_;
};

void HipLanguage::RunSimulation(const SimulationParams &params,
                                std::optional<std::string> initial_program,
                                std::function<bool(const SimulationState &)> callback) const {
  size_t num_programs = params.num_programs;
  size_t reset_index = 1;
  size_t epoch = 0;

  FILE *load_file = nullptr;
  if (params.load_from.has_value()) {  // header: common_language.h:354-360
    load_file = CheckFopen(params.load_from->c_str(), "r");
    CHECK(fread(&reset_index, sizeof(reset_index), 1, load_file) == 1);
    CHECK(fread(&num_programs, sizeof(num_programs), 1, load_file) == 1);
    CHECK(fread(&epoch, sizeof(epoch), 1, load_file) == 1);
  }
  CHECK(num_programs % 2 == 0);

  cubff_ctx *ctx = nullptr;
  hip_check(cubff_ctx_create(&ctx, lang_, num_programs, 0), "cubff_ctx_create");
  auto seed = [&](uint64_t x) { return cubff_epoch_seed(params.seed, x); };

  hip_check(cubff_init_programs(ctx, seed(0), params.zero_init, num_programs),
            "cubff_init_programs");
  if (initial_program.has_value()) {
    std::vector<uint8_t> parsed = p_->Parse(*initial_program);
    const size_t len = std::min(parsed.size(), num_programs * kSingleTapeSize);
    hip_check(cubff_write_soup(ctx, 0, parsed.data(), len), "cubff_write_soup");
  }
  unsigned long long total_ops = 0;

  SimulationState state;
  state.soup.reserve(num_programs * kSingleTapeSize + 16);
  state.soup.resize(num_programs * kSingleTapeSize);
  state.replication_per_prog.resize(num_programs);
  state.shuffle_idx.resize(num_programs);
  p_->InitByteColors(state.byte_colors);

  if (params.save_to.has_value()) {
    CHECK(mkdir(params.save_to->c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH) != -1 ||
          errno == EEXIST);
  }
  if (load_file) {  // common_language.h:398-403
    CHECK(fread(state.soup.data(), 1, num_programs * kSingleTapeSize, load_file) ==
          num_programs * kSingleTapeSize);
    fclose(load_file);
    hip_check(cubff_write_soup(ctx, 0, state.soup.data(), state.soup.size()),
              "cubff_write_soup");
  }

  // Pairing mode (common_language.h:435-486).
  int mode = CUBFF_PAIR_SHUFFLE;
  if (!params.allowed_interactions.empty()) {
    mode = CUBFF_PAIR_2D;
    std::vector<uint64_t> row_off(num_programs + 1, 0);
    std::vector<uint32_t> nbr;
    for (size_t i = 0; i < num_programs; i++) {
      if (i < params.allowed_interactions.size())
        for (uint32_t v : params.allowed_interactions[i]) nbr.push_back(v);
      row_off[i + 1] = nbr.size();
    }
    if (nbr.empty()) nbr.push_back(0);
    hip_check(cubff_set_interactions(ctx, row_off.data(), nbr.data()),
              "cubff_set_interactions");
  } else if (params.permute_programs) {
    mode = params.fixed_shuffle ? CUBFF_PAIR_FIXED : CUBFF_PAIR_SHUFFLE;
  } else {
    mode = CUBFF_PAIR_CYCLIC;
  }

  std::vector<uint8_t> brotlified(BrotliEncoderMaxCompressedSize(num_programs * kSingleTapeSize));
  std::vector<uint64_t> selfrep(num_programs);
  // CheckSelfRep and the reset re-init are launched for floor(N / 32) * 32
  // programs (common_language.h:565, :592).
  const size_t launched = num_programs / 32 * 32;

  // Epochs run in batches up to the next callback or reset epoch
  // (cubff_run_epochs picks the fastest multi-epoch path).  At a callback
  // epoch the soup, pairing, counters, byte histogram (on the GPU) and
  // self-rep scores are taken first; then the next batch of epochs starts
  // on a worker thread while this thread compresses the soup, fills the
  // state and runs the callback, so brotli overlaps the GPU.  The reset
  // re-init is queued before that batch, as the reference applies it after
  // the callback and before the next epoch.  If the callback stops the run,
  // the batch already under way is finished and dropped.
  CHECK(params.callback_interval > 0);
  const size_t interval = params.callback_interval;
  const bool resets = params.reset_interval.has_value() && *params.reset_interval > 0;
  auto next_event = [&](size_t e) {
    size_t n = (e + interval - 1) / interval * interval;
    if (resets) {
      const size_t r = *params.reset_interval;
      n = std::min(n, (e + r - 1) / r * r);
    }
    return n;
  };
  auto run_through = [&](size_t from, size_t to) {  // epochs from..to
    hip_check(cubff_run_epochs(ctx, mode, params.seed, from, to - from + 1, params.mutation_prob),
              "cubff_run_epochs");
  };

  auto start = std::chrono::high_resolution_clock::now();
  const auto simulation_start = start;
  // The first batch starts at epoch 0, or at the epoch a checkpoint stored.
  const size_t first_epoch = epoch;
  epoch = next_event(first_epoch);
  run_through(first_epoch, epoch);
  for (;;) {
    const bool at_callback = epoch % interval == 0;
    const bool at_reset = resets && epoch % *params.reset_interval == 0;
    uint64_t insn = 0, num_runs = 0, nidx = 0;
    uint64_t counts[256] = {};
    const size_t reset_index_here = reset_index;
    std::chrono::high_resolution_clock::time_point stop;
    if (at_callback) {  // common_language.h:496-588
      stop = std::chrono::high_resolution_clock::now();
      hip_check(cubff_read_counters(ctx, &insn, &num_runs, nullptr), "cubff_read_counters");
      total_ops += insn;
      hip_check(cubff_read_soup(ctx, 0, state.soup.data(), state.soup.size()), "cubff_read_soup");
      hip_check(cubff_read_shuffle(ctx, state.shuffle_idx.data(), &nidx), "cubff_read_shuffle");
      hip_check(cubff_byte_histogram(ctx, counts), "cubff_byte_histogram");
      if (params.eval_selfrep) {
        hip_check(cubff_selfrep(ctx, seed(epoch), launched, selfrep.data()), "cubff_selfrep");
        for (size_t i = 0; i < launched; i++) state.replication_per_prog[i] = selfrep[i];
      }
      hip_check(cubff_reset_counters(ctx), "cubff_reset_counters");
    }
    if (at_reset) {  // common_language.h:590-595
      hip_check(cubff_init_programs(ctx, seed(reset_index), params.zero_init, launched),
                "cubff_init_programs");
      reset_index++;
    }
    const size_t next = next_event(epoch + 1);
    if (!at_callback) {
      run_through(epoch + 1, next);
      epoch = next;
      continue;
    }
    const auto batch_start = std::chrono::high_resolution_clock::now();
    std::thread batch([&run_through, from = epoch + 1, to = next] { run_through(from, to); });

    size_t brotli_size = brotlified.size();
    BrotliEncoderCompress(2, 24, BROTLI_MODE_GENERIC, state.soup.size(), state.soup.data(),
                          &brotli_size, brotlified.data());
    const float elapsed_s =
        std::chrono::duration_cast<std::chrono::microseconds>(stop - start).count() * 1e-6;
    const float mops_s = insn * 1.0 / elapsed_s * 1e-6;
    const float sim_elapsed_s =
        std::chrono::duration_cast<std::chrono::microseconds>(stop - simulation_start).count() *
        1e-6;
    std::vector<uint8_t> sorted(256);
    double h0 = 0;
    for (size_t i = 0; i < 256; i++) {
      sorted[i] = (uint8_t)i;
      const double frac = counts[i] * 1.0 / state.soup.size();
      h0 -= counts[i] ? frac * std::log2(frac) : 0.0;
    }
    std::sort(sorted.begin(), sorted.end(), [&](uint8_t a, uint8_t b) {
      return std::make_pair(counts[b], b) < std::make_pair(counts[a], a);
    });
    const double brotli_bpb = brotli_size * 8.0 / (num_programs * kSingleTapeSize);

    state.elapsed_s = sim_elapsed_s;
    state.total_ops = total_ops;
    state.mops_s = mops_s;
    state.epoch = epoch + 1;
    state.ops_per_run = insn * 1.0 / num_runs;
    state.brotli_size = brotli_size;
    state.brotli_bpb = brotli_bpb;
    state.bytes_per_prog = brotli_size * 1.0 / num_programs;
    state.h0 = h0;
    state.higher_entropy = h0 - brotli_bpb;
    for (size_t i = 0; i < state.frequent_bytes.size(); i++) {
      const uint8_t c = sorted[i];
      char chmem[32];
      state.frequent_bytes[i].first = p_->MapChar(c, chmem);
      state.frequent_bytes[i].second = counts[c] * 1.0 / state.soup.size();
    }
    for (size_t i = 0; i < state.uncommon_bytes.size(); i++) {
      const uint8_t c = sorted[256 - state.uncommon_bytes.size() + i];
      char chmem[32];
      state.uncommon_bytes[i].first = p_->MapChar(c, chmem);
      state.uncommon_bytes[i].second = counts[c] * 1.0 / state.soup.size();
    }
    if (params.save_to.has_value() && (epoch % params.save_interval == 0)) {
      std::vector<char> path(params.save_to->size() + 20);
      snprintf(path.data(), path.size(), "%s/%010zu.dat", params.save_to->c_str(), epoch);
      FILE *f = CheckFopen(path.data(), "w");
      const size_t epoch_to_save = epoch + 1;
      fwrite(&reset_index_here, sizeof(reset_index_here), 1, f);
      fwrite(&num_programs, sizeof(num_programs), 1, f);
      fwrite(&epoch_to_save, sizeof(epoch_to_save), 1, f);
      fwrite(state.soup.data(), 1, state.soup.size(), f);
      fclose(f);
    }
    const bool done = callback(state);
    batch.join();
    if (done) break;
    start = batch_start;
    epoch = next;
  }
  cubff_ctx_destroy(ctx);
}

std::map<std::string, std::unique_ptr<LanguageInterface>> &registry() {
  static std::map<std::string, std::unique_ptr<LanguageInterface>> r = [] {
    std::map<std::string, std::unique_ptr<LanguageInterface>> m;
    for (int i = 0; i < cubff_num_langs(); i++) {
      const std::string name = cubff_lang_name(i);
      m[name] = std::make_unique<HipLanguage>(name, i, MakePresenter(name));
    }
    return m;
  }();
  return r;
}

}  // namespace

void RegisterLanguage(const char *lang, std::unique_ptr<LanguageInterface> interface) {
  registry()[lang] = std::move(interface);
}

const LanguageInterface *GetLanguage(const std::string &language) {
  auto &map = registry();
  auto it = map.find(language);
  if (it == map.end()) {
    fprintf(stderr, "Unknown language `%s`\nAvailable languages:\n", language.c_str());
    for (const auto &kv : map) fprintf(stderr, "%s\n", kv.first.c_str());
    exit(1);
  }
  return it->second.get();
}

std::vector<std::string> LanguageNames() {
  std::vector<std::string> out;
  for (const auto &kv : registry()) out.push_back(kv.first);
  return out;
}
