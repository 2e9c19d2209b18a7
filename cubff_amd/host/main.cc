// Command-line front end with the reference's flags, checks, dashboard, CSV
// log, frame dumps and stopping rules (main.cc of
// paradigms-of-intelligence/cubff).  The flag parser is table driven: each
// flag is one row naming its type, storage and help text.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <optional>
#include <string>
#include <variant>
#include <vector>

#include "cubff.h"

namespace {

using FlagValue =
    std::variant<bool *, size_t *, double *, std::string *, std::optional<size_t> *,
                 std::optional<double> *, std::optional<std::string> *>;

struct FlagRow {
  const char *name;
  FlagValue value;
  const char *help;
};

struct Options {
  std::optional<std::string> run;
  size_t run_steps = 32 * 1024;
  bool debug = false;
  size_t num = 128 * 1024;
  std::optional<size_t> max_epochs;
  size_t seed = 0;
  double mutation_prob = 1.0 / (256 * 16);
  std::optional<double> stopping_bpb;
  std::optional<std::string> initial_program;
  std::optional<std::string> log;
  std::optional<std::string> load;
  std::optional<std::string> checkpoint_dir;
  std::optional<size_t> reset_interval;
  std::string lang;
  std::string interaction_pattern;
  bool permute_programs = true;
  bool fixed_shuffle = false;
  bool zero_init = false;
  bool eval_selfrep = false;
  size_t print_interval = 64;
  size_t save_interval = 256;
  size_t clear_interval = 2048;
  std::string draw_to;
  std::string draw_to_2d;
  size_t grid_width_2d = 0;
  bool disable_output = false;
  std::optional<size_t> stopping_selfrep_count;
};

// main.cc:173-213 of the reference: same names, defaults and help strings.
std::vector<FlagRow> flag_table(Options &o) {
  return {
      {"run", &o.run, "run a program"},
      {"run_steps", &o.run_steps, "max number of steps for running a program"},
      {"debug", &o.debug, "print execution step by step"},
      {"num", &o.num, "number of programs to evolve"},
      {"max_epochs", &o.max_epochs, "max epochs"},
      {"seed", &o.seed, "seed"},
      {"mutation_prob", &o.mutation_prob, "mutation_prob"},
      {"stopping_bpb", &o.stopping_bpb, "bits per byte below which to stop execution"},
      {"initial_program", &o.initial_program, "program to seed the soup with"},
      {"log", &o.log, "log file"},
      {"load", &o.load, "load a previous save"},
      {"checkpoint_dir", &o.checkpoint_dir, "directory to store checkpoints"},
      {"reset_interval", &o.reset_interval, "reset interval"},
      {"lang", &o.lang, "language to run"},
      {"interaction_pattern", &o.interaction_pattern,
       "file containing the allowed interactions, with two integers `a b` per line, "
       "representing that program `a` is allowed to interact with `b` (in that order)."},
      {"permute_programs", &o.permute_programs,
       "do not shuffle programs between runs (cyclic interactions)"},
      {"fixed_shuffle", &o.fixed_shuffle, "deterministic shuffling pattern"},
      {"zero_init", &o.zero_init, "zero init"},
      {"eval_selfrep", &o.eval_selfrep, "evaluate self replication in every epoch"},
      {"print_interval", &o.print_interval, "interval between prints"},
      {"save_interval", &o.save_interval, "interval between saves"},
      {"clear_interval", &o.clear_interval, "interval between clears"},
      {"draw_to", &o.draw_to, "directory to save 1d-drawn frames to (must exist)"},
      {"draw_to_2d", &o.draw_to_2d,
       "directory to save 2d-drawn frames to (must exist, and num must be a square number)"},
      {"grid_width_2d", &o.grid_width_2d, "width of the 2d grid"},
      {"disable_output", &o.disable_output, "disable printing to stdout"},
      {"stopping_selfrep_count", &o.stopping_selfrep_count,
       "stop when that many programs appear to be self-replicators"},
  };
}

struct PrintValue {
  void operator()(bool *v) const { fprintf(stderr, "%s", *v ? "true" : "false"); }
  void operator()(size_t *v) const { fprintf(stderr, "%zu", *v); }
  void operator()(double *v) const { fprintf(stderr, "%8.5f", *v); }
  void operator()(std::string *v) const { fprintf(stderr, "%s", v->c_str()); }
  template <typename T>
  void operator()(std::optional<T> *v) const {
    if (!v->has_value()) {
      fprintf(stderr, "<no value>");
    } else {
      T x = **v;
      (*this)(&x);
    }
  }
};

[[noreturn]] void usage(const char *argv0, const std::vector<FlagRow> &flags) {
  fprintf(stderr, "%s flags\n\n", argv0);
  std::vector<const FlagRow *> sorted;
  for (const auto &f : flags) sorted.push_back(&f);
  std::sort(sorted.begin(), sorted.end(),
            [](const FlagRow *a, const FlagRow *b) { return strcmp(a->name, b->name) < 0; });
  for (const FlagRow *f : sorted) {
    fprintf(stderr, "     --%s: %s; default: ", f->name, f->help);
    std::visit(PrintValue{}, f->value);
    fprintf(stderr, "\n");
  }
  exit(1);
}

void parse_flags(int argc, char **argv, Options &o) {
  const std::vector<FlagRow> flags = flag_table(o);
  // "--name" for every flag, "--noname" for booleans.
  std::map<std::string, std::pair<const FlagRow *, bool>> by_spelling;
  for (const auto &f : flags) {
    by_spelling[std::string("--") + f.name] = {&f, true};
    if (std::holds_alternative<bool *>(f.value))
      by_spelling[std::string("--no") + f.name] = {&f, false};
  }
  for (int i = 1; i < argc; i++) {
    const std::string arg = argv[i];
    if (arg == "--help" || arg == "-h") usage(argv[0], flags);
    auto it = by_spelling.find(arg);
    if (it == by_spelling.end()) {
      fprintf(stderr, "Unknown flag %s\n", argv[i]);
      usage(argv[0], flags);
    }
    const FlagRow *f = it->second.first;
    if (auto *b = std::get_if<bool *>(&f->value)) {
      **b = it->second.second;
      continue;
    }
    if (i + 1 >= argc) {
      fprintf(stderr, "missing argument for flag %s\n", argv[i]);
      exit(1);
    }
    const char *v = argv[++i];
    std::visit(
        [&](auto *p) {
          using T = std::remove_pointer_t<decltype(p)>;
          if constexpr (std::is_same_v<T, size_t> || std::is_same_v<T, std::optional<size_t>>)
            *p = (size_t)std::stoul(v);
          else if constexpr (std::is_same_v<T, double> ||
                             std::is_same_v<T, std::optional<double>>)
            *p = std::stod(v);
          else if constexpr (std::is_same_v<T, std::string> ||
                             std::is_same_v<T, std::optional<std::string>>)
            *p = std::string(v);
        },
        f->value);
  }
}

void write_ppm(const std::string &dir, size_t frame, size_t xs, size_t ys,
               const std::vector<uint8_t> &rgb) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/%012lld.ppm", dir.c_str(), (long long)frame);
  FILE *f = CheckFopen(path, "w");
  fprintf(f, "P6\n%lld %lld\n255\n", (long long)xs, (long long)ys);
  fwrite(rgb.data(), 1, rgb.size(), f);
  fclose(f);
}

}  // namespace

int main(int argc, char **argv) {
  Options o;
  parse_flags(argc, argv, o);

  SimulationParams params;
  params.num_programs = o.num;
  params.reset_interval = o.reset_interval;
  params.seed = o.seed;
  params.load_from = o.load;
  params.mutation_prob = (uint32_t)std::round(o.mutation_prob * (1 << 30));
  params.permute_programs = o.permute_programs;
  params.fixed_shuffle = o.fixed_shuffle;
  params.zero_init = o.zero_init;
  params.eval_selfrep = o.eval_selfrep;
  params.save_to = o.checkpoint_dir;
  params.save_interval = o.save_interval;
  const size_t n = params.num_programs;

  if (params.fixed_shuffle && (n & (n - 1)) != 0) {
    fprintf(stderr, "#programs must be a power of two for fixed shuffle\n");
    return 1;
  }
  if (params.fixed_shuffle && !o.interaction_pattern.empty()) {
    fprintf(stderr, "fixed shuffle and interaction pattern are incompatible\n");
    return 1;
  }

  // 1-D frames: programs in columns of 64 pixels with 4 pixels of padding,
  // about as many columns as rows (main.cc:244-257).
  constexpr size_t kPad1d = 4;
  size_t rows_1d = 0, cols_1d = 0;
  if (!o.draw_to.empty()) {
    rows_1d = n / std::ceil(std::sqrt(n / (kSingleTapeSize + kPad1d)));
    cols_1d = (n + rows_1d - 1) / rows_1d;
  }
  const size_t width_1d = cols_1d ? cols_1d * (kSingleTapeSize + kPad1d) - kPad1d : 0;
  std::vector<uint8_t> frame_1d(rows_1d * 3 * width_1d);

  // 2-D frames: each program an 8x8 block on a grid_width_2d-wide grid.
  size_t grid_w = 0;
  if (!o.draw_to_2d.empty()) {
    if (o.grid_width_2d) {
      grid_w = o.grid_width_2d;
      if (n % grid_w != 0) {
        fprintf(stderr, "grid width must divide num_programs\n");
        return 1;
      }
    } else {
      grid_w = (size_t)std::sqrt(n);
      if (grid_w * grid_w != n) {
        fprintf(stderr, "number of programs must be a square\n");
        return 1;
      }
    }
  }
  std::vector<uint8_t> frame_2d(o.draw_to_2d.empty() ? 0 : n * 3 * kSingleTapeSize);

  if (!o.interaction_pattern.empty()) {
    FILE *f = fopen(o.interaction_pattern.c_str(), "r");
    if (!f) {
      fprintf(stderr, "could not open interaction pattern file\n");
      return 1;
    }
    params.allowed_interactions.resize(n);
    long long a, b;
    while (fscanf(f, "%lld%lld", &a, &b) == 2) {
      if ((size_t)a >= n || (size_t)b >= n) {
        fprintf(stderr, "invalid interaction pattern: programs not in [0, n) range.\n");
        return 1;
      }
      params.allowed_interactions[a].push_back((uint32_t)b);
    }
    fclose(f);
  }

  const uint32_t print_interval = (uint32_t)o.print_interval;
  const uint32_t clear_interval = (uint32_t)o.clear_interval;
  // The reference keeps the bpb threshold in an optional<size_t>, so the
  // flag's value is truncated to an integer (main.cc:306).
  std::optional<size_t> stopping_bpb;
  if (o.stopping_bpb.has_value()) stopping_bpb = (size_t)*o.stopping_bpb;

  if (o.stopping_selfrep_count.has_value() && !params.eval_selfrep) {
    fprintf(stderr, "stopping_selfrep_count requires eval_selfrep\n");
    return 1;
  }
  if (params.save_interval % print_interval != 0) {
    fprintf(stderr, "save interval must be divisible by print interval\n");
    return 1;
  }
  if (clear_interval % print_interval != 0) {
    fprintf(stderr, "clear interval must be divisible by print interval\n");
    return 1;
  }
  params.callback_interval = print_interval;

  const LanguageInterface *language = GetLanguage(o.lang);
  if (o.run.has_value()) {
    printf("%s", ResetColors());
    language->RunSingleProgram(*o.run, o.run_steps, o.debug);
    return 0;
  }

  FILE *logfile = nullptr;
  if (o.log.has_value()) {
    logfile = CheckFopen(o.log->c_str(), "w");
    fprintf(logfile, params.eval_selfrep
                         ? "epoch,brotli_size,soup_size,higher_entropy,number_selfreps\n"
                         : "epoch,brotli_size,soup_size,higher_entropy\n");
  }

  auto callback = [&](const SimulationState &st) {
    int replicators = 0;
    for (size_t r : st.replication_per_prog) replicators += r >= (size_t)kSelfrepThreshold;

    if (!o.disable_output) {
      if (st.epoch % clear_interval == 1) printf("%s\033[2J\033[H", ResetColors());
      printf("%s\033[0;0H    Elapsed: %10.3f        ops: %23zu     "
             "MOps/s: %12.3f Epochs: %13zu ops/prog/epoch: %10.3f\n"
             "Brotli size: %10zu Brotli bpb: %23.4f bytes/prog: %12.4f     H0: "
             "%13.4f higher entropy: %10.6f number of repicators: %10d\n",
             ResetColors(), st.elapsed_s, st.total_ops, st.mops_s, st.epoch, st.ops_per_run,
             st.brotli_size, st.brotli_bpb, st.bytes_per_prog, st.h0, st.higher_entropy,
             params.eval_selfrep ? replicators : -1);
      for (const auto &row : {&st.frequent_bytes, &st.uncommon_bytes}) {
        for (const auto &[glyph, frac] : *row)
          printf("\033[37;1m%s%s %5.2f%% ", glyph.c_str(), ResetColors(), frac * 100.0);
        printf("\n");
      }
      printf("\n\n");
      const size_t pairs = std::min<size_t>(48, n / 2);
      for (size_t i = 0; i < pairs; i++) {
        const size_t sep[1] = {kSingleTapeSize};
        if (params.eval_selfrep)
          printf("%02d %02d ", (int)st.replication_per_prog[2 * i],
                 (int)st.replication_per_prog[2 * i + 1]);
        language->PrintProgram(2 * kSingleTapeSize, st.soup.data() + i * 2 * kSingleTapeSize,
                               2 * kSingleTapeSize, sep, 1);
      }
      fflush(stdout);
    }

    if (logfile) {
      if (params.eval_selfrep)
        fprintf(logfile, "%zu,%zu,%zu,%f,%d\n", st.epoch, st.brotli_size,
                st.soup.size() / kSingleTapeSize, st.higher_entropy, replicators);
      else
        fprintf(logfile, "%zu,%zu,%zu,%f\n", st.epoch, st.brotli_size,
                st.soup.size() / kSingleTapeSize, st.higher_entropy);
      fflush(logfile);
    }

    if (!o.draw_to.empty()) {
      for (size_t i = 0; i < n; i++) {
        const size_t x0 = (i / rows_1d) * (kSingleTapeSize + kPad1d), y = i % rows_1d;
        for (size_t j = 0; j < kSingleTapeSize; j++)
          memcpy(&frame_1d[(y * width_1d + x0 + j) * 3],
                 st.byte_colors[st.soup[i * kSingleTapeSize + j]].data(), 3);
      }
      write_ppm(o.draw_to, st.epoch, width_1d, rows_1d, frame_1d);
    }

    if (!o.draw_to_2d.empty()) {
      const size_t xs = grid_w * 8;
      for (size_t i = 0; i < n; i++) {
        const size_t gx = i % grid_w, gy = i / grid_w;
        for (size_t j = 0; j < kSingleTapeSize; j++) {
          const size_t px = j % 8, py = j / 8;
          uint8_t *dst = &frame_2d[((gy * 8 + py) * xs + gx * 8 + px) * 3];
          memcpy(dst, st.byte_colors[st.soup[i * kSingleTapeSize + j]].data(), 3);
          if (px == 0 || py == 0)  // darker block borders
            for (int c = 0; c < 3; c++) dst[c] = (uint8_t)std::max(dst[c] - 32, 0);
        }
      }
      write_ppm(o.draw_to_2d, st.epoch, xs, n / grid_w * 8, frame_2d);
    }

    if (o.max_epochs.has_value() && st.epoch > *o.max_epochs) return true;
    if (stopping_bpb.has_value() && st.brotli_bpb < *stopping_bpb) return true;
    if (o.stopping_selfrep_count.has_value() &&
        (size_t)replicators >= *o.stopping_selfrep_count)
      return true;
    return false;
  };

  language->RunSimulation(params, o.initial_program, callback);
  if (logfile) fclose(logfile);
  return 0;
}
