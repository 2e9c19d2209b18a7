// Test-infrastructure harness around the compiled reference (not product
// code).  This translation unit is compiled once per reference language file:
// it includes that .cu file (REF_LANG_FILE) so the language struct, which the
// reference keeps in an anonymous namespace, is reachable, and exports plain C
// entry points named after REF_LANG_TAG.  Only oracle/_ref/ receives the build.
#include REF_LANG_FILE

#define REF_CAT2(a, b) a##b
#define REF_CAT(a, b) REF_CAT2(a, b)

// Language::Evaluate (e.g. bff.inc.h:356-389) on one 128-byte tape.
extern "C" unsigned long long REF_CAT(ref_eval_, REF_LANG_TAG)(
    uint8_t *tape, unsigned long long steps) {
  return REF_LANG_STRUCT::Evaluate(tape, steps, false);
}

// CheckSelfRep (common_language.h:203-277) for programs [0, count), serially.
extern "C" void REF_CAT(ref_selfrep_, REF_LANG_TAG)(
    uint8_t *programs, unsigned long long seed, unsigned long long num_programs,
    unsigned long long count, unsigned long long *result) {
  for (size_t i = 0; i < count; i++) {
    IndexThreadLocal() = i;
    CheckSelfRep<REF_LANG_STRUCT>(programs, seed, num_programs,
                                  reinterpret_cast<size_t *>(result), false);
  }
}

// MutateAndRunPrograms (common_language.h:157-193) for every slot, serially.
extern "C" unsigned long long REF_CAT(ref_mutate_run_, REF_LANG_TAG)(
    uint8_t *programs, const uint32_t *shuf, unsigned long long seed,
    uint32_t mutation_prob, unsigned long long num_programs,
    unsigned long long num_indices) {
  unsigned long long count = 0;
  for (size_t i = 0; i < num_programs / 2; i++) {
    IndexThreadLocal() = i;
    MutateAndRunPrograms<REF_LANG_STRUCT>(programs, shuf, seed, mutation_prob,
                                          &count, num_programs, num_indices);
  }
  return count;
}
