/*
 * cubff CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference simulator's per-epoch algorithm
 * (pairing, mutation, the 13 language interpreters, the self-replicator
 * check).  It exists so that tests, the smoke check and the CPU-baseline leg
 * of bench.py can check the HIP product path against an independent CPU
 * implementation.  Nothing in the product (cubff_amd/, include/) may link,
 * import or call it.
 *
 * Every function cites the reference file:line it restates.  Each one is
 * pinned by golden vectors recorded from the compiled reference in
 * oracle/_ref, closed forms or invariants (tests/test_cpu_oracle.py).
 */
#ifndef CUBFF_ORACLE_H_
#define CUBFF_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_BFF = 0,
  ORC_BFF_NOHEADS,
  ORC_BFF8,
  ORC_BFF8_NOHEADS,
  ORC_BFF_NOHEADS_4BIT,
  ORC_BFF_PERM,
  ORC_BFF_SELFMOVE,
  ORC_FORTH,
  ORC_FORTHCOPY,
  ORC_FORTHTRIVIAL,
  ORC_FORTHTRIVIAL_RESET,
  ORC_SUBLEQ,
  ORC_RSUBLEQ4,
  ORC_NUM_LANGS
};

int orc_num_langs(void);
const char *orc_lang_name(int lang);
int orc_lang_id(const char *name);

uint64_t orc_splitmix64(uint64_t x);
uint64_t orc_epoch_seed(uint64_t params_seed, uint64_t x);

/* Runs one interaction in place; returns ops (steps minus no-op steps).
 * If steps_out is non-null, the loop-iteration count (steps incl. no-ops) is
 * stored there. */
uint64_t orc_trace_bff(int lang, uint8_t *tape, uint64_t stepcount, uint8_t *out,
                       uint64_t cap, uint64_t *states, uint64_t *steps_out);
uint64_t orc_evaluate(int lang, uint8_t *tape, uint64_t stepcount,
                      uint64_t *steps_out);

void orc_init_programs(uint8_t *soup, uint64_t num_programs,
                       uint64_t seed_value, int zero_init);

void orc_pairing_shuffle(uint32_t *s, uint64_t num_programs,
                         uint64_t params_seed, uint64_t epoch);
void orc_pairing_fixed(uint32_t *s, uint64_t num_programs, uint64_t epoch);
void orc_pairing_cyclic(uint32_t *s, uint64_t num_programs, uint64_t epoch);
/* Neighbour lists in CSR form: programs i's neighbours are
 * nbr[row_off[i] .. row_off[i+1]) in file order.  Returns num_indices. */
uint64_t orc_pairing_2d(uint32_t *s, uint64_t num_programs,
                        const uint64_t *row_off, const uint32_t *nbr,
                        uint64_t params_seed, uint64_t epoch);

/* One epoch of mutate + evaluate + write back.  Returns the summed ops.
 * pair_ops / pair_steps (nullable) receive per-slot values. */
uint64_t orc_mutate_and_run(int lang, uint8_t *soup, const uint32_t *s,
                            uint64_t num_programs, uint64_t num_indices,
                            uint64_t seed_value, uint32_t mutation_prob,
                            uint32_t *pair_ops, uint32_t *pair_steps,
                            int nthreads);

/* Programs in two or more different slots of an epoch's pairing (0 for a
 * permutation); debug builds (-DORC_DEBUG_CHECKS) check every epoch of
 * orc_mutate_and_run with it and count what they find. */
uint64_t orc_pairing_repeats(const uint32_t *s, uint64_t num_programs);
uint64_t orc_debug_repeats_seen(void);

/* Self-replication score for programs [0, count). */
void orc_check_selfrep(int lang, const uint8_t *soup, uint64_t num_programs,
                       uint64_t count, uint64_t seed_value, uint64_t *result,
                       int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* CUBFF_ORACLE_H_ */
