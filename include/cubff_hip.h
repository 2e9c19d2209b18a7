/*
 * cubff on MI355X -- the C-ABI between the host simulation loop and the HIP
 * kernels of the per-epoch hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  Every
 * function returns CUBFF_OK (0) or a negative CUBFF_E* code; the message of
 * the last failure on the calling thread is returned by cubff_last_error().
 * Host buffers belong to the caller.  A context is used by one thread at a
 * time.  Reads are synchronous; launches are asynchronous on the context's
 * stream until a read or cubff_sync().
 *
 * Each entry point names the reference interface it replaces
 * (paradigms-of-intelligence/cubff, file:line).  The reference host loop
 * (common_language.h:344-597) calls CUDA directly; here the same loop calls
 * these functions.
 */
#ifndef CUBFF_HIP_H_
#define CUBFF_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUBFF_OK 0
#define CUBFF_EINVAL -1   /* bad argument */
#define CUBFF_EHIP -2     /* HIP runtime error */
#define CUBFF_ENOMEM -3   /* device allocation failed */
#define CUBFF_ESTATE -4   /* call not valid in the context's current state */

/* Pairing modes of RunSimulation (common_language.h:435-486). */
#define CUBFF_PAIR_SHUFFLE 0  /* Fisher-Yates, :418-424 + :476 */
#define CUBFF_PAIR_FIXED 1    /* --fixed_shuffle, :468-474 */
#define CUBFF_PAIR_CYCLIC 2   /* --nopermute_programs, :478-486 */
#define CUBFF_PAIR_2D 3       /* allowed_interactions greedy, :435-463 */
#define CUBFF_PAIR_HOST 4     /* caller-supplied via cubff_set_shuffle */

typedef struct cubff_ctx cubff_ctx;

/* Message of the last failed call on this thread ("" if none).
 * Replaces CUCHECK / CHECK printing (common_language.h:35-42, :125-129). */
const char *cubff_last_error(void);

/* Language registry by name.  Replaces GetLanguage's lookup over the REGISTER
 * table (common.cc:28-40, common.h:41-45); -1 for an unknown name. */
int cubff_lang_id(const char *name);
const char *cubff_lang_name(int lang);
int cubff_num_langs(void);

/* Context holding one soup of num_programs x 64 bytes on `device`, plus its
 * shuffle buffer, counters and per-slot outputs.  Replaces the DeviceMemory
 * buffers of RunSimulation (common_language.h:362-363, :405). */
int cubff_ctx_create(cubff_ctx **out, int lang, uint64_t num_programs,
                     int device);
int cubff_ctx_destroy(cubff_ctx *ctx);

/* InitPrograms launch (common_language.h:138-155, :371-373, :592-593): fills
 * programs [0, count) with SplitMix64 bytes of seed_value, or zeros. */
int cubff_init_programs(cubff_ctx *ctx, uint64_t seed_value, int zero_init,
                        uint64_t count);

/* Soup copies (common_language.h:377, :402, :502). Offsets/lengths in bytes. */
int cubff_write_soup(cubff_ctx *ctx, uint64_t offset, const uint8_t *src,
                     uint64_t len);
int cubff_read_soup(cubff_ctx *ctx, uint64_t offset, uint8_t *dst,
                    uint64_t len);

/* Caller-computed pairing for the next epoch: the shuf_idx.Write of
 * common_language.h:488, plus num_indices (:433, :442-456). */
int cubff_set_shuffle(cubff_ctx *ctx, const uint32_t *s, uint64_t num_indices);

/* Neighbour lists for CUBFF_PAIR_2D in CSR form: program i may interact with
 * nbr[row_off[i] .. row_off[i+1]), in file order.  Replaces
 * SimulationParams::allowed_interactions (common.h:71, main.cc:281-299). */
int cubff_set_interactions(cubff_ctx *ctx, const uint64_t *row_off,
                           const uint32_t *nbr);

/* Computes the pairing of `epoch` on the device in the given mode (any mode
 * but CUBFF_PAIR_HOST).  Replaces the host pairing block,
 * common_language.h:433-488. */
int cubff_pair(cubff_ctx *ctx, int mode, uint64_t params_seed, uint64_t epoch);

/* One MutateAndRunPrograms pass over the current pairing
 * (common_language.h:157-193, launched at :490-493): gather, mutate with
 * probability mutation_prob / 2^30 per byte, evaluate up to 8192 steps for
 * slots below num_indices, scatter, accumulate ops.  seed_value is the
 * epoch seed S(epoch) of :367-369. */
int cubff_run_epoch(cubff_ctx *ctx, uint64_t seed_value,
                    uint32_t mutation_prob);

/* `count` whole epochs starting at first_epoch: device pairing in `mode`
 * (not CUBFF_PAIR_HOST) followed by cubff_run_epoch with S(epoch), all on the
 * context stream.  The loop body of common_language.h:432-494 without the
 * callback block. */
int cubff_run_epochs(cubff_ctx *ctx, int mode, uint64_t params_seed,
                     uint64_t first_epoch, uint64_t count,
                     uint32_t mutation_prob);

/* cubff_run_epochs with HIP events recorded on the context stream: *total_ms
 * receives the device time of the whole sequence, *epoch_kernel_ms the summed
 * duration of the epoch (mutate + evaluate) launches alone.  Blocks until
 * done.  No reference counterpart: the reference times wall-clock intervals
 * between callbacks (common_language.h:497, :508-517). */
/* Multi-epoch runs (cubff_run_epochs / _timed with the shuffle, fixed or
 * cyclic pairing and count >= 2) can go through a dataflow kernel: the
 * pairings of up to 64 epochs are drawn first, then one launch runs every
 * interaction as soon as the interactions of the previous epoch that wrote
 * its two programs are done.  Results are bit-identical to the
 * epoch-by-epoch loop.  mode CUBFF_FLOW_OFF selects that loop, _ON the
 * dataflow kernel, _AUTO (the default) the dataflow kernel for the
 * languages and sizes where it measured faster, _DEFER one epoch kernel
 * per epoch that gives up on interactions running longer than a command
 * budget (and on those whose programs they still owe) and hands them to a
 * resolver kernel running concurrently on a second stream.  wait_limit_s bounds any
 * single wait inside the kernel (default 10 s), past which the run stops
 * with CUBFF_EHIP and an undefined soup.  No reference counterpart. */
#define CUBFF_FLOW_OFF 0
#define CUBFF_FLOW_ON 1
#define CUBFF_FLOW_AUTO 2
#define CUBFF_FLOW_DEFER 3  /* epoch kernels defer long interactions to a resolver */
int cubff_set_pipeline(cubff_ctx *ctx, int mode, double wait_limit_s);
/* Since the context was created: interactions run by the dataflow kernel,
 * the runs of queued interactions (up to 64, one wavefront) claimed to run
 * them, and interactions handed to the resolver by CUBFF_FLOW_DEFER runs. */
int cubff_flow_stats(cubff_ctx *ctx, uint64_t *items, uint64_t *claims, uint64_t *deferred);
/* Byte histogram of the whole soup (counts[256]): the h0 and frequent /
 * uncommon byte metrics of the callback block (common_language.h:519-540)
 * without copying the soup. */
int cubff_byte_histogram(cubff_ctx *ctx, uint64_t *counts);
/* Programs per chunk of cubff_selfrep's levelled check (0: the default,
 * 2^22, at most 2^22): its scratch holds 13 final tapes of 128 bytes per
 * program of a chunk (7 GiB at 2^22).  Frees the scratch; the next check
 * allocates it.  No reference counterpart: the reference's CheckSelfRep
 * (common_language.h:203-277) keeps no per-chain state. */
int cubff_set_selfrep_chunk(cubff_ctx *ctx, uint64_t programs);
/* Command budget of an interaction in CUBFF_FLOW_DEFER epoch kernels
 * (default 512). */
int cubff_set_defer_budget(cubff_ctx *ctx, uint32_t budget);
/* How long the resolver of CUBFF_FLOW_DEFER runs waits, with nothing
 * deferred and no epoch kernel finishing, before it leaves (default 0.1 s;
 * then the epoch kernels run every interaction themselves).  0 makes it
 * leave at once unless something is already deferred.  No reference
 * counterpart. */
int cubff_set_resolver_idle(cubff_ctx *ctx, double seconds);

int cubff_run_epochs_timed(cubff_ctx *ctx, int mode, uint64_t params_seed,
                           uint64_t first_epoch, uint64_t count,
                           uint32_t mutation_prob, float *total_ms,
                           float *epoch_kernel_ms);

/* Interval counters: ops summed since the last reset (insn_count,
 * common_language.h:499-501), the number of evaluated slots (num_runs,
 * :494) and the loop iterations including no-ops (no reference counterpart;
 * the work measure of the roofline).  Each pointer may be null. */
int cubff_read_counters(cubff_ctx *ctx, uint64_t *ops, uint64_t *num_runs,
                        uint64_t *steps);
int cubff_reset_counters(cubff_ctx *ctx);  /* :585-587 */

/* Per-slot results of the last epoch: ops (the Evaluate return value) and
 * loop iterations including no-ops.  No reference counterpart: the reference
 * only keeps the sum (common_language.h:192). */
int cubff_read_pair_ops(cubff_ctx *ctx, uint32_t *ops, uint32_t *steps,
                        uint64_t num_slots);

/* Pairing of the last epoch (SimulationState::shuffle_idx, common.h:76,
 * filled at common_language.h:407-486) and its num_indices. */
int cubff_read_shuffle(cubff_ctx *ctx, uint32_t *s, uint64_t *num_indices);

/* CheckSelfRep over programs [0, count) with seed_value S(epoch)
 * (common_language.h:203-277, launched at :563-569); result is
 * SimulationState::replication_per_prog (common.h:90). */
int cubff_selfrep(cubff_ctx *ctx, uint64_t seed_value, uint64_t count,
                  uint64_t *result);

/* Evaluates `count` independent 128-byte tapes of language `lang` in place
 * (tapes is count*128 bytes, host memory), up to `steps` steps each; ops and
 * steps (nullable) receive per-tape results.  The batched form of
 * RunOneProgram (common_language.h:195-201) / RunSingleParsedProgram
 * (:279-296), used by --run and by parity checks. */
int cubff_eval_tapes(int lang, int device, uint8_t *tapes, uint64_t count,
                     uint64_t steps, uint32_t *ops, uint32_t *steps_out);

/* cubff_eval_tapes without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The tapes and per-tape results are staged in a fixed device
 * array that each GPU holds until the process ends
 * (CUBFF_NOALLOC_MAX_TAPES * 136 bytes, about 2.1 MiB); count above
 * CUBFF_NOALLOC_MAX_TAPES is rejected with CUBFF_EINVAL.  Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU race.
 * Same arguments and results as cubff_eval_tapes. */
#define CUBFF_NOALLOC_MAX_TAPES 16384
int cubff_eval_tapes_noalloc(int lang, int device, uint8_t *tapes, uint64_t count,
                             uint64_t steps, uint32_t *ops, uint32_t *steps_out);

/* Step trace of one BFF-family tape for the reference's visualiser: the
 * file image python/bff_interpreter.py:evaluate_and_save (:327-393) writes
 * -- "BFF\0", u32 version 1, u32 tape size 128, then before every step u32
 * pc, head0, head1 and the 128 tape bytes (little-endian).  out must hold
 * 12 + 140 * steps bytes; *file_bytes receives the image's length and *ops
 * the Evaluate result (bff.inc.h:356-389). */
int cubff_trace_bff(int lang, int device, const uint8_t *tape, uint64_t steps, uint8_t *out,
                    uint64_t out_bytes, uint64_t *file_bytes, uint64_t *ops);

/* Step-by-step record of one Evaluate of a 128-byte tape in any language:
 * before every step, the state the reference prints under `debug`
 * (bff.inc.h:368-372, forth.inc.h:385-388, subleq.cu:114-118,
 * rsubleq4.cu:122-126), as CUBFF_STEP_RECORD_BYTES-byte records: u32 pc,
 * u32 r0, r1, r2 (BFF: head0, head1; Forth: stack depth; SUBLEQ: the two
 * operand addresses; RSUBLEQ4: the three), then the 128 tape bytes and 128
 * stack bytes (Forth; zeros otherwise).  Up to `cap` records are kept
 * (records holds cap * CUBFF_STEP_RECORD_BYTES bytes); the run continues to
 * its end.  tape is updated in place; *ops receives the Evaluate result,
 * *steps_run the steps taken, *num_records the records written.  Serves
 * --debug of RunSingleProgram (common_language.h:196-201, main.cc:179). */
#define CUBFF_STEP_RECORD_BYTES 272
int cubff_trace_steps(int lang, int device, uint8_t *tape, uint64_t steps, uint8_t *records,
                      uint64_t cap, uint64_t *num_records, uint64_t *ops, uint64_t *steps_run);

/* Single-program self-replication check of EvalParsedSelfrep
 * (common_language.h:326-342): prog is zero-padded to 64 bytes and scored
 * with seed_value = S(epoch) computed from `seed`. */
int cubff_eval_selfrep(int lang, int device, const uint8_t *prog,
                       uint64_t len, uint64_t epoch, uint64_t seed,
                       uint64_t *result);

/* The same check, also returning the tapes CheckSelfRep prints under
 * `debug` (common_language.h:225-255): for each of the 13 trials and each of
 * its 5 generations, the 128-byte tape before and after the Evaluate, in
 * that order (tapes holds CUBFF_SELFREP_TRACE_BYTES bytes). */
#define CUBFF_SELFREP_TRACE_BYTES (13 * 5 * 2 * 128)
int cubff_eval_selfrep_trace(int lang, int device, const uint8_t *prog, uint64_t len,
                             uint64_t epoch, uint64_t seed, uint64_t *result, uint8_t *tapes);

/* Waits for all work queued on the context (Synchronize, :65). */
int cubff_sync(cubff_ctx *ctx);

/* Host-side SplitMix64 and epoch seed, so callers compute S(x) exactly as the
 * device does (common_language.h:131-136, :367-369). */
uint64_t cubff_splitmix64(uint64_t x);
uint64_t cubff_epoch_seed(uint64_t params_seed, uint64_t x);


/* ------------------------------------------------------------------------ */
/* Batched soups.  No reference counterpart: the reference runs one soup per
 * process.  A batch advances many independent soups (each with its own
 * language, size, params seed and mutation threshold) with one set of
 * launches per epoch; every soup's bytes, pairings and counters equal those
 * of a separate single-soup run (RunSimulation, common_language.h:344-597,
 * with Fisher-Yates pairing). */
typedef struct cubff_batch cubff_batch;

#define CUBFF_SCHED_PLAIN 0    /* slots in soup order */
#define CUBFF_SCHED_LEARNED 1  /* MLP-predicted length, longest bucket first */

const char *cubff_batch_last_error(void);
int cubff_batch_create(cubff_batch **out, uint32_t num_soups, const int *langs,
                       const uint64_t *num_programs, int device);
int cubff_batch_destroy(cubff_batch *b);
/* Sets each soup's params seed and mutation threshold (of 2^30) and fills it
 * with InitPrograms(S(0)) (common_language.h:371-373); clears counters. */
int cubff_batch_init(cubff_batch *b, const uint64_t *params_seeds,
                     const uint32_t *mutation_probs, int zero_init);
/* Interpreter launch shape: workgroup size (64, 128 or 256 lanes) and tapes
 * evaluated one after another by each lane.  Default 64 lanes x 2 tapes
 * (what cubff_batch_autotune picks for the 64-soup bff_noheads batch). */
int cubff_batch_set_launch(cubff_batch *b, int wg, uint32_t tapes_per_lane);
int cubff_batch_get_launch(cubff_batch *b, int *wg, uint32_t *tapes_per_lane);
/* `count` epochs from first_epoch; use_graph replays one captured epoch. */
int cubff_batch_run_epochs(cubff_batch *b, uint64_t first_epoch, uint64_t count,
                           int sched, int use_graph);
/* As above with HIP events: device time of the whole run and of the
 * interpreter launches alone (-1 with use_graph: the replayed epoch is one
 * graph launch; its capture happens before the timed region).  Blocks until
 * done. */
int cubff_batch_run_epochs_timed(cubff_batch *b, uint64_t first_epoch,
                                 uint64_t count, int sched, int use_graph,
                                 float *total_ms, float *interp_ms);
/* Runs epochs_per_config live epochs under each of the 12 launch shapes
 * (workgroup 64/128/256 x 1-4 tapes per lane), in rounds of one epoch per
 * shape (two when runs go through deferring segments, which then run
 * 2 * epochs_per_config per shape) so that every shape sees the soups at
 * the same stage, keeps the fastest by total time and reports it; ms_table
 * (nullable, 12 floats) receives ms/epoch per shape.  Results do not depend
 * on the launch shape, so the tuning epochs are real epochs.  For the
 * persistent interpreters "tapes per lane" is the wavefronts per SIMD of
 * their grid. */
int cubff_batch_autotune(cubff_batch *b, uint64_t first_epoch,
                         uint64_t epochs_per_config, int sched, int *best_wg,
                         uint32_t *best_tapes_per_lane, float *ms_table);
int cubff_batch_sync(cubff_batch *b);
int cubff_batch_read_soup(cubff_batch *b, uint32_t soup, uint8_t *dst);
int cubff_batch_write_soup(cubff_batch *b, uint32_t soup, const uint8_t *src);
int cubff_batch_read_counters(cubff_batch *b, uint32_t soup, uint64_t *ops,
                              uint64_t *num_runs, uint64_t *steps);
int cubff_batch_reset_counters(cubff_batch *b);
/* Per-interaction results of the last epoch of one soup, in pair order
 * (each pointer nullable, n/2 values): ops and steps as RunSingleProgram
 * reports them (common_language.h:232-242), and work = the commands the
 * interpreter actually executed (skipped loop iterations and no-op runs not
 * counted; the learned scheduler's training target). */
int cubff_batch_read_pair_stats(cubff_batch *b, uint32_t soup, uint32_t *ops,
                                uint32_t *steps, uint32_t *work);
int cubff_batch_read_shuffle(cubff_batch *b, uint32_t soup, uint32_t *s);
/* Learned-scheduler statistics: trainings so far, the loss of the last
 * training round, and the last epoch's slot count per length bucket
 * (8 counts, bucket 0 = longest predicted). */
/* Multi-epoch path of runs of two or more epochs without graph replay:
 * deferring segments run the first language of the batch that the
 * deferring epoch kernel supports (bff, bff_noheads, bff_noheads_4bit,
 * bff_perm) with a persistent kernel whose lanes take new pairs as they
 * finish and hand interactions past the command budget to a resolver that
 * finishes them beside the later epochs (as cubff_set_pipeline's
 * CUBFF_FLOW_DEFER does for one soup).  CUBFF_FLOW_AUTO (default) uses them
 * under CUBFF_SCHED_PLAIN, CUBFF_FLOW_DEFER under both schedules,
 * CUBFF_FLOW_OFF never (every epoch through the interpreter rounds).
 * Results are the same. */
int cubff_batch_set_pipeline(cubff_batch *b, int mode);
/* The epoch after the last one run (0 before any run). */
int cubff_batch_next_epoch(cubff_batch *b, uint64_t *next);
/* Command budget of an interaction in the deferring epoch kernel (default
 * 768; >= 1). */
int cubff_batch_set_defer_budget(cubff_batch *b, uint32_t budget);
/* Interactions handed to the resolver since the batch was created. */
int cubff_batch_defer_stats(cubff_batch *b, uint64_t *deferred);
int cubff_batch_sched_stats(cubff_batch *b, uint64_t *trainings,
                            float *last_loss, uint32_t *bucket_counts);

/* ------------------------------------------------------------------------ */
/* The MLP cost model of the learned scheduler on its own.  No reference
 * counterpart.  Input: 128-byte tapes (features: byte histogram / 128);
 * output: a predicted log2 length (the scheduler trains it on the commands
 * each interaction ran).
 * Parameters are one fp32 vector of cubff_mlp_num_params() values:
 * E[256][256], W1[256][256], b1[256], W2[256][256], b2[256], w3[256], b3. */
typedef struct cubff_mlp cubff_mlp;
const char *cubff_mlp_last_error(void);
uint64_t cubff_mlp_num_params(void);
int cubff_mlp_create(cubff_mlp **out, uint64_t capacity, uint64_t init_seed,
                     int device);
int cubff_mlp_destroy(cubff_mlp *h);
int cubff_mlp_set_params(cubff_mlp *h, const float *params);
int cubff_mlp_get_params(cubff_mlp *h, float *params);
int cubff_mlp_forward(cubff_mlp *h, const uint8_t *tapes, uint64_t count,
                      float *pred);
/* The scheduler's inference path: one fused kernel (features, both hidden
 * layers with E W1 folded, output).  Agrees with cubff_mlp_forward to bf16
 * rounding of the folded first layer. */
int cubff_mlp_predict(cubff_mlp *h, const uint8_t *tapes, uint64_t count,
                      float *pred);
/* The same kernel reading each tape as two 64-byte pieces: tape r is pieces
 * index[2r] then index[2r + 1] of `pieces` (num_pieces * 64 bytes, at most
 * 2 * capacity pieces; every index below num_pieces).  This is how the
 * batched learned scheduler reads the next epoch's tapes out of the current
 * epoch's stage.  Predictions equal cubff_mlp_predict's on the assembled
 * tapes. */
int cubff_mlp_predict_pieces(cubff_mlp *h, const uint8_t *pieces, uint64_t num_pieces,
                             const uint32_t *index, uint64_t count, float *pred);
/* Loss 0.5 * mean((pred - target)^2) and its gradient (same layout as the
 * parameters). */
int cubff_mlp_gradients(cubff_mlp *h, const uint8_t *tapes,
                        const float *targets, uint64_t count, float *loss,
                        float *grads);
/* `steps` full-batch Adam steps; losses (nullable) receives each step's
 * loss before its update. */
int cubff_mlp_train(cubff_mlp *h, const uint8_t *tapes, const float *targets,
                    uint64_t count, uint32_t steps, float lr, float *losses);

#ifdef __cplusplus
}
#endif

#endif /* CUBFF_HIP_H_ */
