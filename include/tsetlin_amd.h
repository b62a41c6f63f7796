/*
 * tsetlin_amd.h -- C ABI of the MI355X (gfx950) Tsetlin-machine engine.
 *
 * Drop-in boundary for the classifier/encoder train and predict path of
 * santoku.tsetlin (module santoku.tsetlin.capi, lib/santoku/tsetlin/capi.c).
 * Every entry point names the reference function it replaces.  Plain C types
 * only: pointers, sizes, a callback.  Host pointers unless the name ends in
 * _dev (device pointers, data already resident in HBM).
 *
 * Error convention: every call returns TM_OK (0) or a negative TM_E* code;
 * tm_last_error() returns a thread-local message for the last failure.  The
 * reference raises luaL_error / tk_lua_verror instead; a Lua shim maps a
 * non-zero status to luaL_error(L, "%s", tm_last_error()).
 *
 * Handles are not re-entrant: one thread at a time per handle (the reference
 * is called from one Lua thread).  Each handle owns a HIP stream.
 */
#ifndef TSETLIN_AMD_H
#define TSETLIN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* capi.c:1802 (lua_pushinteger(L, TK_CVEC_BITS); "align"). */
#define TM_ALIGN 8

/* capi.c:41-44 (tk_tsetlin_type_t); persisted as int32 at capi.c:1543. */
#define TM_CLASSIFIER 0
#define TM_ENCODER 1

#define TM_OK 0
#define TM_E_ARG (-1)        /* bad argument (tk_lua_verror in the reference) */
#define TM_E_ALLOC (-2)      /* host or device allocation failed */
#define TM_E_STATE (-3)      /* operation not valid in this state */
#define TM_E_FORMAT (-4)     /* malformed persisted data */
#define TM_E_SIZE (-5)       /* buffer size mismatch */
#define TM_E_DEVICE (-6)     /* HIP runtime error / no device */
#define TM_E_UNSUPPORTED (-7)/* valid for the reference, not built here */

typedef struct tm_machine tm_machine;

/* Options of create/reconfigure.  Field names are the Lua option keys of
 * capi.c:631-660 (classifier) and :668-697 (encoder): for an encoder,
 * `classes` carries "hidden" and `features` carries "visible". */
typedef struct {
  uint32_t classes;          /* required, > 0 */
  uint32_t features;         /* required */
  uint32_t clauses;          /* required, > 0; rounded up to a multiple of 8 */
  uint32_t clause_tolerance; /* required, > 0 */
  uint32_t clause_maximum;   /* required, > 0 */
  uint32_t state;            /* "state" (state bits), default 8, >= 2; at most 16
                                here (TM_E_UNSUPPORTED above); rows up to 4096 bytes
                                train at <= 8 state bits, up to 2048 bytes above */
  uint32_t include_bits;     /* default 1 (0 means 1) */
  double target;             /* default -1: sqrt(clauses/2)*tolerance */
  double negative;           /* default -1: 1/classes */
  double specificity;        /* required, > 0 */
  int32_t reusable;          /* default 0 */
  int32_t individualized;    /* default 0; individualized machines are created
                                by tm_create_*_ind (1 here is TM_E_ARG) */
  uint64_t seed;             /* RNG seed of train (no counterpart: the reference
                                seeds each OpenMP thread with its thread id) */
} tm_options;

/* Fill `o` with the reference's defaults (state 8, include_bits 1, target -1,
 * negative -1, reusable 0, seed 0); required fields are zeroed. */
void tm_options_default(tm_options *o);

/* Epoch callback: the `each` option of train (capi.c:1168-1182).  Called on the
 * calling thread after each epoch with the 1-based epoch number, after the
 * epoch's actions are visible to tm_predict*.  Return 1 to continue, 0 to stop
 * (the Lua callback returning false), negative to report an error (printed
 * to stderr; training stops, and train still returns TM_OK, as the reference
 * does when lua_pcall fails). */
typedef int (*tm_epoch_fn)(void *user, uint32_t epoch);

/* create("classifier", opts) -> tk_tsetlin_create_classifier, capi.c:627-660,
 * init at capi.c:449-506. */
int tm_create_classifier(const tm_options *opts, tm_machine **out);

/* create("encoder", opts) -> tk_tsetlin_create_encoder, capi.c:662-697,
 * init at capi.c:508-526. */
int tm_create_encoder(const tm_options *opts, tm_machine **out);

/* machine:train{samples, problems, solutions, iterations, each}
 * -> tk_tsetlin_train_classifier, capi.c:1112-1191.
 * problems: n rows of input_chunks bytes (flip-interleaved literals);
 * solutions: n int64 class labels. */
int tm_train_classifier(tm_machine *m, uint32_t n, const uint8_t *problems, const int64_t *solutions,
                        uint32_t iterations, tm_epoch_fn each, void *user);

/* machine:train{samples, sentences, codes, iterations, each}
 * -> tk_tsetlin_train_encoder, capi.c:1193-1282.
 * codes: n rows of ceil(hidden/8) bytes, bit h of a row is the target of
 * hidden bit h. */
int tm_train_encoder(tm_machine *m, uint32_t n, const uint8_t *sentences, const uint8_t *codes,
                     uint32_t iterations, tm_epoch_fn each, void *user);

/* Same as above with problems/solutions/codes already in device memory
 * (same byte layouts).  No counterpart in the reference (its data is host
 * memory); for callers that keep datasets resident in HBM. */
int tm_train_classifier_dev(tm_machine *m, uint32_t n, const uint8_t *problems_dev, const int64_t *solutions_dev,
                            uint32_t iterations, tm_epoch_fn each, void *user);
int tm_train_encoder_dev(tm_machine *m, uint32_t n, const uint8_t *sentences_dev, const uint8_t *codes_dev,
                         uint32_t iterations, tm_epoch_fn each, void *user);

/* machine:predict(problems, n) for a classifier -> tk_tsetlin_predict_classifier,
 * capi.c:903-948.  out: n int64 class ids. */
int tm_predict_classifier(tm_machine *m, const uint8_t *problems, uint32_t n, int64_t *out);

/* machine:predict(problems, n) for an encoder -> tk_tsetlin_predict_encoder,
 * capi.c:950-995.  out: n * ceil(hidden/8) bytes (padding bits zero). */
int tm_predict_encoder(tm_machine *m, const uint8_t *problems, uint32_t n, uint8_t *out);

/* Device-pointer variants of predict (inputs and outputs in device memory). */
int tm_predict_classifier_dev(tm_machine *m, const uint8_t *problems_dev, uint32_t n, int64_t *out_dev);
int tm_predict_encoder_dev(tm_machine *m, const uint8_t *problems_dev, uint32_t n, uint8_t *out_dev);

/* tm_predict_classifier_dev without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * class sums, and the rows when input_chunks is not a multiple of 4 or the
 * rows are not 16-byte aligned (then repacked into whole words), live in a
 * library-owned device array of 32 MiB per GPU, held until the process ends:
 * TM_E_SIZE when n * classes * 8 bytes (plus n * ceil(input_chunks / 4) * 4
 * for repacked rows) exceed it.  Every call on one GPU shares that array, so
 * calls that overlap on one GPU race.  Same results as
 * tm_predict_classifier_dev. */
int tm_predict_classifier_dev_noalloc(tm_machine *m, const uint8_t *problems_dev, uint32_t n, int64_t *out_dev);

/* ---------------------------------------------------------------------------
 * Individualized machines: class h (hidden bit h of an encoder) has a feature
 * set of its own, k_h = feat_offsets[h + 1] - feat_offsets[h] features, and
 * is a one-class machine over rows of ceil(2 k_h / 8) bytes.  The plain
 * train/predict entry points refuse these machines and the _ind ones refuse
 * plain machines; persist, load, checkpoint, restore, reconfigure, type and
 * destroy work on both.
 *
 * Problems of an individualized call: one buffer of problems_bytes bytes;
 * class h's row of sample s is the ceil(2 k_h / 8) bytes at byte
 * dim_offsets[h] + s * ceil(2 k_h / 8) (classes entries in dim_offsets).
 * TM_E_SIZE if some class's rows run past problems_bytes.
 */

/* create("classifier", {individualized = true, feat_offsets, ...})
 * -> tk_tsetlin_init_classifier_ind, capi.c:528-606 (dispatch capi.c:631-645).
 * n_offsets must be classes + 1; opts->features and opts->individualized are
 * not read.  An empty feature set (k_h = 0) is TM_E_ARG. */
int tm_create_classifier_ind(const tm_options *opts, const int64_t *feat_offsets, uint32_t n_offsets,
                             tm_machine **out);

/* create("encoder", {individualized = true, feat_offsets, hidden, ...})
 * -> tk_tsetlin_init_encoder_ind, capi.c:608-626 (dispatch capi.c:668-682). */
int tm_create_encoder_ind(const tm_options *opts, const int64_t *feat_offsets, uint32_t n_offsets,
                          tm_machine **out);

/* machine:train{samples, problems, solutions, dim_offsets, iterations, each}
 * -> tk_tsetlin_train_classifier_ind, capi.c:1387-1482. */
int tm_train_classifier_ind(tm_machine *m, uint32_t n, const uint8_t *problems, size_t problems_bytes,
                            const int64_t *solutions, const int64_t *dim_offsets, uint32_t iterations,
                            tm_epoch_fn each, void *user);

/* machine:train{samples, sentences, codes, dim_offsets, iterations, each}
 * -> tk_tsetlin_train_encoder_ind, capi.c:1284-1385. */
int tm_train_encoder_ind(tm_machine *m, uint32_t n, const uint8_t *sentences, size_t sentences_bytes,
                         const uint8_t *codes, const int64_t *dim_offsets, uint32_t iterations, tm_epoch_fn each,
                         void *user);

/* machine:predict(problems, dim_offsets, n) -> tk_tsetlin_predict_classifier_ind,
 * capi.c:997-1043: n int64 class ids. */
int tm_predict_classifier_ind(tm_machine *m, const uint8_t *problems, size_t problems_bytes,
                              const int64_t *dim_offsets, uint32_t n, int64_t *out);

/* machine:predict(problems, dim_offsets, n) -> tk_tsetlin_predict_encoder_ind,
 * capi.c:1045-1086: n * ceil(hidden/8) bytes. */
int tm_predict_encoder_ind(tm_machine *m, const uint8_t *problems, size_t problems_bytes,
                           const int64_t *dim_offsets, uint32_t n, uint8_t *out);

/* The four _ind calls above with problems (and solutions / codes, and the
 * predictions) in device memory: the rows are repacked on the device from
 * problems_dev, with the same dim_offsets (host) and size checks. */
int tm_train_classifier_ind_dev(tm_machine *m, uint32_t n, const uint8_t *problems_dev, size_t problems_bytes,
                                const int64_t *solutions_dev, const int64_t *dim_offsets, uint32_t iterations,
                                tm_epoch_fn each, void *user);
int tm_train_encoder_ind_dev(tm_machine *m, uint32_t n, const uint8_t *sentences_dev, size_t sentences_bytes,
                             const uint8_t *codes_dev, const int64_t *dim_offsets, uint32_t iterations,
                             tm_epoch_fn each, void *user);
int tm_predict_classifier_ind_dev(tm_machine *m, const uint8_t *problems_dev, size_t problems_bytes,
                                  const int64_t *dim_offsets, uint32_t n, int64_t *out_dev);
int tm_predict_encoder_ind_dev(tm_machine *m, const uint8_t *problems_dev, size_t problems_bytes,
                               const int64_t *dim_offsets, uint32_t n, uint8_t *out_dev);

/* machine:restrict(selected_dims) -> tk_tsetlin_restrict, capi.c:752-900:
 * keep the classes listed in `selected` (in that order), with their actions
 * and, while the machine has them, counts.  Errors as the reference: not
 * individualized, empty, out of range, duplicate (all TM_E_ARG). */
int tm_restrict(tm_machine *m, const int64_t *selected, uint32_t n_selected);

/* Feature sizes k_h of an individualized machine (the feat_sizes the
 * reference persists, capi.c:1528-1530); TM_E_SIZE if cap < classes. */
int tm_get_feature_sizes(const tm_machine *m, uint64_t *out, uint32_t cap);

/* machine:persist(nil) -> tk_tsetlin_persist, capi.c:1538-1562 (layout
 * capi.c:1509-1531).  Writes the persisted bytes into `buf` when cap is large
 * enough; *len always receives the required size.  TM_E_SIZE if cap < *len. */
int tm_persist(tm_machine *m, uint8_t *buf, size_t cap, size_t *len);

/* machine:persist(path) -> same, to a file. */
int tm_persist_file(tm_machine *m, const char *path);

/* tsetlin.load(data, true) -> tk_tsetlin_load, capi.c:1751-1772 (reader
 * capi.c:1675-1737).  The loaded machine has no automata state: train on it is
 * TM_E_STATE (the reference dereferences NULL state there). */
int tm_load(const uint8_t *data, size_t len, tm_machine **out);

/* tsetlin.load(path) -> same, from a file. */
int tm_load_file(const char *path, tm_machine **out);

/* machine:checkpoint(cvec) -> tk_tsetlin_checkpoint, capi.c:1564-1576: copy
 * the actions (action_chunks bytes, reference layout) into buf. *len gets the
 * size; TM_E_SIZE if cap is too small. */
int tm_checkpoint(tm_machine *m, uint8_t *buf, size_t cap, size_t *len);

/* machine:restore(cvec) -> tk_tsetlin_restore, capi.c:1578-1588: TM_E_SIZE
 * unless len == action_chunks. */
int tm_restore(tm_machine *m, const uint8_t *buf, size_t len);

/* checkpoint/restore with the buffer in device memory (same byte layout as
 * tm_checkpoint/tm_restore, capi.c:1564-1588).  No reference counterpart;
 * for callers that keep a best-epoch checkpoint in HBM (a search trial's
 * early stop, optimize.lua:616,648) without a round trip through the host. */
int tm_checkpoint_dev(tm_machine *m, uint8_t *buf_dev, size_t cap, size_t *len);
int tm_restore_dev(tm_machine *m, const uint8_t *buf_dev, size_t len);

/* The counters in the reference's layout: state_chunks bytes,
 * counts[clause][plane][input_chunks] with plane 0 the least significant bit
 * (automata.h:7-31; an individualized machine's classes one after another,
 * capi.c:576-593).  No reference counterpart: the reference neither persists
 * nor returns its counts (capi.c:1509-1531); this is a diagnostic read for
 * parity checks of the whole automata state.  TM_E_STATE when the machine
 * holds no counters (loaded, or released after training); TM_E_SIZE if cap
 * is too small (*len always receives the size). */
int tm_counts(tm_machine *m, uint8_t *buf, size_t cap, size_t *len);

/* machine:reconfigure(opts) -> tk_tsetlin_reconfigure, capi.c:1590-1673.
 * Reads clauses, clause_tolerance, clause_maximum, include_bits, target,
 * specificity (and seed) from opts; TM_E_STATE unless created reusable. */
int tm_reconfigure(tm_machine *m, const tm_options *opts);

/* machine:type() -> tk_tsetlin_type, capi.c:1774-1788: TM_CLASSIFIER or
 * TM_ENCODER, or a negative error code. */
int tm_type(const tm_machine *m);

/* machine:destroy() / __gc -> tk_tsetlin_destroy, capi.c:743-750.  Frees the
 * handle; NULL is ignored. */
void tm_destroy(tm_machine *m);

/* Derived sizes (the persisted header fields, capi.c:1511-1525). */
typedef struct {
  uint32_t type, classes, features, clauses, clause_tolerance, clause_maximum, target, state_bits,
           include_bits, input_bits, input_chunks, clause_chunks, state_chunks, action_chunks, class_chunks;
  double specificity, negative;
  uint8_t tail_mask, reusable, has_state, trained, individualized;
} tm_info;
int tm_get_info(const tm_machine *m, tm_info *out);

/* Thread-local message of the last failing call on this thread. */
const char *tm_last_error(void);

/* Timing of the last train/predict call, measured with HIP events on the
 * handle's stream: milliseconds inside train kernels (all epochs), and inside
 * the predict kernels. */
int tm_last_timing(const tm_machine *m, double *train_ms, double *predict_ms);

/* Milliseconds of the last train call spent building the classifier's
 * per-epoch step queues (the pre-pass kernels ahead of each train launch;
 * 0 when none was built).  No reference counterpart. */
int tm_last_queue_timing(const tm_machine *m, double *queue_ms);

/* Work counters of the last train call, summed over its epochs (no reference
 * counterpart; used to price the kernels against the VALU roofline):
 * [0] chunk-steps evaluated (negative-sampling skips excluded),
 * [1] Type I events on a true clause, [2] of which grew the clause
 * (literals <= clause_maximum), [3] Type II events on a true clause,
 * [4] Type I events on a false clause, [5] random-literal draws. */
#define TM_NSTATS 6
int tm_last_train_stats(const tm_machine *m, uint64_t out[TM_NSTATS]);

/* ---------------------------------------------------------------------------
 * Device-resident data preparation and dev metrics (device pointers in and
 * out; each call runs on the legacy default stream and returns when done).
 */

/* ivec:bits_to_cvec(n, features, true) of santoku-matrix (layout
 * doc/sth.md:85-91; the reference's callers, classify/mnist.lua:60): set_bits
 * holds sample * features + feature for each present feature, in any order,
 * duplicates allowed.  rows_dev receives n flip-interleaved rows of
 * ceil(2 features / 8) bytes (x bits, then not-x bits, LSB first, padding
 * zero).  Indices outside [0, n * features) are skipped and counted in
 * *n_invalid (may be NULL). */
int tm_bits_to_cvec_dev(const int64_t *set_bits_dev, uint64_t n_set, uint32_t n, uint32_t features, uint8_t *rows_dev,
                        uint64_t *n_invalid);

/* ivec:bits_to_cvec_ind(ind_offsets, feat_offsets, n, true) of santoku-matrix
 * (doc/sth.md:39-55; the reference's callers, classify/newsgroups.lua:97,
 * encoder/mnist.lua:397): the packer of individualized machines' problems.
 * ind_dev holds, for class h, the entries ind[ind_offsets[h] ..
 * ind_offsets[h + 1]) = sample * k_h + feature (k_h = feat_offsets[h + 1] -
 * feat_offsets[h], the feature an index into class h's own feature list), in
 * any order, duplicates allowed; ind_offsets and feat_offsets are host arrays
 * of classes + 1 entries starting at 0.  Class h's block, n flip-interleaved
 * rows of ceil(2 k_h / 8) bytes as tm_bits_to_cvec_dev writes them, goes to
 * bitmap_dev + dim_offsets[h], the blocks back to back in class order;
 * dim_offsets (host, classes + 1 entries) receives the offsets and the total,
 * which must not exceed bitmap_bytes (TM_E_SIZE).  Entries outside [0, n k_h)
 * are skipped and counted in *n_invalid (may be NULL).  The bitmap and
 * dim_offsets are what the _ind train and predict calls take. */
int tm_bits_to_cvec_ind_dev(const int64_t *ind_dev, const int64_t *ind_offsets, const int64_t *feat_offsets,
                            uint32_t classes, uint32_t n, uint8_t *bitmap_dev, uint64_t bitmap_bytes,
                            int64_t *dim_offsets, uint64_t *n_invalid);

/* The same rows from bit-packed ones: raw row s is ceil(features / 8) bytes,
 * bit f (LSB first) = feature f.  No reference counterpart (a dense form of
 * the packer above). */
int tm_flip_interleave_dev(const uint8_t *raw_dev, uint32_t n, uint32_t features, uint8_t *rows_dev);

/* evaluator.class_accuracy -> tm_class_accuracy, evaluator/capi.c:119-203:
 * per_class (3 * classes doubles: precision, recall, f1 per class; may be
 * NULL) and macro (precision, recall, f1).  Predictions and labels are int64,
 * converted to unsigned int as the reference does; samples with either
 * outside [0, classes) are skipped. */
int tm_class_accuracy_dev(const int64_t *predicted_dev, const int64_t *expected_dev, uint32_t n, uint32_t classes,
                          double *per_class, double *macro);

/* evaluator.encoding_accuracy -> tm_encoding_accuracy, evaluator/capi.c:497-574:
 * bits (hidden doubles, per-bit error rate; may be NULL) and summary
 * (mean_hamming, ber_min, ber_max, ber_std). */
int tm_encoding_accuracy_dev(const uint8_t *predicted_dev, const uint8_t *expected_dev, uint32_t n, uint32_t hidden,
                             double *bits, double *summary);

/* ---------------------------------------------------------------------------
 * Batched search trials.  No reference counterpart: optimize.lua runs search
 * trials one after another on one reusable machine (optimize.lua:730-768).
 * Here N classifiers (or N encoders, for the encoder search of
 * optimize.lua:842) that share data, classes (hidden bits) and features but
 * differ in clauses, clause_tolerance, clause_maximum, target, specificity,
 * include_bits and seed train together, one launch per epoch for all of
 * them (replayed from a captured HIP graph after the first epoch).  Each
 * trial's automata are bit-identical to a solo tm_train_classifier (or
 * tm_train_encoder) run with the same options.  The tile shape and workgroup
 * order are chosen by timing the candidates on the first epoch (the
 * autotuner); the choice is cached per (row words, state bits, row stride,
 * machine type, workgroup count rounded down to a power of two, capped at
 * 4096), so later batches of similar size reuse it.
 */
typedef struct tm_batch tm_batch;

/* Called after each epoch with the 1-based epoch number; active[t] is 1 for
 * trials still training.  Clearing active[t] freezes trial t (early stop);
 * returning 0 stops the whole batch, negative reports an error and stops. */
typedef int (*tm_batch_epoch_fn)(void *user, uint32_t epoch, uint8_t *active);

int tm_batch_create(const tm_options *opts, uint32_t ntrials, tm_batch **out);
/* Borrowed handle of trial t (valid until tm_batch_destroy): predict,
 * checkpoint, restore and persist work on it as on any machine. */
int tm_batch_machine(tm_batch *b, uint32_t trial, tm_machine **out);
int tm_batch_train_classifier(tm_batch *b, uint32_t n, const uint8_t *problems, const int64_t *solutions,
                              uint32_t iterations, tm_batch_epoch_fn each, void *user);
int tm_batch_train_classifier_dev(tm_batch *b, uint32_t n, const uint8_t *problems_dev,
                                  const int64_t *solutions_dev, uint32_t iterations, tm_batch_epoch_fn each,
                                  void *user);
/* A batch of encoder trials (the options' classes are the hidden bits), and
 * its training on sentences and codes laid out as tm_train_encoder takes
 * them (codes: [n][(classes + 7) / 8] bytes).  A batch trains only as its own type;
 * the accuracy calls below score classifier batches (an encoder trial is
 * scored through tm_batch_machine: tm_predict_encoder_dev and
 * tm_encoding_accuracy_dev). */
int tm_batch_create_encoder(const tm_options *opts, uint32_t ntrials, tm_batch **out);
int tm_batch_train_encoder(tm_batch *b, uint32_t n, const uint8_t *sentences, const uint8_t *codes,
                           uint32_t iterations, tm_batch_epoch_fn each, void *user);
int tm_batch_train_encoder_dev(tm_batch *b, uint32_t n, const uint8_t *sentences_dev, const uint8_t *codes_dev,
                               uint32_t iterations, tm_batch_epoch_fn each, void *user);
/* Batches of individualized trials (the default of the reference's encoder
 * and newsgroups specs, encoder/mnist.lua:23, classify/newsgroups.lua:29):
 * every trial takes the same feat_offsets (classes + 1 offsets, as
 * tm_create_classifier_ind / tm_create_encoder_ind), and training takes the
 * problems' byte size and dim_offsets as tm_train_classifier_ind /
 * tm_train_encoder_ind do.  One launch per epoch trains every class of every
 * trial; each trial equals its solo individualized machine.  Classifier
 * batches are scored by tm_batch_class_accuracy_ind_dev, encoder batches
 * through tm_batch_machine (tm_predict_encoder_ind_dev). */
int tm_batch_create_classifier_ind(const tm_options *opts, uint32_t ntrials, const int64_t *feat_offsets,
                                   uint32_t n_offsets, tm_batch **out);
int tm_batch_create_encoder_ind(const tm_options *opts, uint32_t ntrials, const int64_t *feat_offsets,
                                uint32_t n_offsets, tm_batch **out);
int tm_batch_train_classifier_ind(tm_batch *b, uint32_t n, const uint8_t *problems, size_t problems_bytes,
                                  const int64_t *solutions, const int64_t *dim_offsets, uint32_t iterations,
                                  tm_batch_epoch_fn each, void *user);
int tm_batch_train_classifier_ind_dev(tm_batch *b, uint32_t n, const uint8_t *problems_dev, size_t problems_bytes,
                                      const int64_t *solutions_dev, const int64_t *dim_offsets, uint32_t iterations,
                                      tm_batch_epoch_fn each, void *user);
int tm_batch_train_encoder_ind(tm_batch *b, uint32_t n, const uint8_t *sentences, size_t sentences_bytes,
                               const uint8_t *codes, const int64_t *dim_offsets, uint32_t iterations,
                               tm_batch_epoch_fn each, void *user);
int tm_batch_train_encoder_ind_dev(tm_batch *b, uint32_t n, const uint8_t *sentences_dev, size_t sentences_bytes,
                                   const uint8_t *codes_dev, const int64_t *dim_offsets, uint32_t iterations,
                                   tm_batch_epoch_fn each, void *user);
/* tm_batch_class_accuracy_dev for a batch of individualized classifiers: the
 * dev problems buffer read at dim_offsets (as tm_predict_classifier_ind_dev),
 * macro[3 t .. 3 t + 2] of trial t; one synchronisation for all trials. */
int tm_batch_class_accuracy_ind_dev(tm_batch *b, const uint8_t *problems_dev, size_t problems_bytes,
                                    const int64_t *solutions_dev, const int64_t *dim_offsets, uint32_t n,
                                    double *macro);
/* Held-out accuracy of every trial on device-resident rows and labels. */
int tm_batch_accuracy_dev(tm_batch *b, const uint8_t *problems_dev, const int64_t *solutions_dev, uint32_t n,
                          double *accuracy);
/* evaluator.class_accuracy of every trial (evaluator/capi.c:119-203, the
 * search metric of the reference's classifier specs, classify/mnist.lua:93-97):
 * macro[3 t .. 3 t + 2] = macro precision, recall, F1 of trial t, the
 * reference's doubles.  One synchronisation for all trials. */
int tm_batch_class_accuracy_dev(tm_batch *b, const uint8_t *problems_dev, const int64_t *solutions_dev, uint32_t n,
                                double *macro);
/* The autotuner's choice: lanes per clause, words per lane, workgroup order
 * (0 trial-major, 1 heaviest trials first) and its epoch time. */
int tm_batch_tuning(const tm_batch *b, uint32_t *lanes, uint32_t *wpl, uint32_t *order, double *epoch_ms);
/* Milliseconds of the last train call inside train kernels and of the last
 * accuracy call. */
int tm_batch_timing(const tm_batch *b, double *train_ms, double *eval_ms);
void tm_batch_destroy(tm_batch *b);

/* ---------------------------------------------------------------------------
 * Search surrogate: an FP32 MLP 12 -> 64 -> 64 -> 64 -> 1 (ReLU hidden layers,
 * linear output) trained with full-batch mean squared error and Adam by the
 * engine's own HIP kernels.  No reference counterpart: it predicts a trial's
 * dev accuracy from its hyper-parameters so the search can propose trials.
 * Parameters: one float array of TM_MLP_PARAMS values laid out as
 * W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4.
 */
#define TM_MLP_INPUTS 12
#define TM_MLP_HIDDEN 64
#define TM_MLP_PARAMS 9217
typedef struct tm_mlp tm_mlp;
int tm_mlp_create(const float *params, tm_mlp **out);
/* Start again from `params` with zeroed Adam moments and step count, keeping
 * the network's stream and buffers (a search refits every round). */
int tm_mlp_reset(tm_mlp *m, const float *params);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8) on n samples X[n][12] -> Y[n];
 * *loss receives the mean squared error of the last step (may be NULL). */
int tm_mlp_train(tm_mlp *m, const float *X, const float *Y, uint32_t n, uint32_t steps, float lr, float *loss);
/* tm_mlp_train of k distinct networks at once, network i on X[i][n[i]][12] ->
 * Y[i][n[i]]: up to 16 networks on one device share one launch per pass
 * (more, or several devices: each network's steps on its own stream), so the
 * fits run side by side; each ends with the parameters tm_mlp_train would
 * give it.  loss[i] receives network i's last-step loss (loss may be NULL). */
int tm_mlp_train_many(tm_mlp *const *nets, uint32_t k, const float *const *X, const float *const *Y,
                      const uint32_t *n, uint32_t steps, float lr, float *loss);
/* Gradient of the mean squared error at the current parameters (no update). */
int tm_mlp_gradient(tm_mlp *m, const float *X, const float *Y, uint32_t n, float *grad, float *loss);
int tm_mlp_predict(tm_mlp *m, const float *X, uint32_t n, float *Y);
int tm_mlp_params(tm_mlp *m, float *out);
const char *tm_mlp_last_error(void);
void tm_mlp_destroy(tm_mlp *m);

#ifdef __cplusplus
}
#endif

#endif
