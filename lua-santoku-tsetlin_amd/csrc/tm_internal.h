// tm_internal.h -- launch interface between the host library (tm_api.cpp) and
// the gfx950 kernels (tm_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Training launch.  Device layouts (u32 words, Wd = ceil(input_chunks/4)):
//   counts  [clause][plane 0..m-1][Wd]   bit-sliced counters, plane 0 = LSB
//   actions [clause][Wd]                 the include plane (MSB of each counter)
//   rows    [n][row_stride]              flip-interleaved literal rows
//   perm    [epoch - perm_epoch0][n]     sample order of each epoch
//   ekeys   [epoch - perm_epoch0][4]     per-epoch key words
//
// Per-machine ("trial") parameters live in tm_train_trial; a launch covers one
// machine (`one`) or a table of machines sharing the data (`trials`, with
// `block_map` giving each workgroup's (trial, chunk)).
struct tm_train_trial {
  uint32_t *counts;
  uint32_t *actions;
  const uint32_t *perm;
  const uint32_t *ekeys;
  unsigned long long *stats;      // optional work counters, TM_NSTATS entries (atomically added)
  uint64_t skip_threshold;        // negative sampling: skip iff draw < threshold
  uint32_t perm_epoch0;           // epoch of perm/ekeys row 0
  uint32_t clause_chunks;
  uint32_t tolerance, maximum, target, s_draws, include_bits;
  uint32_t seed_lo, seed_hi;
  uint32_t active;                // 0: the trial is frozen (its workgroups return at once)
  const uint32_t *rows;           // literal rows of this trial [n][row_stride]
  uint32_t row_stride, wd;        // words between rows (a multiple of the tile's words per lane), per row
  uint32_t input_bits;            // literal bits (the range of the random-literal draws)
  uint32_t last_live, last_eval;  // masks of the last word (Wd-1)
  uint32_t class_base;            // class (label / code bit) of the trial's chunk 0
  uint32_t chunk_base;            // index of the trial's chunk 0 in the random stream
};

struct tm_train_launch {
  tm_train_trial one;
  const tm_train_trial *trials;   // device table, or NULL for `one`
  const uint32_t *block_map;      // device, 2 words per workgroup: trial, chunk (NULL: trial 0, chunk = block)
  const uint32_t *epoch_dev;      // if set, the launch's epoch is read here (graph replay); init = epoch == 0
  const uint32_t *rows;
  const int64_t *labels;          // classifier targets (NULL for the encoder)
  const uint8_t *codes;           // encoder targets [n][class_chunks] (NULL for the classifier)
  uint32_t n, wd;
  uint32_t row_stride;            // words between rows: >= wd, a multiple of the tile's words per lane
  uint32_t epoch_first, epochs;   // absolute index of the first epoch, count in this launch
  uint32_t class_chunks, total_blocks;
  uint32_t input_bits, state_bits;
  uint32_t encoder;
  uint32_t last_live, last_eval;  // masks of the last real word (Wd-1)
  uint32_t init;                  // 1: start from the setup state (capi.c:1136-1146)
  uint32_t lanes, wpl;            // tile: lanes per clause, words per lane (0: default shape)
  uint32_t all_ib1;               // every trial has include_bits == 1
  uint32_t all_full;              // every trial has width and row stride wd and all-ones last-word masks
  // Step queue (classifier, one machine, one epoch per launch; NULL qid: the
  // kernel scans the epoch order itself).  Chunk ch trains on entries
  // [qoff[ch * qseg], qoff[(ch + 1) * qseg]) in order.
  const uint64_t *qoff;
  const uint32_t *qid;            // sample | target << 31
  const uint32_t *qdraw;          // [8][qcap] step draws, clause j's in plane j (Philox TAG_STEP, TAG_STEP + 1)
  uint64_t qcap;
  uint32_t qseg;                  // 1024-position segments per chunk
};

// Per-epoch step queue of a classifier: which epoch positions each chunk
// trains on (target or kept negative), in epoch order, with their step draws.
struct tm_queue_launch {
  const uint32_t *perm;           // this epoch's order [n]
  const int64_t *labels;
  uint32_t n, chunks, clause_chunks, nseg, epoch;
  uint32_t ek0, ek1, seed_lo, seed_hi;
  uint64_t skip_threshold;
  uint64_t *counts;               // [chunks * nseg + 1]
  uint64_t *offsets;              // [chunks * nseg + 1], exclusive sums; the last is the total
  uint32_t *qid, *qdraw;
  uint64_t qcap;
  // per epoch position: the sample's half of its skip hash and its label as
  // uint32 [2][n], written by the count launch and read by both passes
  uint32_t *pos_keys;
};

// Work counters of a train launch: TM_NSTATS entries, as documented at
// tm_last_train_stats in tsetlin_amd.h.

struct tm_predict_launch {
  const uint32_t *actT;    // [Wd][clauses_total] transposed include plane
  const uint32_t *lit;     // [clauses_total] literal counts
  const uint32_t *rows;    // [n][Wd]
  int64_t *sums;           // [n][sums_stride]; class k of this launch adds into column class_offset + k
  uint32_t n, wd, clauses_total, clauses, classes, tolerance, last_eval;
  uint32_t sums_stride;    // 0: classes
  uint32_t class_offset;
  uint32_t pass_groups;    // set by the launcher: tile groups per L2-resident pass
};

// Record `msg` as the calling thread's tm_last_error() text; returns code.
int tm_set_error(int code, const char *msg);

// Returns a hipError_t.  `ms` receives the time of the train kernels only.
int tm_launch_train(const tm_train_launch &L, hipStream_t st);
int tm_launch_setup(uint32_t *counts, uint32_t *actions, uint64_t nclauses, uint32_t m, uint32_t wd, hipStream_t st);
int tm_launch_perm(uint64_t seed, uint32_t epoch_first, uint32_t epochs, uint32_t n, uint32_t *perm_out,
                   void *scratch, size_t scratch_bytes, hipStream_t st);
size_t tm_perm_scratch_bytes(uint32_t n);
// Queue: count + exclusive scan (offsets[chunks * nseg] = total entries), then
// the write pass into qid / qdraw (capacity qcap >= total).
uint32_t tm_queue_segments(uint32_t n);
size_t tm_queue_scratch_bytes(uint64_t items);
int tm_launch_queue_count(const tm_queue_launch &Q, void *scratch, size_t scratch_bytes, hipStream_t st);
int tm_launch_queue_write(const tm_queue_launch &Q, hipStream_t st);
int tm_launch_prepare_predict(const uint32_t *actions, uint32_t *actT, uint32_t *lit, uint32_t clauses_total,
                              uint32_t wd, uint32_t last_eval, hipStream_t st);
int tm_launch_predict(const tm_predict_launch &L, hipStream_t st);
int tm_launch_argmax(const int64_t *sums, uint32_t n, uint32_t classes, int64_t *out, hipStream_t st);
int tm_launch_encode(const int64_t *sums, uint32_t n, uint32_t classes, uint8_t *out, hipStream_t st);
int tm_launch_repack(const uint8_t *src, uint32_t n, uint32_t bytes, uint32_t stride_words, uint32_t *dst, hipStream_t st);
int tm_launch_zero(void *p, size_t bytes, hipStream_t st);
// State bits the train kernels are built for: 2 .. TM_MAX_STATE_BITS.  A tile
// is built for state bits b when b x words per lane <= 128 (the chunk state
// of a lane in VGPRs).
#define TM_MAX_STATE_BITS 16
// Train kernel tile for a row of wd words: lanes per clause and words per lane.
int tm_train_shape(uint32_t wd, uint32_t state_bits, uint32_t *lanes, uint32_t *words_per_lane);
// Every instantiated tile that covers wd words (for the autotuner); returns the count.
int tm_train_shapes(uint32_t wd, uint32_t state_bits, uint32_t (*out)[2], int cap);
int tm_launch_epoch_advance(uint32_t *epoch_dev, hipStream_t st);
// hits[t] += number of s < n with pred[s] == labels[s] (one counter per launch)
int tm_launch_count_hits(const int64_t *pred, const int64_t *labels, uint32_t n, unsigned long long *hits,
                         hipStream_t st);
// Per-class TP / FP / FN of predictions against labels (classes outside
// [0, classes) skipped, the reference's unsigned conversion) added into
// counts[3][classes] on stream st; and the host finish of
// evaluator/capi.c:156-173 over such counts, in the reference's order.
int tm_launch_class_counts(const int64_t *pred, const int64_t *expect, uint32_t n, uint32_t classes,
                           unsigned long long *counts, hipStream_t st);
void tm_finish_class_accuracy(const uint64_t *counts, uint32_t classes, double *per_class, double *macro);
