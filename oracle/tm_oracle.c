/*
 * tm_oracle.c -- CPU oracle for the Tsetlin-machine train/predict path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The shipped engine (libtsetlin_amd.so)
 * never links, loads or calls it.
 *
 * It restates, in plain C, the algorithm of the reference library
 * (lib/santoku/tsetlin/capi.c and lib/santoku/tsetlin/automata.h of
 * lua-santoku-tsetlin 0.0.274).  Every function cites the reference lines it
 * follows.  Two training front-ends exist:
 *
 *   orc_train_replay   -- the engine's own counter-based RNG stream (Philox +
 *                         lowbias32, see "RNG stream" below).  The GPU engine
 *                         must reproduce its automata state bit for bit.
 *   orc_train_refshape -- the reference's execution shape: OpenMP static
 *                         schedule over 8-clause chunks, one sequential RNG
 *                         and one inside-out Fisher-Yates shuffle per thread
 *                         and epoch, clause evaluation before the negative
 *                         sampling test.  The reference's RNG (santoku's
 *                         tk_fast_*) is not available, so each thread runs
 *                         PCG32 seeded from (seed, thread id).  This is the
 *                         timed CPU baseline and the accuracy reference.
 *
 * Parity pins for each part live in tests/test_oracle.py (brute force
 * byte-per-literal counter model, closed forms, Random123 known-answer
 * vectors for Philox4x32-10).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>
#include <math.h>
#include <omp.h>

#define ORC_CVEC_BITS 8

/* ------------------------------------------------------------------------ */
/* Machine parameters (reference: tk_tsetlin_t, capi.c:46-83)               */
/* ------------------------------------------------------------------------ */

typedef struct {
  /* user options */
  uint32_t type;              /* 0 classifier, 1 encoder */
  uint32_t classes;           /* classes, or hidden bits for the encoder */
  uint32_t features;          /* features, or visible bits */
  uint32_t clauses;           /* per class, rounded up to a multiple of 8 */
  uint32_t clause_tolerance;
  uint32_t clause_maximum;
  uint32_t state_bits;
  uint32_t include_bits;
  double target_opt;
  double negative_opt;
  double specificity;
  uint64_t seed;
  /* derived (capi.c:473-493) */
  uint32_t target;
  uint32_t input_bits;
  uint32_t input_chunks;
  uint32_t clause_chunks;
  uint32_t class_chunks;
  uint32_t state_chunks;
  uint32_t action_chunks;
  uint8_t tail_mask;
  uint8_t pad_[7];
  double negative;
} orc_params;

static inline uint32_t bits_bytes (uint64_t n) { return (uint32_t)((n + ORC_CVEC_BITS - 1) / ORC_CVEC_BITS); }

/* Derived sizes and defaults: capi.c:473-493 (classifier) and :508-526
 * (encoder = classifier with hidden/visible).  The target formula keeps the
 * reference's float fmaxf and truncating unsigned conversion (capi.c:480-483). */
void orc_derive (orc_params *p)
{
  p->negative = p->negative_opt < 0 ? 1.0 / (double) p->classes : p->negative_opt;
  p->class_chunks = bits_bytes(p->classes);
  p->clauses = bits_bytes(p->clauses) * ORC_CVEC_BITS;
  double t;
  if (p->target_opt < 0)
    t = sqrt((double) p->clauses / 2.0) * (double) p->clause_tolerance;
  else
    t = ceil(p->target_opt >= 1 ? p->target_opt : fmaxf(1.0f, (float)((double) p->clauses * p->target_opt)));
  p->target = (uint32_t) t;
  if (p->include_bits == 0) p->include_bits = 1;
  p->input_bits = 2 * p->features;
  uint32_t tail_bits = p->input_bits & (ORC_CVEC_BITS - 1);
  p->tail_mask = tail_bits ? (uint8_t)((1u << tail_bits) - 1) : 0xFF;
  p->input_chunks = bits_bytes(p->input_bits);
  p->clause_chunks = bits_bytes(p->clauses);
  p->state_chunks = p->classes * p->clauses * (p->state_bits - 1) * p->input_chunks;
  p->action_chunks = p->classes * p->clauses * p->input_chunks;
}

/* ------------------------------------------------------------------------ */
/* RNG stream (engine-defined; the device kernels compute the same values)   */
/* ------------------------------------------------------------------------ */

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11; multipliers and Weyl key increments from that paper. */
void orc_philox (const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t) 0xD2511F53u * c0;
    uint64_t p1 = (uint64_t) 0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t) p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t) p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* lowbias32 integer mixer (C. Wellons, "hash-prospector"): a bijection on
 * 32-bit words. */
uint32_t orc_lowbias32 (uint32_t x)
{
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

#define ORC_TAG_EPOCH 0x45504f43u
#define ORC_TAG_PERM  0x5045524du
#define ORC_TAG_STEP  0x53544550u
#define ORC_GOLDEN    0x9E3779B9u

static inline void seed_key (uint64_t seed, uint32_t key[2])
{
  key[0] = (uint32_t) seed;
  key[1] = (uint32_t)(seed >> 32);
}

/* Four per-epoch key words. */
void orc_epoch_keys (uint64_t seed, uint32_t epoch, uint32_t ek[4])
{
  uint32_t key[2], ctr[4] = { epoch, 0, 0, ORC_TAG_EPOCH };
  seed_key(seed, key);
  orc_philox(ctr, key, ek);
}

/* Sort key of sample position i in epoch e. */
static inline uint32_t perm_key (const uint32_t key[2], uint32_t epoch, uint32_t i)
{
  uint32_t ctr[4] = { i, epoch, 0, ORC_TAG_PERM }, out[4];
  orc_philox(ctr, key, out);
  return out[0];
}

typedef struct { uint32_t k, i; } orc_kv;
static int kv_cmp (const void *a, const void *b)
{
  const orc_kv *x = a, *y = b;
  if (x->k != y->k) return x->k < y->k ? -1 : 1;
  return x->i < y->i ? -1 : (x->i > y->i);
}

/* Epoch sample order: positions sorted by (sort key, position).  Stands in for
 * the per-thread Fisher-Yates of capi.c:245-256; one order is shared by every
 * chunk of the epoch. */
void orc_perm (uint64_t seed, uint32_t epoch, uint32_t n, uint32_t *perm)
{
  uint32_t key[2];
  seed_key(seed, key);
  orc_kv *kv = malloc((size_t) n * sizeof(orc_kv) + 1);
  for (uint32_t i = 0; i < n; i++) { kv[i].k = perm_key(key, epoch, i); kv[i].i = i; }
  qsort(kv, n, sizeof(orc_kv), kv_cmp);
  for (uint32_t i = 0; i < n; i++) perm[i] = kv[i].i;
  free(kv);
}

/* Negative-sampling draw for (sample, chunk) in an epoch (capi.c:215). */
uint32_t orc_skip_draw (const uint32_t ek[4], uint32_t sample, uint32_t chunk)
{
  return orc_lowbias32(orc_lowbias32(sample ^ ek[0]) ^ (chunk * ORC_GOLDEN + ek[1]));
}

/* The eight 32-bit feedback draws for (sample, chunk) in an epoch, d[j] for
 * clause j of the chunk (capi.c:221-239): clauses 0-3 from one Philox block,
 * 4-7 from the next counter tag. */
void orc_step_draws (uint64_t seed, uint32_t epoch, uint32_t sample, uint32_t chunk, uint32_t d[8])
{
  uint32_t key[2], ctr[4] = { sample, chunk, epoch, ORC_TAG_STEP };
  seed_key(seed, key);
  orc_philox(ctr, key, d);
  ctr[3] = ORC_TAG_STEP + 1;
  orc_philox(ctr, key, d + 4);
}

/* Key of the random-literal draws of a Type I / output 0 event on a clause. */
uint32_t orc_literal_key (const uint32_t ek[4], uint32_t sample, uint32_t clause)
{
  return orc_lowbias32(orc_lowbias32(sample ^ ek[2]) ^ (clause * ORC_GOLDEN + ek[3]));
}

/* Draw r of that event: a literal index in [0, input_bits) (capi.c:188). */
uint32_t orc_literal_draw (uint32_t kev, uint32_t r, uint32_t input_bits)
{
  uint32_t h = orc_lowbias32(kev + r * ORC_GOLDEN);
  return (uint32_t)(((uint64_t) h * input_bits) >> 32);
}

/* chance(q) threshold for 32-bit draws: draw < thr  <=>  event. */
uint64_t orc_threshold32 (double q)
{
  if (!(q > 0)) return 0;
  if (q >= 1) return (uint64_t) 1 << 32;
  return (uint64_t) floor(q * 4294967296.0);
}

/* ------------------------------------------------------------------------ */
/* Bit-plane automata (automata.h:7-31 layout, :72-232 and :499-742 ops)     */
/* ------------------------------------------------------------------------ */

static inline uint8_t *act_row (const orc_params *p, uint8_t *actions, uint64_t clause)
{
  return actions + clause * p->input_chunks;                       /* automata.h:17-22 */
}

static inline uint8_t *cnt_plane (const orc_params *p, uint8_t *counts, uint64_t clause, uint64_t plane)
{
  return counts + (clause * (p->state_bits - 1) + plane) * p->input_chunks;  /* automata.h:24-31 */
}

/* automata.h:72-86: every count plane all-ones, actions zero (value 2^m - 1). */
void orc_setup (const orc_params *p, uint8_t *counts, uint8_t *actions, uint64_t clause_first, uint64_t clause_last)
{
  uint64_t m = p->state_bits - 1;
  for (uint64_t c = clause_first; c <= clause_last; c++) {
    for (uint64_t b = 0; b < m; b++)
      memset(cnt_plane(p, counts, c, b), 0xFF, p->input_chunks);
    memset(act_row(p, actions, c), 0, p->input_chunks);
  }
}

/* Saturating +1 on the literals set in `active` within one byte; the carry into
 * the action plane is masked by tail_mask on the last byte (automata.h:88-116). */
void orc_inc_byte (const orc_params *p, uint8_t *counts, uint8_t *actions, uint64_t clause, uint64_t chunk, uint8_t active)
{
  if (!active) return;
  uint64_t m = p->state_bits - 1;
  uint8_t *a = act_row(p, actions, clause);
  uint8_t mask = (chunk == p->input_chunks - 1) ? p->tail_mask : 0xFF;
  uint8_t carry = active;
  for (uint64_t b = 0; b < m; b++) {
    uint8_t *pl = cnt_plane(p, counts, clause, b);
    uint8_t nx = pl[chunk] & carry;
    pl[chunk] ^= carry;
    carry = nx;
  }
  uint8_t cm = carry & mask;
  uint8_t nx = a[chunk] & cm;
  a[chunk] ^= cm;
  carry = nx;
  for (uint64_t b = 0; b < m; b++)
    cnt_plane(p, counts, clause, b)[chunk] |= carry;
  a[chunk] |= carry;
}

/* Saturating -1, same masking (automata.h:118-146). */
void orc_dec_byte (const orc_params *p, uint8_t *counts, uint8_t *actions, uint64_t clause, uint64_t chunk, uint8_t active)
{
  if (!active) return;
  uint64_t m = p->state_bits - 1;
  uint8_t *a = act_row(p, actions, clause);
  uint8_t mask = (chunk == p->input_chunks - 1) ? p->tail_mask : 0xFF;
  uint8_t borrow = active;
  for (uint64_t b = 0; b < m; b++) {
    uint8_t *pl = cnt_plane(p, counts, clause, b);
    uint8_t nx = (uint8_t)(~pl[chunk]) & borrow;
    pl[chunk] ^= borrow;
    borrow = nx;
  }
  uint8_t bm = borrow & mask;
  uint8_t nx = (uint8_t)(~a[chunk]) & bm;
  a[chunk] ^= bm;
  borrow = nx;
  for (uint64_t b = 0; b < m; b++)
    cnt_plane(p, counts, clause, b)[chunk] &= (uint8_t) ~borrow;
  a[chunk] &= (uint8_t) ~borrow;
}

/* "included" = actions AND the top include_bits-1 count planes (automata.h:499-512). */
static inline uint8_t included_byte (const orc_params *p, uint8_t *counts, uint8_t *actions, uint64_t clause, uint64_t chunk)
{
  uint8_t inc = act_row(p, actions, clause)[chunk];
  uint64_t m = p->state_bits - 1;
  for (uint64_t i = 1; i < p->include_bits && i <= m; i++)
    inc &= cnt_plane(p, counts, clause, m - i)[chunk];
  return inc;
}

/* The reference's 128-bit bulk loops (automata.h:164-194, :556-594, :660-698)
 * run over whole 16-byte blocks with no tail mask; the remaining bytes go
 * through the masked byte path.  bulk_end() is where the byte path starts.
 * Here each 16-byte block is processed as two 64-bit words, which is the same
 * bitwise computation. */
static inline uint64_t bulk_end (uint64_t n_chunks) { return n_chunks & ~(uint64_t) 15; }

static inline uint64_t ld64 (const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
static inline void st64 (uint8_t *p, uint64_t v) { memcpy(p, &v, 8); }

#define ORC_MAX_PLANES 64

/* One 64-bit word of +1 (carry) or -1 (borrow) ripple over the m count planes
 * and the action plane, with saturation (automata.h:164-194, :575-594). */
static inline void word_add (uint8_t **planes, uint64_t m, uint8_t *a, uint64_t k, uint64_t carry)
{
  for (uint64_t b = 0; b < m; b++) {
    uint64_t c = ld64(planes[b] + k);
    st64(planes[b] + k, c ^ carry);
    carry &= c;
  }
  uint64_t av = ld64(a + k);
  st64(a + k, av ^ carry);
  carry &= av;
  for (uint64_t b = 0; b < m; b++) st64(planes[b] + k, ld64(planes[b] + k) | carry);
  st64(a + k, ld64(a + k) | carry);
}

static inline void word_sub (uint8_t **planes, uint64_t m, uint8_t *a, uint64_t k, uint64_t borrow)
{
  for (uint64_t b = 0; b < m; b++) {
    uint64_t c = ld64(planes[b] + k);
    st64(planes[b] + k, c ^ borrow);
    borrow &= ~c;
  }
  uint64_t av = ld64(a + k);
  st64(a + k, av ^ borrow);
  borrow &= ~av;
  for (uint64_t b = 0; b < m; b++) st64(planes[b] + k, ld64(planes[b] + k) & ~borrow);
  st64(a + k, ld64(a + k) & ~borrow);
}

static inline uint64_t included_word (uint8_t **planes, uint64_t m, uint64_t ib, const uint8_t *a, uint64_t k)
{
  uint64_t inc = ld64(a + k);
  for (uint64_t i = 1; i < ib && i <= m; i++) inc &= ld64(planes[m - i] + k);
  return inc;
}

static inline void planes_of (const orc_params *p, uint8_t *counts, uint64_t clause, uint8_t **planes)
{
  for (uint64_t b = 0; b < p->state_bits - 1; b++) planes[b] = cnt_plane(p, counts, clause, b);
}

/* tk_automata_inc (automata.h:147-232): +1 on every literal set in input. */
void orc_inc (const orc_params *p, uint8_t *counts, uint8_t *actions, uint64_t clause, const uint8_t *input)
{
  uint64_t n = p->input_chunks, k = 0, be = bulk_end(n), m = p->state_bits - 1;
  uint8_t *planes[ORC_MAX_PLANES], *a = act_row(p, actions, clause);
  planes_of(p, counts, clause, planes);
  for (; k < be; k += 8) word_add(planes, m, a, k, ld64(input + k));
  for (; k < n; k++) orc_inc_byte(p, counts, actions, clause, k, input[k]);
}

/* tk_automata_dec_not_excluded (automata.h:538-640): -1 on every literal that
 * is false in input and not included. */
void orc_dec_not_excluded (const orc_params *p, uint8_t *counts, uint8_t *actions, uint64_t clause, const uint8_t *input)
{
  uint64_t n = p->input_chunks, k = 0, be = bulk_end(n), m = p->state_bits - 1;
  uint8_t *planes[ORC_MAX_PLANES], *a = act_row(p, actions, clause);
  planes_of(p, counts, clause, planes);
  for (; k < be; k += 8)
    word_sub(planes, m, a, k, ~ld64(input + k) & ~included_word(planes, m, p->include_bits, a, k));
  for (; k < n; k++) {
    uint8_t active = (uint8_t)(~input[k] & ~included_byte(p, counts, actions, clause, k));
    orc_dec_byte(p, counts, actions, clause, k, active);
  }
}

/* tk_automata_inc_not_excluded (automata.h:642-742): +1 on every literal that
 * is false in input and not included. */
void orc_inc_not_excluded (const orc_params *p, uint8_t *counts, uint8_t *actions, uint64_t clause, const uint8_t *input)
{
  uint64_t n = p->input_chunks, k = 0, be = bulk_end(n), m = p->state_bits - 1;
  uint8_t *planes[ORC_MAX_PLANES], *a = act_row(p, actions, clause);
  planes_of(p, counts, clause, planes);
  for (; k < be; k += 8)
    word_add(planes, m, a, k, ~ld64(input + k) & ~included_word(planes, m, p->include_bits, a, k));
  for (; k < n; k++) {
    uint8_t active = (uint8_t)(~input[k] & ~included_byte(p, counts, actions, clause, k));
    orc_inc_byte(p, counts, actions, clause, k, active);
  }
}

/* ------------------------------------------------------------------------ */
/* Clause evaluation, chunk sums, predict (capi.c:85-161, :903-995)          */
/* ------------------------------------------------------------------------ */

/* popcount(a) and popcount(a & ~x) over `nbytes` whole bytes: the semantics of
 * santoku-matrix's tk_cvec_bits_popcount_andnot_serial as used at capi.c:103. */
static inline void popcount_andnot (const uint8_t *a, const uint8_t *x, uint64_t nbytes, uint64_t *lit, uint64_t *fail)
{
  uint64_t l = 0, f = 0, k = 0;
  for (; k + 8 <= nbytes; k += 8) {
    uint64_t av, xv;
    memcpy(&av, a + k, 8);
    memcpy(&xv, x + k, 8);
    l += (uint64_t) __builtin_popcountll(av);
    f += (uint64_t) __builtin_popcountll(av & ~xv);
  }
  for (; k < nbytes; k++) {
    l += (uint64_t) __builtin_popcount(a[k]);
    f += (uint64_t) __builtin_popcount((uint8_t)(a[k] & ~x[k]));
  }
  *lit = l;
  *fail = f;
}

/* tk_tsetlin_calculate (capi.c:85-146): literals, votes and the 8 output bits
 * of one chunk.  Empty clauses vote +tolerance. */
uint8_t orc_calculate (const orc_params *p, const uint8_t *actions, const uint8_t *input, uint32_t chunk, uint32_t *literals, uint32_t *votes)
{
  uint8_t out = 0;
  const uint32_t B = p->input_chunks, tol = p->clause_tolerance;
  for (uint32_t j = 0; j < ORC_CVEC_BITS; j++) {
    const uint8_t *a = actions + ((uint64_t) chunk * ORC_CVEC_BITS + j) * B;
    uint64_t l64 = 0, f64 = 0;
    if (B > 1) popcount_andnot(a, input, B - 1, &l64, &f64);
    uint8_t la = a[B - 1] & p->tail_mask;
    uint32_t lit = (uint32_t) l64 + (uint32_t) __builtin_popcount(la);
    uint32_t fail = (uint32_t) f64 + (uint32_t) __builtin_popcount((uint8_t)(la & ~input[B - 1]));
    long v;
    if (lit == 0) v = (long) tol;
    else {
      v = (long)(lit < tol ? lit : tol) - (long) fail;
      if (v < 0) v = 0;
    }
    if (v > 0) out |= (uint8_t)(1u << j);
    literals[j] = lit;
    votes[j] = (uint32_t) v;
  }
  return out;
}

/* tk_tsetlin_sums (capi.c:148-161). */
long orc_sums (uint8_t out, const uint32_t *votes)
{
  long s = 0;
  for (uint32_t j = 0; j < ORC_CVEC_BITS; j += 2) if (out & (1u << j)) s += (long) votes[j];
  for (uint32_t j = 1; j < ORC_CVEC_BITS; j += 2) if (out & (1u << j)) s -= (long) votes[j];
  return s;
}

/* Per-sample class sums (the quantity both predict paths reduce). */
void orc_class_sums (const orc_params *p, const uint8_t *actions, const uint8_t *rows, uint32_t n, int64_t *sums)
{
  const uint32_t total = p->clause_chunks * p->classes;
  #pragma omp parallel for schedule(static)
  for (uint32_t s = 0; s < n; s++) {
    uint32_t lit[8], vot[8];
    int64_t *row = sums + (uint64_t) s * p->classes;
    for (uint32_t c = 0; c < p->classes; c++) row[c] = 0;
    const uint8_t *x = rows + (uint64_t) s * p->input_chunks;
    for (uint32_t ch = 0; ch < total; ch++) {
      uint8_t out = orc_calculate(p, actions, x, ch, lit, vot);
      row[ch / p->clause_chunks] += orc_sums(out, vot);
    }
  }
}

/* tk_tsetlin_predict_classifier (capi.c:903-948): argmax with strict '>' so
 * ties go to the lowest class. */
void orc_predict_classifier (const orc_params *p, const uint8_t *actions, const uint8_t *rows, uint32_t n, int64_t *out)
{
  int64_t *sums = malloc((size_t) n * p->classes * sizeof(int64_t) + 8);
  orc_class_sums(p, actions, rows, n, sums);
  for (uint32_t s = 0; s < n; s++) {
    int64_t best = -INT64_MAX;
    uint32_t bc = 0;
    for (uint32_t c = 0; c < p->classes; c++) {
      if (sums[(uint64_t) s * p->classes + c] > best) { best = sums[(uint64_t) s * p->classes + c]; bc = c; }
    }
    out[s] = bc;
  }
  free(sums);
}

/* tk_tsetlin_predict_encoder (capi.c:950-995): bit h = (sum_h > 0); padding
 * bits of the last code byte are written as zero. */
void orc_predict_encoder (const orc_params *p, const uint8_t *actions, const uint8_t *rows, uint32_t n, uint8_t *out)
{
  int64_t *sums = malloc((size_t) n * p->classes * sizeof(int64_t) + 8);
  orc_class_sums(p, actions, rows, n, sums);
  for (uint32_t s = 0; s < n; s++) {
    uint8_t *e = out + (uint64_t) s * p->class_chunks;
    memset(e, 0, p->class_chunks);
    for (uint32_t c = 0; c < p->classes; c++)
      if (sums[(uint64_t) s * p->classes + c] > 0) e[c / 8] |= (uint8_t)(1u << (c % 8));
  }
  free(sums);
}

/* ------------------------------------------------------------------------ */
/* Training, replay front-end (capi.c:163-243 feedback, :1112-1282 loops)    */
/* ------------------------------------------------------------------------ */

/* Work counters of a replay run: [0] chunk-steps evaluated, [1] Type I with
 * output 1 events, [2] Type II with output 1 events, [3] Type I with output 0
 * events, [4] random-literal draws, [5] negative-sampling skips. */
#define ORC_NSTATS 6

/* One chunk-step.  `chunk` indexes the state arrays; `chunk_r` is the chunk's
 * index in the random stream (equal for a plain machine; for an
 * individualized one, class h's chunk c is chunk h * clause_chunks + c of the
 * stream while its state is class h's own array). */
static void replay_step (const orc_params *p, uint8_t *counts, uint8_t *actions, const uint8_t *x,
                         uint32_t sample, uint32_t chunk, uint32_t chunk_r, uint32_t epoch, bool tgt,
                         const uint32_t ek[4], uint32_t s_draws, uint64_t *st)
{
  uint32_t lit[8], vot[8];
  uint8_t out = orc_calculate(p, actions, x, chunk, lit, vot);     /* capi.c:1162 */
  long T = (long) p->target;
  long v = orc_sums(out, vot);                                      /* capi.c:217-220 */
  if (v > T) v = T;
  if (v < -T) v = -T;
  uint64_t num = (uint64_t)(tgt ? T - v : T + v);                   /* capi.c:221-225 */
  uint64_t den = 2 * (uint64_t) T;
  uint32_t d[8];
  orc_step_draws(p->seed, epoch, sample, chunk_r, d);
  st[0]++;
  for (uint32_t j = 0; j < 8; j++) {                                /* capi.c:226-242 */
    /* chance(num / den): u / 2^32 < num / den, exactly in integers */
    if (!((uint64_t) d[j] * den < (num << 32))) continue;
    bool type1 = tgt ? ((j & 1) == 0) : ((j & 1) == 1);
    uint64_t clause = (uint64_t) chunk * 8 + j;
    bool output = vot[j] > 0;                                       /* capi.c:176 */
    if (type1) {
      if (output) {                                                 /* capi.c:179-183 */
        if (lit[j] <= p->clause_maximum)
          orc_inc(p, counts, actions, clause, x);
        orc_dec_not_excluded(p, counts, actions, clause, x);
        st[1]++;
      } else {                                                      /* capi.c:184-192 */
        uint32_t kev = orc_literal_key(ek, sample, chunk_r * 8 + j);
        for (uint32_t r = 0; r < s_draws; r++) {
          uint32_t bit = orc_literal_draw(kev, r, p->input_bits);
          orc_dec_byte(p, counts, actions, clause, bit / 8, (uint8_t)(1u << (bit % 8)));
        }
        st[3]++;
        st[4] += s_draws;
      }
    } else if (output) {                                            /* capi.c:194-197 */
      orc_inc_not_excluded(p, counts, actions, clause, x);
      st[2]++;
    }
  }
}

/* Number of random decrements of a Type I / output 0 event (capi.c:185). */
uint32_t orc_s_draws (const orc_params *p)
{
  unsigned int twof = 2 * p->features;
  return (uint32_t)((double) twof / p->specificity);
}

/* Train for `epochs` epochs from the setup state (capi.c:1136-1146 runs setup
 * on every train call).  Classifier: labels[n] (int64).  Encoder: codes
 * [n][class_chunks].  Chunks are independent (each one's feedback depends only
 * on its own 8 clauses), so they run in parallel here in any order.
 * epoch_first lets a caller replay epochs [epoch_first, epoch_first+epochs)
 * on an existing state (no setup when epoch_first > 0). */
/* The replay of the chunks in `chunk_list` (all chunks when NULL); the other
 * chunks' state is left as it is. */
static void train_replay_list (const orc_params *p, uint8_t *counts, uint8_t *actions,
                               uint32_t n, const uint8_t *rows, const int64_t *labels, const uint8_t *codes,
                               uint32_t epoch_first, uint32_t epochs, const uint32_t *chunk_list,
                               uint32_t n_list, uint64_t *stats)
{
  const uint32_t total = chunk_list ? n_list : p->clause_chunks * p->classes;
  if (epoch_first == 0)
    for (uint32_t i = 0; i < total; i++) {
      uint32_t ch = chunk_list ? chunk_list[i] : i;
      orc_setup(p, counts, actions, (uint64_t) ch * 8, (uint64_t) ch * 8 + 7);
    }
  uint32_t *perm = malloc((size_t) n * sizeof(uint32_t) + 4);
  uint64_t thr = orc_threshold32(1.0 - p->negative);
  uint32_t s_draws = orc_s_draws(p);
  uint64_t st[ORC_NSTATS] = { 0 };
  for (uint32_t e = epoch_first; e < epoch_first + epochs; e++) {
    orc_perm(p->seed, e, n, perm);
    uint32_t ek[4];
    orc_epoch_keys(p->seed, e, ek);
    #pragma omp parallel
    {
      uint64_t lst[ORC_NSTATS] = { 0 };
      #pragma omp for schedule(dynamic, 1)
      for (uint32_t li = 0; li < total; li++) {
        uint32_t ch = chunk_list ? chunk_list[li] : li;
        uint32_t cls = ch / p->clause_chunks;
        for (uint32_t i = 0; i < n; i++) {
          uint32_t s = perm[i];
          const uint8_t *x = rows + (uint64_t) s * p->input_chunks;
          bool tgt;
          if (p->type == 0) {
            uint32_t sc = (uint32_t) labels[s];
            tgt = sc == cls;
            if (!tgt && (uint64_t) orc_skip_draw(ek, s, ch) < thr) { lst[5]++; continue; }
          } else {
            tgt = (codes[(uint64_t) s * p->class_chunks + cls / 8] >> (cls % 8)) & 1;
          }
          replay_step(p, counts, actions, x, s, ch, ch, e, tgt, ek, s_draws, lst);
        }
      }
      #pragma omp critical
      for (int k = 0; k < ORC_NSTATS; k++) st[k] += lst[k];
    }
  }
  if (stats) for (int k = 0; k < ORC_NSTATS; k++) stats[k] = st[k];
  free(perm);
}

void orc_train_replay (const orc_params *p, uint8_t *counts, uint8_t *actions,
                       uint32_t n, const uint8_t *rows, const int64_t *labels, const uint8_t *codes,
                       uint32_t epoch_first, uint32_t epochs, uint64_t *stats)
{
  train_replay_list(p, counts, actions, n, rows, labels, codes, epoch_first, epochs, NULL, 0, stats);
}

/* Replay of some chunks only.  Chunks are independent (capi.c:201-243: a
 * chunk's feedback reads its own 8 votes), so each listed chunk ends exactly
 * as in a replay of the whole machine, at a cost that does not grow with the
 * number of classes and clauses. */
void orc_train_replay_chunks (const orc_params *p, uint8_t *counts, uint8_t *actions,
                              uint32_t n, const uint8_t *rows, const int64_t *labels, const uint8_t *codes,
                              uint32_t epoch_first, uint32_t epochs, const uint32_t *chunk_list,
                              uint32_t n_list, uint64_t *stats)
{
  train_replay_list(p, counts, actions, n, rows, labels, codes, epoch_first, epochs, chunk_list, n_list, stats);
}

/* ------------------------------------------------------------------------ */
/* Training, reference-shaped front-end (timed CPU baseline)                 */
/* ------------------------------------------------------------------------ */

/* PCG32 (M. O'Neill, "PCG: A family of simple fast space-efficient
 * statistically good algorithms for random number generation", 2014). */
typedef struct { uint64_t state, inc; } orc_pcg;
static inline uint32_t pcg32 (orc_pcg *g)
{
  uint64_t old = g->state;
  g->state = old * 6364136223846793005ULL + g->inc;
  uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((-rot) & 31));
}
static inline void pcg_seed (orc_pcg *g, uint64_t seed, uint64_t stream)
{
  g->state = 0; g->inc = (stream << 1u) | 1u;
  pcg32(g); g->state += seed; pcg32(g);
}
static inline bool pcg_chance (orc_pcg *g, double p) { return (double) pcg32(g) * (1.0 / 4294967296.0) < p; }

/* capi.c:201-243 with a sequential per-thread stream: the negative-sampling
 * test happens after evaluation (capi.c:1162-1164, :215-216), and each pair
 * draws twice in the reference's order. */
static void refshape_step (const orc_params *p, uint8_t *counts, uint8_t *actions, const uint8_t *x,
                           uint32_t chunk, uint32_t sample_class, bool target_vote, uint32_t chunk_class,
                           orc_pcg *g, uint32_t s_draws)
{
  uint32_t lit[8], vot[8];
  uint8_t out = orc_calculate(p, actions, x, chunk, lit, vot);
  if (chunk_class != sample_class && pcg_chance(g, 1 - p->negative))
    return;
  long T = (long) p->target;
  long v = orc_sums(out, vot);
  if (v > T) v = T;
  if (v < -T) v = -T;
  bool tgt = chunk_class == sample_class && target_vote;
  double prob = tgt ? (double)(T - v) / (2.0 * T) : (double)(T + v) / (2.0 * T);
  for (uint32_t j = 0; j < 8; j += 2) {
    uint32_t first = tgt ? j : j + 1, second = tgt ? j + 1 : j;
    for (int w = 0; w < 2; w++) {
      uint32_t cj = w == 0 ? first : second;
      if (!pcg_chance(g, prob)) continue;
      uint64_t clause = (uint64_t) chunk * 8 + cj;
      bool output = vot[cj] > 0;
      if (w == 0) {
        if (output) {
          if (lit[cj] <= p->clause_maximum) orc_inc(p, counts, actions, clause, x);
          orc_dec_not_excluded(p, counts, actions, clause, x);
        } else {
          for (uint32_t r = 0; r < s_draws; r++) {
            uint32_t bit = pcg32(g) % p->input_bits;
            orc_dec_byte(p, counts, actions, clause, bit / 8, (uint8_t)(1u << (bit % 8)));
          }
        }
      } else if (output) {
        orc_inc_not_excluded(p, counts, actions, clause, x);
      }
    }
  }
}

/* capi.c:1112-1282 with OpenMP static scheduling over chunks and a per-thread
 * inside-out Fisher-Yates each epoch (capi.c:245-256, :1155).  sample_limit
 * caps the samples visited per epoch (0 = all) so a bounded slice of a large
 * workload can be timed; the shuffle still runs over all n. */
void orc_train_refshape (const orc_params *p, uint8_t *counts, uint8_t *actions,
                         uint32_t n, const uint8_t *rows, const int64_t *labels, const uint8_t *codes,
                         uint32_t epochs, int nthreads, uint32_t sample_limit)
{
  const uint32_t total = p->clause_chunks * p->classes;
  const uint32_t visit = (sample_limit && sample_limit < n) ? sample_limit : n;
  uint32_t s_draws = orc_s_draws(p);
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  uint32_t **shuf = malloc((size_t) nthreads * sizeof(uint32_t *));
  for (int t = 0; t < nthreads; t++) shuf[t] = malloc((size_t) n * sizeof(uint32_t) + 4);
  orc_pcg *gens = malloc((size_t) nthreads * sizeof(orc_pcg));
  #pragma omp parallel num_threads(nthreads)
  {
    int tid = omp_get_thread_num();
    pcg_seed(&gens[tid], p->seed, (uint64_t) tid);
    #pragma omp for schedule(static)
    for (uint32_t ch = 0; ch < total; ch++)
      orc_setup(p, counts, actions, (uint64_t) ch * 8, (uint64_t) ch * 8 + 7);
  }
  for (uint32_t e = 0; e < epochs; e++) {
    #pragma omp parallel num_threads(nthreads)
    {
      int tid = omp_get_thread_num();
      orc_pcg *g = &gens[tid];
      uint32_t *sh = shuf[tid];
      for (uint32_t i = 0; i < n; i++) {
        sh[i] = i;
        uint32_t j = i == 0 ? 0 : pcg32(g) % (i + 1);
        uint32_t t = sh[i]; sh[i] = sh[j]; sh[j] = t;
      }
      #pragma omp for schedule(static)
      for (uint32_t ch = 0; ch < total; ch++) {
        uint32_t cls = ch / p->clause_chunks;
        for (uint32_t i = 0; i < visit; i++) {
          uint32_t s = sh[i];
          const uint8_t *x = rows + (uint64_t) s * p->input_chunks;
          if (p->type == 0)
            refshape_step(p, counts, actions, x, ch, (uint32_t) labels[s], true, cls, g, s_draws);
          else {
            bool tv = (codes[(uint64_t) s * p->class_chunks + cls / 8] >> (cls % 8)) & 1;
            refshape_step(p, counts, actions, x, ch, cls, tv, cls, g, s_draws);
          }
        }
      }
    }
  }
  for (int t = 0; t < nthreads; t++) free(shuf[t]);
  free(shuf);
  free(gens);
}

/* Predict with an explicit thread count (CPU baseline timing, capi.c:920). */
void orc_predict_classifier_threads (const orc_params *p, const uint8_t *actions, const uint8_t *rows, uint32_t n, int64_t *out, int nthreads)
{
  int prev = omp_get_max_threads();
  if (nthreads > 0) omp_set_num_threads(nthreads);
  orc_predict_classifier(p, actions, rows, n, out);
  omp_set_num_threads(prev);
}

void orc_predict_encoder_threads (const orc_params *p, const uint8_t *actions, const uint8_t *rows, uint32_t n, uint8_t *out, int nthreads)
{
  int prev = omp_get_max_threads();
  if (nthreads > 0) omp_set_num_threads(nthreads);
  orc_predict_encoder(p, actions, rows, n, out);
  omp_set_num_threads(prev);
}

/* ------------------------------------------------------------------------ */
/* Individualized machines (capi.c:258-402, :528-626, :997-1086, :1284-1482) */
/* ------------------------------------------------------------------------ */

/* Class h of an individualized machine is a one-class machine over its own
 * feat_sizes[h] features (row width, tail mask and random-draw range of its
 * own, capi.c:576-604) sharing clauses, tolerance, maximum, target,
 * specificity, negative sampling and seed with the others.  Its counts and
 * actions sit one class after another, in class order (capi.c:585-588). */
void orc_ind_class (const orc_params *p, uint64_t k_h, orc_params *out)
{
  *out = *p;
  out->classes = 1;
  out->features = (uint32_t) k_h;
  orc_derive(out);
  out->target = p->target;        /* target and negative are the machine's (capi.c:555-566) */
  out->negative = p->negative;
  out->class_chunks = p->class_chunks;
}

static void ind_offsets (const orc_params *p, const uint64_t *feat_sizes, uint32_t h, uint64_t *coff, uint64_t *aoff)
{
  uint64_t c = 0, a = 0;
  for (uint32_t i = 0; i < h; i++) {
    uint64_t b = bits_bytes(2 * feat_sizes[i]);
    c += (uint64_t) p->clauses * b * (p->state_bits - 1);
    a += (uint64_t) p->clauses * b;
  }
  *coff = c;
  *aoff = a;
}

/* Training (capi.c:1284-1482) with the engine's stream: one epoch order for
 * every chunk; class h's chunk c draws as chunk h * clause_chunks + c.
 * Sample s of class h reads problems + dim_offsets[h] + s * input_chunks_h.
 * Classifier: negative sampling as the plain machine; encoder: every sample
 * (capi.c:1354 passes 0.0). */
void orc_train_replay_ind (const orc_params *p, const uint64_t *feat_sizes, uint8_t *counts, uint8_t *actions,
                           uint32_t n, const uint8_t *problems, const int64_t *dim_offsets,
                           const int64_t *labels, const uint8_t *codes, uint32_t epoch_first, uint32_t epochs,
                           uint64_t *stats)
{
  const uint32_t K = p->classes, cc = p->clause_chunks;
  if (epoch_first == 0)
    for (uint32_t h = 0; h < K; h++) {
      orc_params ph;
      orc_ind_class(p, feat_sizes[h], &ph);
      uint64_t co, ao;
      ind_offsets(p, feat_sizes, h, &co, &ao);
      orc_setup(&ph, counts + co, actions + ao, 0, (uint64_t) cc * 8 - 1);
    }
  uint32_t *perm = malloc((size_t) n * sizeof(uint32_t) + 4);
  uint64_t thr = orc_threshold32(1.0 - p->negative);
  uint64_t st[ORC_NSTATS] = { 0 };
  for (uint32_t e = epoch_first; e < epoch_first + epochs; e++) {
    orc_perm(p->seed, e, n, perm);
    uint32_t ek[4];
    orc_epoch_keys(p->seed, e, ek);
    #pragma omp parallel
    {
      uint64_t lst[ORC_NSTATS] = { 0 };
      #pragma omp for schedule(dynamic, 1)
      for (uint32_t chg = 0; chg < K * cc; chg++) {
        const uint32_t h = chg / cc, c = chg % cc;
        orc_params ph;
        orc_ind_class(p, feat_sizes[h], &ph);
        uint64_t co, ao;
        ind_offsets(p, feat_sizes, h, &co, &ao);
        const uint32_t s_draws = orc_s_draws(&ph);
        for (uint32_t i = 0; i < n; i++) {
          uint32_t s = perm[i];
          const uint8_t *x = problems + dim_offsets[h] + (uint64_t) s * ph.input_chunks;
          bool tgt;
          if (p->type == 0) {
            tgt = (uint32_t) labels[s] == h;
            if (!tgt && (uint64_t) orc_skip_draw(ek, s, chg) < thr) { lst[5]++; continue; }
          } else {
            tgt = (codes[(uint64_t) s * p->class_chunks + h / 8] >> (h % 8)) & 1;
          }
          replay_step(&ph, counts + co, actions + ao, x, s, c, chg, e, tgt, ek, s_draws, lst);
        }
      }
      #pragma omp critical
      for (int k = 0; k < ORC_NSTATS; k++) st[k] += lst[k];
    }
  }
  if (stats) for (int k = 0; k < ORC_NSTATS; k++) stats[k] = st[k];
  free(perm);
}

/* Class sums of an individualized machine (capi.c:1010-1029, :1057-1073):
 * class h evaluates its own input row. */
void orc_class_sums_ind (const orc_params *p, const uint64_t *feat_sizes, const uint8_t *actions,
                         const uint8_t *problems, const int64_t *dim_offsets, uint32_t n, int64_t *sums)
{
  const uint32_t K = p->classes, cc = p->clause_chunks;
  #pragma omp parallel for schedule(static)
  for (uint32_t s = 0; s < n; s++) {
    uint32_t lit[8], vot[8];
    for (uint32_t h = 0; h < K; h++) {
      orc_params ph;
      orc_ind_class(p, feat_sizes[h], &ph);
      uint64_t co, ao;
      ind_offsets(p, feat_sizes, h, &co, &ao);
      const uint8_t *x = problems + dim_offsets[h] + (uint64_t) s * ph.input_chunks;
      int64_t sum = 0;
      for (uint32_t c = 0; c < cc; c++) {
        uint8_t out = orc_calculate(&ph, actions + ao, x, c, lit, vot);
        sum += orc_sums(out, vot);
      }
      sums[(uint64_t) s * K + h] = sum;
    }
  }
}

/* tk_tsetlin_predict_classifier_ind / _encoder_ind (capi.c:997-1086): argmax
 * (strict '>', ties to the lowest class) or bit h = sum_h > 0. */
void orc_predict_ind (const orc_params *p, const uint64_t *feat_sizes, const uint8_t *actions,
                      const uint8_t *problems, const int64_t *dim_offsets, uint32_t n, int64_t *out_class,
                      uint8_t *out_codes)
{
  const uint32_t K = p->classes;
  int64_t *sums = malloc((size_t) n * K * sizeof(int64_t) + 8);
  orc_class_sums_ind(p, feat_sizes, actions, problems, dim_offsets, n, sums);
  for (uint32_t s = 0; s < n; s++) {
    const int64_t *row = sums + (uint64_t) s * K;
    if (out_class) {
      int64_t best = -INT64_MAX;
      uint32_t bc = 0;
      for (uint32_t h = 0; h < K; h++) if (row[h] > best) { best = row[h]; bc = h; }
      out_class[s] = bc;
    }
    if (out_codes) {
      uint8_t *e = out_codes + (uint64_t) s * p->class_chunks;
      memset(e, 0, p->class_chunks);
      for (uint32_t h = 0; h < K; h++) if (row[h] > 0) e[h / 8] |= (uint8_t)(1u << (h % 8));
    }
  }
  free(sums);
}

int orc_max_threads (void) { return omp_get_max_threads(); }
