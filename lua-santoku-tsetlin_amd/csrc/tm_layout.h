// tm_layout.h -- the host side's pure layout and size arithmetic.
//
// Plain C++ with no HIP dependency: tm_api.cpp and tm_kernels.hip use it, and
// a CPU sanitizer build (AddressSanitizer + UndefinedBehaviorSanitizer, under
// tests/sanitize/) exercises the same code.  Nothing here touches a device.
#pragma once
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

namespace tm_layout {

// ---------------------------------------------------------------------------
// train tiles

// Built tiles of the train kernel: lanes per clause x words per lane.
constexpr uint32_t kShapes[][2] = { { 8, 1 }, { 8, 2 }, { 16, 1 }, { 16, 2 }, { 32, 1 }, { 32, 2 },
                                    { 64, 1 }, { 64, 2 }, { 64, 4 }, { 64, 8 }, { 64, 16 } };
constexpr uint32_t kMaxStateBits = 16;

// A tile is built for a plane count when the chunk state of one lane,
// state bits x words per lane registers, fits in 128 VGPRs.
inline bool tile_fits(uint32_t wpl, uint32_t state_bits) { return (uint64_t) state_bits * wpl <= 128; }

// Default tile: the smallest lanes x words that covers the row, one clause
// group per wave once a row needs 64 lanes.  0, or -1 when none is built.
inline int train_shape(uint32_t wd, uint32_t state_bits, uint32_t *lanes, uint32_t *wpl)
{
  static const uint32_t order[][2] = { { 8, 1 }, { 16, 1 }, { 32, 1 }, { 64, 1 }, { 64, 2 }, { 64, 4 }, { 64, 8 }, { 64, 16 } };
  if (state_bits < 2 || state_bits > kMaxStateBits) return -1;
  for (auto &s : order)
    if ((uint64_t) s[0] * s[1] >= wd && tile_fits(s[1], state_bits)) { *lanes = s[0]; *wpl = s[1]; return 0; }
  return -1;
}

// Every built tile that covers wd words and is at most about four times the
// row (the autotuner's candidates); returns the count written (<= cap).
inline int train_shapes(uint32_t wd, uint32_t state_bits, uint32_t (*out)[2], int cap)
{
  int n = 0;
  if (state_bits < 2 || state_bits > kMaxStateBits) return 0;
  for (auto &s : kShapes) {
    const uint64_t words = (uint64_t) s[0] * s[1];
    if (words >= wd && words < 4ull * wd + 8 && tile_fits(s[1], state_bits) && n < cap) {
      out[n][0] = s[0];
      out[n][1] = s[1];
      n++;
    }
  }
  return n;
}

// ---------------------------------------------------------------------------
// word masks of a literal row held as wd u32 words

// Literal bits < input_bits inside the last word of a row of wd words.
inline uint32_t eval_mask(uint32_t wd, uint32_t input_bits)
{
  const uint64_t bits = (uint64_t) input_bits - 32ull * (wd - 1);
  return bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1);
}

// Bits of the last word whose automata evolve.  Bytes < input_chunks are live.
// The last byte keeps only its tail bits when the reference's masked byte path
// handles it (input_chunks % 16 != 0, automata.h:97,106); inside the unmasked
// 128-bit bulk region all 8 bits of the byte evolve (automata.h:164-194).
inline uint32_t live_mask(uint32_t wd, uint32_t input_chunks, uint8_t tail_mask)
{
  uint32_t m = 0;
  for (uint32_t b = 0; b < 4; b++) {
    const uint64_t byte = 4ull * (wd - 1) + b;
    if (byte >= input_chunks) break;
    uint8_t bm = 0xFF;
    if (byte == input_chunks - 1ull && (input_chunks % 16) != 0) bm = tail_mask;
    m |= (uint32_t) bm << (8 * b);
  }
  return m;
}

// ---------------------------------------------------------------------------
// options and derived sizes (capi.c:473-493)

// Widths of the kernels' integer arithmetic: a vote is at most the tolerance
// and a predict tile sums 64 of them in 32 bits; the train kernel's fire test
// computes T +- v and 2T in 32 bits.
constexpr uint32_t kMaxTolerance = (1u << 25) - 1;
constexpr double kMaxTarget = (double) ((1u << 30) - 1);

// The target as the reference derives it (float fmaxf, ceil; capi.c:480-483),
// as a double.  The caller refuses values above kMaxTarget before converting.
inline double target_value(double targetf, uint32_t clauses, uint32_t tolerance)
{
  if (targetf < 0) return sqrt((double) clauses / 2.0) * (double) tolerance;
  return ceil(targetf >= 1 ? targetf : fmaxf(1.0f, (float)((double) clauses * targetf)));
}

struct sizes {
  uint32_t input_bits, input_chunks, wd;
  uint8_t tail_mask;
};

// Literal row of 2 * features bits (capi.c:487-490): bytes, words, tail mask.
// false when the row does not fit the 32-bit size fields.
inline bool row_sizes(uint64_t features, sizes *out)
{
  if (features == 0 || features > 0x7FFFFFFFull) return false;
  out->input_bits = (uint32_t)(2 * features);
  // in 64 bits: input_bits + 7 wraps 32 bits from 2^31 - 3 features on
  out->input_chunks = (uint32_t)(((uint64_t) out->input_bits + 7) / 8);
  const uint32_t tail_bits = out->input_bits & 7;
  out->tail_mask = tail_bits ? (uint8_t)((1u << tail_bits) - 1) : 0xFF;
  out->wd = (uint32_t)(((uint64_t) out->input_chunks + 3) / 4);
  return true;
}

// The persisted state sizes of a machine (actions = classes x clauses x
// input_chunks bytes, counts state_bits - 1 times that) fit the persisted
// 32-bit fields (capi.c:1511-1525).
inline bool state_fits(uint64_t classes, uint64_t clauses, uint64_t input_chunks, uint64_t state_bits)
{
  if (classes == 0 || clauses == 0 || input_chunks == 0 || state_bits < 2) return false;
  if (classes > 0xFFFFFFFFull / clauses) return false;
  const uint64_t cc = classes * clauses;
  if (cc > 0xFFFFFFFFull / input_chunks) return false;
  const uint64_t act = cc * input_chunks;
  return act * (state_bits - 1) <= 0xFFFFFFFFull;
}

// ---------------------------------------------------------------------------
// persisted layout (capi.c:1509-1531 written, :1675-1737 read)

// int32 type, 13 u32 fields, f64 specificity, u8 tail mask.
constexpr size_t kHeaderBytes = 4 + 13 * 4 + 8 + 1;

struct header {
  int32_t type;
  uint32_t classes, features, clauses, clause_tolerance, clause_maximum, target, state_bits, include_bits,
      input_bits, input_chunks, clause_chunks, state_chunks, action_chunks;
  double specificity;
  uint8_t tail_mask;
};

// Reads the header at the front of `data`.  0, or a message on failure (the
// bytes are shorter than a header, or the type is neither 0 nor 1).
inline const char *read_header(const uint8_t *data, size_t len, header *h)
{
  if (!data || len < 4) return "load: truncated data";
  memcpy(&h->type, data, 4);
  if (h->type != 0 && h->type != 1) return "unexpected tsetlin machine type in load";
  if (len < kHeaderBytes) return "load: truncated header";
  uint32_t f[13];
  memcpy(f, data + 4, sizeof f);
  h->classes = f[0]; h->features = f[1]; h->clauses = f[2]; h->clause_tolerance = f[3];
  h->clause_maximum = f[4]; h->target = f[5]; h->state_bits = f[6]; h->include_bits = f[7];
  h->input_bits = f[8]; h->input_chunks = f[9]; h->clause_chunks = f[10]; h->state_chunks = f[11];
  h->action_chunks = f[12];
  memcpy(&h->specificity, data + 4 + sizeof f, 8);
  h->tail_mask = data[kHeaderBytes - 1];
  return nullptr;
}

// Consistency of the fields every machine has: clauses a positive multiple
// of 8 and clause_chunks its eighth.  0 or a message.
inline const char *check_common(const header &h)
{
  if (!h.classes || !h.clauses || h.clauses % 8 || h.clause_chunks != h.clauses / 8) return "load: inconsistent header";
  return nullptr;
}

// The plain machine's row and state fields agree with each other.
inline const char *check_plain(const header &h)
{
  if (h.input_bits == 0 || h.input_chunks != (uint32_t)(((uint64_t) h.input_bits + 7) / 8) ||
      (uint64_t) h.action_chunks != (uint64_t) h.classes * h.clauses * h.input_chunks)
    return "load: inconsistent header";
  return nullptr;
}

// ---------------------------------------------------------------------------
// individualized machines (capi.c:576-593): class h has its own row width and
// its state follows the classes before it

struct ind_class {
  uint64_t features = 0;                  // k_h
  uint32_t input_bits = 0, input_chunks = 0, wd = 0;
  uint8_t tail_mask = 0xFF;
  uint64_t aoff = 0, coff = 0;            // u32-word offsets of its actions and counts
  uint64_t hoff = 0;                      // byte offset of its actions in the persisted layout
};

struct ind_totals {
  uint64_t action_words;                  // device words of the include plane
  uint32_t action_chunks, state_chunks;   // persisted sizes
  uint32_t wd_max;                        // the widest class's words per row
};

// Fills each class's sizes and offsets from its feature count.  0, or the index
// of the first bad class plus one with *why set: "empty" (k_h = 0), "width"
// (too many features), "size" (a class's state does not fit the 32-bit
// fields); or classes + 1 with why "total" when the machine's totals do not.
inline size_t ind_layout(ind_class *cls, size_t classes, uint32_t clauses, uint32_t state_bits, ind_totals *t,
                         const char **why)
{
  uint64_t aw = 0, hb = 0;
  uint32_t wmax = 1;
  for (size_t h = 0; h < classes; h++) {
    ind_class &c = cls[h];
    sizes z;
    if (c.features == 0) { *why = "empty"; return h + 1; }
    if (!row_sizes(c.features, &z)) { *why = "width"; return h + 1; }
    if (!state_fits(1, clauses, z.input_chunks, state_bits)) { *why = "size"; return h + 1; }
    c.input_bits = z.input_bits;
    c.input_chunks = z.input_chunks;
    c.tail_mask = z.tail_mask;
    c.wd = z.wd;
    c.aoff = aw;
    c.coff = aw * (state_bits - 1);
    c.hoff = hb;
    aw += (uint64_t) clauses * c.wd;
    hb += (uint64_t) clauses * c.input_chunks;
    if (c.wd > wmax) wmax = c.wd;
    if (hb > 0xFFFFFFFFull) { *why = "total"; return classes + 1; }
  }
  if (hb * (state_bits - 1) > 0xFFFFFFFFull) { *why = "total"; return classes + 1; }
  t->action_words = aw;
  t->action_chunks = (uint32_t) hb;
  t->state_chunks = (uint32_t)(hb * (state_bits - 1));
  t->wd_max = wmax;
  return 0;
}

} // namespace tm_layout
