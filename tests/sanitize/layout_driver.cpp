// CPU sanitizer driver for the host's layout and size arithmetic
// (lua-santoku-tsetlin_amd/csrc/tm_layout.h, the same header tm_api.cpp and
// tm_kernels.hip use), compiled under AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitize.py.  It sweeps and fuzzes
// every function, checks the invariants the kernels rely on, and exits 0 when
// nothing was reported and every check held.
#include "tm_layout.h"

#include <stdio.h>
#include <stdlib.h>
#include <vector>

using namespace tm_layout;

static int failures = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) { fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); failures++; } \
  } while (0)

static uint64_t rs = 88172645463325252ull;
static uint64_t rnd() { rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17; return rs; }

int main()
{
  // tiles: a default tile covers the row and fits its state; candidates too
  for (uint32_t sb = 0; sb <= 18; sb++)
    for (uint32_t wd = 1; wd <= 5000; wd += (wd < 300 ? 1 : 37)) {
      uint32_t g = 0, w = 0;
      if (train_shape(wd, sb, &g, &w) == 0) {
        CHECK(sb >= 2 && sb <= kMaxStateBits && (uint64_t) g * w >= wd && tile_fits(w, sb));
      }
      uint32_t out[16][2];
      const int n = train_shapes(wd, sb, out, 16);
      CHECK(n >= 0 && n <= 16);
      for (int i = 0; i < n; i++) CHECK((uint64_t) out[i][0] * out[i][1] >= wd && tile_fits(out[i][1], sb));
    }
  // masks: the evaluated and live bits of the last word lie inside the row
  for (uint32_t bits = 2; bits <= 4200; bits += 2) {
    sizes z;
    CHECK(row_sizes(bits / 2, &z));
    const uint32_t em = eval_mask(z.wd, z.input_bits), lm = live_mask(z.wd, z.input_chunks, z.tail_mask);
    const uint32_t last_bits = z.input_bits - 32 * (z.wd - 1);
    CHECK(__builtin_popcount(em) == (int) (last_bits > 32 ? 32 : last_bits));
    CHECK((lm & ~em) == 0 || z.input_chunks % 16 == 0);   // only the bulk region's padding evolves
  }
  sizes z;
  CHECK(!row_sizes(0, &z) && !row_sizes(0x80000000ull, &z) && !row_sizes(0xFFFFFFFFull, &z));
  CHECK(row_sizes(0x7FFFFFFFull, &z) && z.input_bits == 0xFFFFFFFEu && z.input_chunks == 0x20000000u);
  // target: finite, and refused above the limit by the caller's comparison
  const double tf[] = { -1.0, 1e-9, 0.001, 0.25, 0.999, 1.0, 3.5, 1e6, 1e300 };
  const uint32_t cl[] = { 8, 64, 4096, 1u << 20, 0xFFFFFFF8u };
  const uint32_t tol[] = { 1, 64, kMaxTolerance, 0xFFFFFFFFu };
  for (double t : tf)
    for (uint32_t c : cl)
      for (uint32_t o : tol) {
        const double v = target_value(t, c, o);
        CHECK(v >= 1.0 || t < 0);
        if (v <= kMaxTarget) CHECK((uint32_t) v <= (1u << 30));
      }
  // persisted state sizes
  CHECK(state_fits(1, 8, 1, 2) && !state_fits(0, 8, 1, 2) && !state_fits(1, 8, 1, 1));
  CHECK(!state_fits(0xFFFFFFFFull, 0xFFFFFFFFull, 0xFFFFFFFFull, 16));
  CHECK(!state_fits(1u << 16, 1u << 16, 2, 8) && state_fits(1u << 8, 1u << 8, 1024, 8));
  // header: every truncation and random bytes
  std::vector<uint8_t> buf(kHeaderBytes);
  for (int it = 0; it < 200000; it++) {
    for (auto &b : buf) b = (uint8_t) rnd();
    if (it % 3 == 0) { int32_t ty = (int32_t) (rnd() % 2); memcpy(buf.data(), &ty, 4); }
    const size_t len = rnd() % (kHeaderBytes + 2);
    header h;
    const char *e = read_header(len ? buf.data() : nullptr, len > kHeaderBytes ? kHeaderBytes : len, &h);
    if (e) continue;
    if (!check_common(h)) CHECK(h.clauses % 8 == 0 && h.clause_chunks * 8 == h.clauses && h.classes > 0);
    if (!check_common(h) && !check_plain(h))
      CHECK((uint64_t) h.action_chunks == (uint64_t) h.classes * h.clauses * h.input_chunks);
  }
  // individualized layout: offsets follow each other, totals fit, bad classes named
  for (int it = 0; it < 20000; it++) {
    const size_t K = 1 + rnd() % 9;
    std::vector<ind_class> c(K);
    for (auto &x : c) {
      const uint64_t r = rnd() % 100;
      x.features = r < 3 ? 0 : r < 6 ? (rnd() % 0xFFFFFFFFull) : 1 + rnd() % 3000;
    }
    const uint32_t clauses = 8 * (1 + (uint32_t) (rnd() % 512)), sb = 2 + (uint32_t) (rnd() % 15);
    ind_totals t;
    const char *why = "";
    const size_t bad = ind_layout(c.data(), K, clauses, sb, &t, &why);
    if (bad) { CHECK(bad <= K + 1 && why[0]); continue; }
    uint64_t aw = 0, hb = 0;
    for (const auto &x : c) {
      CHECK(x.features > 0 && x.aoff == aw && x.hoff == hb && x.coff == aw * (sb - 1));
      aw += (uint64_t) clauses * x.wd;
      hb += (uint64_t) clauses * x.input_chunks;
    }
    CHECK(t.action_words == aw && t.action_chunks == hb && (uint64_t) t.state_chunks == hb * (sb - 1));
  }
  printf("layout checks: %d failures\n", failures);
  return failures ? 1 : 0;
}
