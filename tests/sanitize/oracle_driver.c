/* CPU sanitizer driver for the oracle (test infrastructure): compiled together
 * with oracle/tm_oracle.c under AddressSanitizer and UndefinedBehaviorSanitizer
 * by tests/test_sanitize.py.  It runs every front-end of the oracle on small
 * machines, including row tails in both padding regimes, state bits 2, 3, 8
 * and 16, include_bits 3, labels outside the classes and individualized
 * machines, and exits 0 when nothing was reported. */
#include "../../oracle/tm_oracle.c"

#include <stdio.h>

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd (void)
{
  rng_state ^= rng_state << 13; rng_state ^= rng_state >> 7; rng_state ^= rng_state << 17;
  return (uint32_t) (rng_state >> 11);
}

static void fill_rows (uint8_t *rows, uint32_t n, uint32_t features, uint32_t bytes)
{
  memset(rows, 0, (size_t) n * bytes);
  for (uint32_t s = 0; s < n; s++)
    for (uint32_t f = 0; f < features; f++) {
      const uint32_t b = (rnd() % 100) < 30 ? f : features + f;   /* x, or not x */
      rows[(size_t) s * bytes + b / 8] |= (uint8_t) (1u << (b % 8));
    }
}

static uint64_t checksum (const uint8_t *p, size_t n)
{
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; i++) h = (h ^ p[i]) * 1099511628211ull;
  return h;
}

static void plain (uint32_t type, uint32_t features, uint32_t classes, uint32_t clauses, uint32_t state, uint32_t ib)
{
  orc_params p;
  memset(&p, 0, sizeof p);
  p.type = type; p.classes = classes; p.features = features; p.clauses = clauses;
  p.clause_tolerance = 5; p.clause_maximum = 6; p.state_bits = state; p.include_bits = ib;
  p.target_opt = 6; p.negative_opt = -1; p.specificity = 1.5; p.seed = 3;
  orc_derive(&p);
  const uint32_t n = 40, B = p.input_chunks;
  uint8_t *rows = malloc((size_t) n * B);
  fill_rows(rows, n, features, B);
  int64_t *labels = malloc(n * sizeof(int64_t));
  uint8_t *codes = malloc((size_t) n * p.class_chunks);
  for (uint32_t s = 0; s < n; s++) labels[s] = s % 11 == 0 ? -1 : (int64_t) (rnd() % (classes + 1));
  for (size_t i = 0; i < (size_t) n * p.class_chunks; i++) codes[i] = (uint8_t) rnd();
  uint8_t *counts = malloc(p.state_chunks), *actions = malloc(p.action_chunks);
  uint64_t st[6];
  orc_train_replay(&p, counts, actions, n, rows, type ? NULL : labels, type ? codes : NULL, 0, 2, st);
  uint32_t list[2] = { 0, p.clause_chunks * classes - 1 };
  orc_train_replay_chunks(&p, counts, actions, n, rows, type ? NULL : labels, type ? codes : NULL, 2, 1, list, 2, st);
  orc_train_refshape(&p, counts, actions, n, rows, type ? NULL : labels, type ? codes : NULL, 2, 2, 0);
  orc_train_refshape(&p, counts, actions, n, rows, type ? NULL : labels, type ? codes : NULL, 1, 3, 7);
  int64_t *sums = malloc((size_t) n * classes * sizeof(int64_t)), *pred = malloc(n * sizeof(int64_t));
  uint8_t *enc = malloc((size_t) n * p.class_chunks);
  orc_class_sums(&p, actions, rows, n, sums);
  if (type == 0) orc_predict_classifier_threads(&p, actions, rows, n, pred, 2);
  else orc_predict_encoder_threads(&p, actions, rows, n, enc, 2);
  printf("plain type %u F %u state %u ib %u: %016llx\n", type, features, state, ib,
         (unsigned long long) (checksum(actions, p.action_chunks) ^ checksum(counts, p.state_chunks)));
  free(rows); free(labels); free(codes); free(counts); free(actions); free(sums); free(pred); free(enc);
}

static void individualized (uint32_t type)
{
  const uint64_t fs[4] = { 5, 64, 3, 100 };
  orc_params p;
  memset(&p, 0, sizeof p);
  p.type = type; p.classes = 4; p.features = 1; p.clauses = 16; p.clause_tolerance = 4; p.clause_maximum = 5;
  p.state_bits = 8; p.include_bits = 1; p.target_opt = 5; p.negative_opt = -1; p.specificity = 2.0; p.seed = 9;
  orc_derive(&p);
  const uint32_t n = 30;
  uint64_t acts = 0, bytes = 0;
  int64_t offs[4];
  for (int h = 0; h < 4; h++) {
    const uint64_t b = (2 * fs[h] + 7) / 8;
    acts += p.clauses * b;
    offs[h] = (int64_t) bytes;
    bytes += n * b + 3;
  }
  uint8_t *prob = malloc(bytes), *actions = malloc(acts), *counts = malloc(acts * 7);
  for (uint64_t i = 0; i < bytes; i++) prob[i] = (uint8_t) rnd();
  int64_t labels[30];
  uint8_t codes[30];
  for (uint32_t s = 0; s < n; s++) { labels[s] = rnd() % 4; codes[s] = (uint8_t) rnd(); }
  uint64_t st[6];
  orc_train_replay_ind(&p, fs, counts, actions, n, prob, offs, type ? NULL : labels, type ? codes : NULL, 0, 2, st);
  int64_t sums[30 * 4], pred[30];
  uint8_t enc[30];
  orc_class_sums_ind(&p, fs, actions, prob, offs, n, sums);
  orc_predict_ind(&p, fs, actions, prob, offs, n, type ? NULL : pred, type ? enc : NULL);
  printf("individualized type %u: %016llx\n", type, (unsigned long long) checksum(actions, acts));
  free(prob); free(actions); free(counts);
}

int main (void)
{
  const uint32_t F[] = { 5, 63, 64, 100, 300 };
  for (uint32_t t = 0; t < 2; t++)
    for (size_t i = 0; i < sizeof F / sizeof F[0]; i++) {
      plain(t, F[i], 3, 16, 8, 1);
      plain(t, F[i], 2, 8, i % 2 ? 3 : 2, 1);
    }
  plain(0, 13, 3, 8, 16, 3);
  plain(1, 9, 9, 24, 5, 3);
  individualized(0);
  individualized(1);
  return 0;
}
