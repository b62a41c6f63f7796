// tm_api.cpp -- host side of the C ABI declared in include/tsetlin_amd.h.
//
// Mirrors the machine object of lib/santoku/tsetlin/capi.c (tk_tsetlin_t,
// capi.c:46-83): option checks and derived sizes, the persisted byte layout,
// checkpoint/restore/reconfigure semantics.  All automata state lives in device
// memory; the kernels are in tm_kernels.hip.
#include "tsetlin_amd.h"
#include "tm_internal.h"
#include "tm_layout.h"
#include "tm_rng.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(expr)                                                                \
  do {                                                                               \
    hipError_t e_ = (hipError_t)(expr);                                              \
    if (e_ != hipSuccess) return fail(TM_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

struct dbuf {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes)
  {
    if (bytes <= cap && p) return TM_OK;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = bytes ? bytes : 4;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(TM_E_ALLOC, "device allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
    }
    cap = want;
    return TM_OK;
  }
  void release()
  {
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <typename T> T *as() const { return (T *) p; }
};

using tm_layout::eval_mask;
using tm_layout::live_mask;

// A trial whose rows are wd words apart, wd words long, with no bit of the
// last word masked: the train kernel's variant for such rows keeps no masks.
uint32_t full_rows(const tm_train_trial &Q, uint32_t wd)
{
  return Q.wd == wd && Q.row_stride == wd && Q.last_live == 0xFFFFFFFFu && Q.last_eval == 0xFFFFFFFFu;
}

} // namespace

int tm_set_error(int code, const char *msg)
{
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

// One class of an individualized machine: a one-class machine over its own
// feature set (capi.c:576-604), stored after the classes before it.
using tm_ind_class = tm_layout::ind_class;

struct tm_machine {
  int type = TM_CLASSIFIER;
  bool has_state = false, trained = false, reusable = false, individualized = false;
  uint32_t classes = 0, features = 0, clauses = 0, clause_tolerance = 0, clause_maximum = 0, target = 0;
  uint32_t state_bits = 8, include_bits = 1, input_bits = 0, input_chunks = 0, clause_chunks = 0;
  uint32_t class_chunks = 0, state_chunks = 0, action_chunks = 0;
  uint8_t tail_mask = 0xFF;
  double negative = 0, specificity = 0;
  uint64_t seed = 0;
  std::vector<tm_ind_class> ind;    // individualized: one entry per class

  int device = 0;
  hipStream_t stream = nullptr;
  bool owns_stream = true;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t wd = 0;                  // u32 words per literal row (individualized: the widest class)
  dbuf counts, actions;             // automata state
  dbuf actT, lit;                   // predict-side views of the include plane
  bool predict_ready = false;
  dbuf t_rows, t_words, t_labels, t_codes, t_perm, t_perm_scratch, t_ekeys;   // training data
  dbuf t_trials, t_map;                                                       // individualized train table
  dbuf p_rows, p_words, p_sums, p_out;                                        // predict data
  dbuf t_stats;
  dbuf q_counts, q_offsets, q_scratch, q_id, q_draw, q_keys;                  // step queue
  uint64_t q_cap = 0;
  hipEvent_t evq0 = nullptr, evq1 = nullptr;
  uint64_t stats[TM_NSTATS] = { 0 };
  double train_ms = 0, predict_ms = 0, queue_ms = 0;

  uint64_t nclauses() const { return (uint64_t) classes * clauses; }
  // u32 words of the include plane (the counts hold state_bits - 1 times as many)
  uint64_t action_words() const
  {
    if (!individualized) return nclauses() * wd;
    uint64_t w = 0;
    for (const tm_ind_class &c : ind) w += (uint64_t) clauses * c.wd;
    return w;
  }
  uint32_t last_eval_mask() const { return eval_mask(wd, input_bits); }
  uint32_t last_live_mask() const { return live_mask(wd, input_chunks, tail_mask); }
};

namespace {

int check_handle(const tm_machine *m)
{
  if (!m) return fail(TM_E_ARG, "machine is NULL");
  return TM_OK;
}

int use_device(tm_machine *m)
{
  HIP_TRY(hipSetDevice(m->device));
  return TM_OK;
}

// The kernels' integer widths (tm_layout.h).  The reference computes in long
// int and double (capi.c:148-161, :217-225) and has no such limits; class
// sums here are 64-bit as there, so only the tolerance and target remain.
using tm_layout::kMaxTolerance;
using tm_layout::kMaxTarget;

// capi.c:473-493: derived sizes and defaults.  The target keeps the reference's
// float fmaxf and truncating conversion to unsigned int (capi.c:480-483).
// Targets above kMaxTarget are refused before the conversion.
int derive_target(double targetf, uint32_t clauses, uint32_t tolerance, const char *what, uint32_t *out)
{
  const double t = tm_layout::target_value(targetf, clauses, tolerance);
  if (!(t <= kMaxTarget))
    return fail(TM_E_UNSUPPORTED, "%s: target %.17g exceeds %.0f, the train kernel's 32-bit feedback arithmetic",
                what, t, kMaxTarget);
  *out = (uint32_t) t;
  return TM_OK;
}

int check_tolerance(uint32_t tolerance, const char *what)
{
  if (tolerance > kMaxTolerance)
    return fail(TM_E_UNSUPPORTED, "%s: clause_tolerance %u exceeds %u, the kernels' 32-bit vote sums", what,
                tolerance, kMaxTolerance);
  return TM_OK;
}

// capi.c:463-472 (plain) and :542-553 (individualized: no features option;
// the feature sets come from feat_offsets).
int validate_opts(const tm_options *o, const char *what, bool ind = false)
{
  if (!o) return fail(TM_E_ARG, "%s: options are NULL", what);
  if (!o->classes) return fail(TM_E_ARG, "%s: classes must be greater than 1", what);
  if (!o->clauses) return fail(TM_E_ARG, "%s: clauses must be greater than 0", what);
  // rounding up to whole bytes (capi.c:477) must not wrap 32 bits
  if (o->classes > 0x7FFFFFFFu || o->clauses > 0x7FFFFFF8u)
    return fail(TM_E_ARG, "%s: classes or clauses too large for the 32-bit size fields", what);
  if (!o->clause_tolerance) return fail(TM_E_ARG, "%s: clause_tolerance must be greater than 0", what);
  if (!o->clause_maximum) return fail(TM_E_ARG, "%s: clause_maximum must be greater than 0", what);
  if (o->state < 2) return fail(TM_E_ARG, "%s: bits must be greater than 1", what);
  if (!(o->specificity > 0)) return fail(TM_E_ARG, "%s: specificity must be a positive number", what);
  if (!ind && !o->features) return fail(TM_E_ARG, "%s: features must be greater than 0", what);
  if (!ind && o->individualized)
    return fail(TM_E_ARG, "%s: individualized machines are created by the _ind entry points (feat_offsets)", what);
  if (o->state > TM_MAX_STATE_BITS)
    return fail(TM_E_UNSUPPORTED, "%s: state bits above %d are not built in this engine", what, TM_MAX_STATE_BITS);
  return check_tolerance(o->clause_tolerance, what);
}

// Size limits of a machine; `train` adds the train tile's (a loaded machine
// only predicts).
int size_check(uint64_t classes, uint64_t clauses, uint64_t input_chunks, uint64_t state_bits, bool train = true)
{
  if (!tm_layout::state_fits(classes, clauses, input_chunks, state_bits))
    return fail(TM_E_ARG, "machine too large for the persisted 32-bit size fields");
  uint32_t lanes, wpl;
  if (train && tm_train_shape((uint32_t)((input_chunks + 3) / 4), (uint32_t) state_bits, &lanes, &wpl))
    return fail(TM_E_UNSUPPORTED, "rows of %llu bytes exceed the largest train tile for %llu state bits (%u bytes)",
                (unsigned long long) input_chunks, (unsigned long long) state_bits,
                state_bits * 16 <= 128 ? 4096u : 2048u);
  return TM_OK;
}

// The predict-side views of the include plane (transposed plane, literal
// counts) are allocated with the machine, so predict itself need not allocate
// them.
int alloc_predict_views(tm_machine *m)
{
  int rc;
  if ((rc = m->actT.ensure(m->action_words() * 4))) return rc;
  if ((rc = m->lit.ensure(m->nclauses() * 4))) return rc;
  m->predict_ready = false;
  return TM_OK;
}

int alloc_device(tm_machine *m)
{
  int rc;
  const uint64_t aw = m->action_words();
  if ((rc = m->actions.ensure(aw * 4))) return rc;
  if ((rc = m->counts.ensure(aw * (m->state_bits - 1) * 4))) return rc;
  if ((rc = alloc_predict_views(m))) return rc;
  HIP_TRY(tm_launch_zero(m->actions.p, aw * 4, m->stream));
  m->has_state = true;
  return TM_OK;
}

// Per-class sizes and offsets of an individualized machine from its feature
// sizes and clause count (capi.c:576-593): class h takes clauses x
// ceil(2 k_h / 8) bytes of actions, after the classes before it.  On the
// device each row is padded to whole u32 words.  `train` adds the train
// tile's limits (a loaded machine only predicts).
int ind_layout(tm_machine *m, const char *what, bool train = true)
{
  tm_layout::ind_totals t;
  const char *why = "";
  const size_t bad = tm_layout::ind_layout(m->ind.data(), m->ind.size(), m->clauses, m->state_bits, &t, &why);
  if (bad) {
    if (bad > m->ind.size()) return fail(TM_E_ARG, "%s: machine too large for the persisted 32-bit size fields", what);
    if (!strcmp(why, "empty")) return fail(TM_E_ARG, "%s: class %zu has an empty feature set", what, bad - 1);
    if (!strcmp(why, "width")) return fail(TM_E_ARG, "%s: class %zu has too many features", what, bad - 1);
    return fail(TM_E_ARG, "%s: class %zu: machine too large for the persisted 32-bit size fields", what, bad - 1);
  }
  if (train)
    for (const tm_ind_class &c : m->ind) {
      int rc = size_check(1, m->clauses, c.input_chunks, m->state_bits, true);
      if (rc) return rc;
    }
  m->action_chunks = t.action_chunks;
  m->state_chunks = t.state_chunks;
  m->wd = t.wd_max;
  return TM_OK;
}

int init_runtime(tm_machine *m, hipStream_t shared = nullptr)
{
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(TM_E_DEVICE, "no HIP device available");
  HIP_TRY(hipGetDevice(&m->device));
  if (shared) {
    m->stream = shared;
    m->owns_stream = false;
  } else {
    HIP_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  }
  HIP_TRY(hipEventCreate(&m->ev0));
  HIP_TRY(hipEventCreate(&m->ev1));
  HIP_TRY(hipEventCreate(&m->evq0));
  HIP_TRY(hipEventCreate(&m->evq1));
  return TM_OK;
}

// tk_tsetlin_init_classifier, capi.c:449-506.
int create_common(const tm_options *o, int type, tm_machine **out, hipStream_t shared = nullptr)
{
  const char *what = type == TM_CLASSIFIER ? "create classifier" : "create encoder";
  if (!out) return fail(TM_E_ARG, "%s: out is NULL", what);
  *out = nullptr;
  int rc = validate_opts(o, what);
  if (rc) return rc;
  tm_machine *m = new tm_machine();
  m->type = type;
  m->negative = o->negative < 0 ? 1.0 / (double) o->classes : o->negative;
  m->classes = o->classes;
  m->class_chunks = (o->classes + 7) / 8;
  m->clauses = ((o->clauses + 7) / 8) * 8;
  m->clause_tolerance = o->clause_tolerance;
  m->clause_maximum = o->clause_maximum;
  if ((rc = derive_target(o->target, m->clauses, o->clause_tolerance, what, &m->target))) { delete m; return rc; }
  m->features = o->features;
  m->state_bits = o->state;
  m->include_bits = o->include_bits ? o->include_bits : 1;
  tm_layout::sizes z;
  if (!tm_layout::row_sizes(o->features, &z)) { delete m; return fail(TM_E_ARG, "%s: too many features", what); }
  m->input_bits = z.input_bits;
  m->tail_mask = z.tail_mask;
  m->input_chunks = z.input_chunks;
  m->clause_chunks = m->clauses / 8;
  if ((rc = size_check(m->classes, m->clauses, m->input_chunks, m->state_bits))) { delete m; return rc; }
  m->state_chunks = m->classes * m->clauses * (m->state_bits - 1) * m->input_chunks;
  m->action_chunks = m->classes * m->clauses * m->input_chunks;
  m->specificity = o->specificity;
  m->reusable = o->reusable != 0;
  m->seed = o->seed;
  m->wd = z.wd;
  if ((rc = init_runtime(m, shared)) || (rc = alloc_device(m))) { tm_destroy(m); return rc; }
  *out = m;
  return TM_OK;
}

// tk_tsetlin_init_classifier_ind / _init_encoder_ind, capi.c:528-626 (create
// dispatch capi.c:631-645, :668-682; reusable is read after init, :659).
// Class h has feat_offsets[h + 1] - feat_offsets[h] features.  The machine-wide
// features, input_bits, input_chunks and tail_mask stay 0 (the reference
// never sets them for an individualized machine).
int create_ind(const tm_options *o, const int64_t *feat_offsets, uint32_t n_offsets, int type, tm_machine **out)
{
  const char *what = type == TM_CLASSIFIER ? "create classifier ind" : "create encoder ind";
  if (!out) return fail(TM_E_ARG, "%s: out is NULL", what);
  *out = nullptr;
  int rc = validate_opts(o, what, true);
  if (rc) return rc;
  if (!feat_offsets || (uint64_t) n_offsets != (uint64_t) o->classes + 1)
    return fail(TM_E_ARG, "%s: feat_offsets must have classes + 1 elements", what);
  for (uint32_t h = 0; h < o->classes; h++)
    if (feat_offsets[h + 1] < feat_offsets[h]) return fail(TM_E_ARG, "%s: feat_offsets must not decrease", what);
  tm_machine *m = new tm_machine();
  m->type = type;
  m->individualized = true;
  m->negative = o->negative < 0 ? 1.0 / (double) o->classes : o->negative;
  m->classes = o->classes;
  m->class_chunks = (o->classes + 7) / 8;
  m->clauses = ((o->clauses + 7) / 8) * 8;
  m->clause_chunks = m->clauses / 8;
  m->clause_tolerance = o->clause_tolerance;
  m->clause_maximum = o->clause_maximum;
  if ((rc = derive_target(o->target, m->clauses, o->clause_tolerance, what, &m->target))) { delete m; return rc; }
  m->state_bits = o->state;
  m->include_bits = o->include_bits ? o->include_bits : 1;
  m->specificity = o->specificity;
  m->reusable = o->reusable != 0;
  m->seed = o->seed;
  m->features = 0;
  m->input_bits = 0;
  m->input_chunks = 0;
  m->tail_mask = 0;
  m->ind.resize(m->classes);
  for (uint32_t h = 0; h < m->classes; h++) m->ind[h].features = (uint64_t)(feat_offsets[h + 1] - feat_offsets[h]);
  if ((rc = ind_layout(m, what))) { delete m; return rc; }
  if ((rc = init_runtime(m)) || (rc = alloc_device(m))) { tm_destroy(m); return rc; }
  *out = m;
  return TM_OK;
}

// Bytes of an individualized problems buffer that a call reads: class h's
// rows are the n rows of input_chunks_h bytes from byte dim_offsets[h]
// (capi.c:1348-1350 train, :1022-1024 predict).
int ind_extent(const tm_machine *m, uint32_t n, const int64_t *dim_offsets, size_t bytes, const char *what,
               size_t *extent)
{
  if (!dim_offsets) return fail(TM_E_ARG, "%s: dim_offsets must not be NULL", what);
  uint64_t e = 0;
  for (size_t h = 0; h < m->ind.size(); h++) {
    if (dim_offsets[h] < 0) return fail(TM_E_ARG, "%s: dim_offsets[%zu] is negative", what, h);
    e = std::max<uint64_t>(e, (uint64_t) dim_offsets[h] + (uint64_t) n * m->ind[h].input_chunks);
  }
  if (e > bytes)
    return fail(TM_E_SIZE, "%s: problems hold %zu bytes, dim_offsets and n need %llu", what, bytes,
                (unsigned long long) e);
  *extent = (size_t) e;
  return TM_OK;
}

// Stage an individualized problems buffer (host, or device with dev) as
// per-class u32 rows: class h's n rows start at word woff[h], strides[h]
// words apart.
int stage_rows_ind(tm_machine *m, const uint8_t *rows, uint32_t n, size_t extent, const int64_t *dim_offsets,
                   const std::vector<uint32_t> &strides, dbuf &bytes_buf, dbuf &words_buf, std::vector<uint64_t> &woff,
                   bool dev)
{
  int rc;
  const uint8_t *src = rows;
  if (!dev) {
    if ((rc = bytes_buf.ensure(extent))) return rc;
    HIP_TRY(hipMemcpyAsync(bytes_buf.p, rows, extent, hipMemcpyHostToDevice, m->stream));
    src = bytes_buf.as<uint8_t>();
  }
  woff.assign(m->ind.size(), 0);
  uint64_t total = 0;
  for (size_t h = 0; h < m->ind.size(); h++) { woff[h] = total; total += (uint64_t) n * strides[h]; }
  if ((rc = words_buf.ensure(total * 4))) return rc;
  for (size_t h = 0; h < m->ind.size(); h++)
    HIP_TRY(tm_launch_repack(src + dim_offsets[h], n, m->ind[h].input_chunks, strides[h],
                             words_buf.as<uint32_t>() + woff[h], m->stream));
  return TM_OK;
}

// Stage host or device rows (n x input_chunks bytes) as u32 words [n][stride]
// (stride >= wd; words past the row are zero).  Rows are used in place when
// the caller's bytes already have that layout and 16-byte alignment.
int stage_rows(tm_machine *m, const uint8_t *rows, uint32_t n, bool dev, uint32_t stride, dbuf &bytes_buf,
               dbuf &words_buf, const uint32_t **words)
{
  const size_t B = m->input_chunks, nbytes = (size_t) n * B;
  const bool in_place = (B % 4 == 0) && (stride == m->wd);
  if (dev) {
    if (in_place && ((uintptr_t) rows % 16 == 0)) { *words = (const uint32_t *) rows; return TM_OK; }
    int rc = words_buf.ensure((size_t) n * stride * 4);
    if (rc) return rc;
    HIP_TRY(tm_launch_repack(rows, n, (uint32_t) B, stride, words_buf.as<uint32_t>(), m->stream));
    *words = words_buf.as<uint32_t>();
    return TM_OK;
  }
  int rc = bytes_buf.ensure(nbytes);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(bytes_buf.p, rows, nbytes, hipMemcpyHostToDevice, m->stream));
  if (in_place) { *words = bytes_buf.as<uint32_t>(); return TM_OK; }
  if ((rc = words_buf.ensure((size_t) n * stride * 4))) return rc;
  HIP_TRY(tm_launch_repack(bytes_buf.as<uint8_t>(), n, (uint32_t) B, stride, words_buf.as<uint32_t>(), m->stream));
  *words = words_buf.as<uint32_t>();
  return TM_OK;
}

// Train tile (lanes per clause x words per lane): the default for the row
// width, or TSETLIN_AMD_TILE=<lanes>x<words> when that tile is built and
// covers the row (for tuning; every tile gives the same machine).
void train_tile(const tm_machine *m, uint32_t *lanes, uint32_t *wpl)
{
  *lanes = 64;
  *wpl = 1;
  tm_train_shape(m->wd, m->state_bits, lanes, wpl);
  const char *v = getenv("TSETLIN_AMD_TILE");
  unsigned g = 0, w = 0;
  if (v && sscanf(v, "%ux%u", &g, &w) == 2) {
    uint32_t shapes[16][2];
    const int n = tm_train_shapes(m->wd, m->state_bits, shapes, 16);
    for (int i = 0; i < n; i++)
      if (shapes[i][0] == g && shapes[i][1] == w) { *lanes = g; *wpl = w; }
  }
}

uint32_t train_row_stride(const tm_machine *m)
{
  uint32_t lanes, wpl;
  train_tile(m, &lanes, &wpl);
  return (m->wd + wpl - 1) / wpl * wpl;
}

int prepare_predict(tm_machine *m)
{
  if (m->predict_ready) return TM_OK;
  int rc;
  if ((rc = m->actT.ensure(m->action_words() * 4))) return rc;
  if ((rc = m->lit.ensure(m->nclauses() * 4))) return rc;
  if (m->individualized) {
    for (size_t h = 0; h < m->ind.size(); h++) {
      const tm_ind_class &c = m->ind[h];
      HIP_TRY(tm_launch_prepare_predict(m->actions.as<uint32_t>() + c.aoff, m->actT.as<uint32_t>() + c.aoff,
                                        m->lit.as<uint32_t>() + h * m->clauses, m->clauses, c.wd,
                                        eval_mask(c.wd, c.input_bits), m->stream));
    }
  } else {
    HIP_TRY(tm_launch_prepare_predict(m->actions.as<uint32_t>(), m->actT.as<uint32_t>(), m->lit.as<uint32_t>(),
                                      (uint32_t) m->nclauses(), m->wd, m->last_eval_mask(), m->stream));
  }
  m->predict_ready = true;
  return TM_OK;
}

// Class sums of an individualized machine (capi.c:1010-1029, :1057-1073): one
// predict launch per class over that class's own rows, each adding into its
// column of the [n][classes] sums.
int predict_sums_ind(tm_machine *m, const uint8_t *rows, size_t bytes, const int64_t *dim_offsets, uint32_t n,
                     bool dev)
{
  int rc;
  size_t extent;
  if ((rc = ind_extent(m, n, dim_offsets, bytes, "predict", &extent))) return rc;
  if ((rc = use_device(m))) return rc;
  if ((rc = prepare_predict(m))) return rc;
  std::vector<uint32_t> strides(m->classes);
  for (uint32_t h = 0; h < m->classes; h++) strides[h] = m->ind[h].wd;
  std::vector<uint64_t> woff;
  if ((rc = stage_rows_ind(m, rows, n, extent, dim_offsets, strides, m->p_rows, m->p_words, woff, dev))) return rc;
  if ((rc = m->p_sums.ensure((size_t) n * m->classes * 8))) return rc;
  HIP_TRY(tm_launch_zero(m->p_sums.p, (size_t) n * m->classes * 8, m->stream));
  HIP_TRY(hipEventRecord(m->ev0, m->stream));
  for (uint32_t h = 0; h < m->classes; h++) {
    const tm_ind_class &c = m->ind[h];
    tm_predict_launch L;
    memset(&L, 0, sizeof L);
    L.actT = m->actT.as<uint32_t>() + c.aoff;
    L.lit = m->lit.as<uint32_t>() + (size_t) h * m->clauses;
    L.rows = m->p_words.as<uint32_t>() + woff[h];
    L.sums = m->p_sums.as<int64_t>();
    L.n = n;
    L.wd = c.wd;
    L.clauses_total = m->clauses;
    L.clauses = m->clauses;
    L.classes = 1;
    L.tolerance = m->clause_tolerance;
    L.last_eval = eval_mask(c.wd, c.input_bits);
    L.sums_stride = m->classes;
    L.class_offset = h;
    HIP_TRY(tm_launch_predict(L, m->stream));
  }
  HIP_TRY(hipEventRecord(m->ev1, m->stream));
  return TM_OK;
}

int predict_sums(tm_machine *m, const uint8_t *rows, uint32_t n, bool dev)
{
  int rc;
  if ((rc = use_device(m))) return rc;
  if ((rc = prepare_predict(m))) return rc;
  const uint32_t *words;
  if ((rc = stage_rows(m, rows, n, dev, m->wd, m->p_rows, m->p_words, &words))) return rc;
  if ((rc = m->p_sums.ensure((size_t) n * m->classes * 8))) return rc;
  HIP_TRY(tm_launch_zero(m->p_sums.p, (size_t) n * m->classes * 8, m->stream));
  tm_predict_launch L;
  memset(&L, 0, sizeof L);
  L.actT = m->actT.as<uint32_t>();
  L.lit = m->lit.as<uint32_t>();
  L.rows = words;
  L.sums = m->p_sums.as<int64_t>();
  L.n = n;
  L.wd = m->wd;
  L.clauses_total = (uint32_t) m->nclauses();
  L.clauses = m->clauses;
  L.classes = m->classes;
  L.tolerance = m->clause_tolerance;
  L.last_eval = m->last_eval_mask();
  HIP_TRY(hipEventRecord(m->ev0, m->stream));
  HIP_TRY(tm_launch_predict(L, m->stream));
  HIP_TRY(hipEventRecord(m->ev1, m->stream));
  return TM_OK;
}

int predict_finish_timing(tm_machine *m)
{
  HIP_TRY(hipStreamSynchronize(m->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
  m->predict_ms = ms;
  return TM_OK;
}

// Enqueue predict of device-resident rows into device output (no sync).
int predict_enqueue_dev(tm_machine *m, const uint8_t *rows, uint32_t n, int64_t *out_dev)
{
  int rc;
  if ((rc = predict_sums(m, rows, n, true))) return rc;
  HIP_TRY(tm_launch_argmax(m->p_sums.as<int64_t>(), n, m->classes, out_dev, m->stream));
  return TM_OK;
}

// Individualized calls pass dim_offsets (never NULL) and the byte size of rows.
int predict_impl(tm_machine *m, const uint8_t *rows, uint32_t n, void *out, bool dev, int want_type,
                 bool ind_call = false, const int64_t *dim_offsets = nullptr, size_t rows_bytes = 0)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (m->type != want_type) return fail(TM_E_ARG, "predict: machine is not a%s", want_type == TM_CLASSIFIER ? " classifier" : "n encoder");
  if (m->individualized && !ind_call)
    return fail(TM_E_ARG, "predict: individualized machine: predict with dim_offsets (the _ind entry points)");
  if (!m->individualized && ind_call)
    return fail(TM_E_ARG, "predict: dim_offsets given to a machine that is not individualized");
  if (n == 0) return TM_OK;
  if (!rows || !out) return fail(TM_E_ARG, "predict: problems and out must not be NULL");
  if ((rc = ind_call ? predict_sums_ind(m, rows, rows_bytes, dim_offsets, n, dev) : predict_sums(m, rows, n, dev)))
    return rc;
  const size_t out_bytes = want_type == TM_CLASSIFIER ? (size_t) n * 8 : (size_t) n * m->class_chunks;
  void *dout = out;
  if (!dev) {
    if ((rc = m->p_out.ensure(out_bytes))) return rc;
    dout = m->p_out.p;
  }
  if (want_type == TM_CLASSIFIER)
    HIP_TRY(tm_launch_argmax(m->p_sums.as<int64_t>(), n, m->classes, (int64_t *) dout, m->stream));
  else
    HIP_TRY(tm_launch_encode(m->p_sums.as<int64_t>(), n, m->classes, (uint8_t *) dout, m->stream));
  if (!dev) HIP_TRY(hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, m->stream));
  return predict_finish_timing(m);
}

// The classifier trains from per-epoch step queues unless TSETLIN_AMD_QUEUE=0.
bool queue_enabled()
{
  const char *v = getenv("TSETLIN_AMD_QUEUE");
  return !(v && v[0] == '0');
}

// Queue entries beyond this many bytes are not worth their memory: the
// epoch then runs with the in-kernel scan.
const uint64_t kQueueMaxBytes = 16ull << 30;

// Build the step queue of one classifier epoch (perm and epoch keys already
// in place) and point L at it.  *built = false: the queue would not fit, L is
// left in scan mode.
int build_queue(tm_machine *m, tm_train_launch &L, const uint32_t ek[4], uint32_t epoch, bool *built)
{
  *built = false;
  L.qoff = nullptr; L.qid = nullptr; L.qdraw = nullptr; L.qcap = 0; L.qseg = 0;
  const uint32_t nseg = tm_queue_segments(L.n);
  const uint64_t items = (uint64_t) L.total_blocks * nseg + 1;
  if (items > 0x7FFFFFFFull) return TM_OK;
  int rc;
  if ((rc = m->q_counts.ensure(items * 8)) || (rc = m->q_offsets.ensure(items * 8)) ||
      (rc = m->q_scratch.ensure(tm_queue_scratch_bytes(items))) || (rc = m->q_keys.ensure((size_t) L.n * 8)))
    return rc;
  tm_queue_launch Q;
  memset(&Q, 0, sizeof Q);
  Q.perm = L.one.perm;
  Q.labels = L.labels;
  Q.n = L.n;
  Q.chunks = L.total_blocks;
  Q.clause_chunks = L.one.clause_chunks;
  Q.nseg = nseg;
  Q.epoch = epoch;
  Q.ek0 = ek[0];
  Q.ek1 = ek[1];
  Q.seed_lo = L.one.seed_lo;
  Q.seed_hi = L.one.seed_hi;
  Q.skip_threshold = L.one.skip_threshold;
  Q.counts = m->q_counts.as<uint64_t>();
  Q.offsets = m->q_offsets.as<uint64_t>();
  Q.pos_keys = m->q_keys.as<uint32_t>();
  HIP_TRY(hipEventRecord(m->evq0, m->stream));
  HIP_TRY(tm_launch_queue_count(Q, m->q_scratch.p, m->q_scratch.cap, m->stream));
  uint64_t total = 0;
  HIP_TRY(hipMemcpyAsync(&total, Q.offsets + (items - 1), 8, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  // an entry: its id and one 32-bit draw per clause
  if (total * 36 > kQueueMaxBytes) return TM_OK;
  if (total > m->q_cap || !m->q_id.p || !m->q_draw.p) {
    // some headroom: the count moves a little from epoch to epoch
    const uint64_t cap = std::max<uint64_t>(64, total + total / 8);
    if ((rc = m->q_id.ensure(cap * 4)) || (rc = m->q_draw.ensure(cap * 32))) {
      m->q_id.release();
      m->q_draw.release();
      m->q_cap = 0;
      (void) hipGetLastError();
      return TM_OK;
    }
    m->q_cap = cap;
  }
  Q.qid = m->q_id.as<uint32_t>();
  Q.qdraw = m->q_draw.as<uint32_t>();
  Q.qcap = m->q_cap;
  HIP_TRY(tm_launch_queue_write(Q, m->stream));
  HIP_TRY(hipEventRecord(m->evq1, m->stream));
  L.qoff = Q.offsets;
  L.qid = Q.qid;
  L.qdraw = Q.qdraw;
  L.qcap = Q.qcap;
  L.qseg = nseg;
  *built = true;
  return TM_OK;
}

// tk_tsetlin_train_classifier / _encoder, capi.c:1112-1282; individualized
// machines (ind_call, with dim_offsets and the byte size of rows):
// tk_tsetlin_train_classifier_ind / _encoder_ind, capi.c:1284-1482.
int train_impl(tm_machine *m, uint32_t n, const uint8_t *rows, const int64_t *labels, const uint8_t *codes,
               uint32_t iterations, tm_epoch_fn each, void *user, bool dev, int want_type,
               bool ind_call = false, const int64_t *dim_offsets = nullptr, size_t rows_bytes = 0)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (m->type != want_type) return fail(TM_E_ARG, "train: machine is not a%s", want_type == TM_CLASSIFIER ? " classifier" : "n encoder");
  if (m->individualized && !ind_call)
    return fail(TM_E_ARG, "train: individualized machine: train with dim_offsets (the _ind entry points)");
  if (!m->individualized && ind_call)
    return fail(TM_E_ARG, "train: dim_offsets given to a machine that is not individualized");
  if (!m->has_state)
    return fail(TM_E_STATE, "can't train a model loaded without state (or released after training; create with reusable=true to train again)");
  if (n > 0 && !rows) return fail(TM_E_ARG, "train: problems must not be NULL");
  if (n > 0 && want_type == TM_CLASSIFIER && !labels) return fail(TM_E_ARG, "train: solutions must not be NULL");
  if (n > 0 && want_type == TM_ENCODER && !codes) return fail(TM_E_ARG, "train: codes must not be NULL");
  if ((rc = use_device(m))) return rc;
  m->train_ms = 0;
  m->queue_ms = 0;
  memset(m->stats, 0, sizeof m->stats);
  m->predict_ready = false;
  if (iterations == 0 || n == 0) {
    // setup still runs (capi.c:1136-1146) even when no sample is visited; it
    // fills every word alike, so the whole state goes in one call
    HIP_TRY(tm_launch_setup(m->counts.as<uint32_t>(), m->actions.as<uint32_t>(), m->action_words(),
                            m->state_bits - 1, 1, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    for (uint32_t e = 0; e < iterations; e++) {
      if (!each) continue;
      int r = each(user, e + 1);
      if (r < 0) { fprintf(stderr, "Error in epoch callback (status %d)\n", r); break; }
      if (r == 0) break;
    }
  } else {
    const uint32_t *words;
    const uint32_t stride = train_row_stride(m);
    std::vector<uint32_t> strides;     // individualized: per-class row strides and word offsets
    std::vector<uint64_t> woff;
    if (ind_call) {
      size_t extent;
      if ((rc = ind_extent(m, n, dim_offsets, rows_bytes, "train", &extent))) return rc;
      uint32_t lanes, wpl;
      train_tile(m, &lanes, &wpl);
      strides.resize(m->classes);
      for (uint32_t h = 0; h < m->classes; h++) strides[h] = (m->ind[h].wd + wpl - 1) / wpl * wpl;
      if ((rc = stage_rows_ind(m, rows, n, extent, dim_offsets, strides, m->t_rows, m->t_words, woff, dev)))
        return rc;
      words = m->t_words.as<uint32_t>();
    } else if ((rc = stage_rows(m, rows, n, dev, stride, m->t_rows, m->t_words, &words))) {
      return rc;
    }
    const int64_t *dlabels = nullptr;
    const uint8_t *dcodes = nullptr;
    if (want_type == TM_CLASSIFIER) {
      if (dev) dlabels = labels;
      else {
        if ((rc = m->t_labels.ensure((size_t) n * 8))) return rc;
        HIP_TRY(hipMemcpyAsync(m->t_labels.p, labels, (size_t) n * 8, hipMemcpyHostToDevice, m->stream));
        dlabels = m->t_labels.as<int64_t>();
      }
    } else {
      if (dev) dcodes = codes;
      else {
        if ((rc = m->t_codes.ensure((size_t) n * m->class_chunks))) return rc;
        HIP_TRY(hipMemcpyAsync(m->t_codes.p, codes, (size_t) n * m->class_chunks, hipMemcpyHostToDevice, m->stream));
        dcodes = m->t_codes.as<uint8_t>();
      }
    }
    // Without a callback, epochs run in groups inside one launch each (the
    // chunk state stays in registers across the epochs of a group).  A
    // classifier instead builds a step queue per epoch and launches once per
    // epoch: the queue takes the epoch-order scan, the negative-sampling
    // draws and the step draws out of the train kernel.
    const bool queue = want_type == TM_CLASSIFIER && queue_enabled() && !m->individualized && m->state_bits == 8;
    const uint32_t group = (each || queue) ? 1u : 8u;
    if ((rc = m->t_perm.ensure((size_t) group * n * 4))) return rc;
    if ((rc = m->t_perm_scratch.ensure(tm_perm_scratch_bytes(n)))) return rc;
    if ((rc = m->t_ekeys.ensure((size_t) group * 16))) return rc;
    if ((rc = m->t_stats.ensure(TM_NSTATS * 8))) return rc;
    HIP_TRY(tm_launch_zero(m->t_stats.p, TM_NSTATS * 8, m->stream));
    tm_train_launch L;
    memset(&L, 0, sizeof L);
    tm_train_trial &P = L.one;
    P.counts = m->counts.as<uint32_t>();
    P.actions = m->actions.as<uint32_t>();
    P.perm = m->t_perm.as<uint32_t>();
    P.ekeys = m->t_ekeys.as<uint32_t>();
    P.stats = m->t_stats.as<unsigned long long>();
    P.clause_chunks = m->clause_chunks;
    P.tolerance = m->clause_tolerance;
    P.maximum = m->clause_maximum;
    P.target = m->target;
    // capi.c:185: unsigned s = (2 * features) / specificity
    P.s_draws = (uint32_t)((double)(2u * m->features) / m->specificity);
    P.include_bits = m->include_bits;
    double q = 1.0 - m->negative;
    P.skip_threshold = !(q > 0) ? 0 : (q >= 1 ? (1ull << 32) : (uint64_t) floor(q * 4294967296.0));
    P.seed_lo = (uint32_t) m->seed;
    P.seed_hi = (uint32_t)(m->seed >> 32);
    P.active = 1;
    P.rows = words;
    P.row_stride = stride;
    P.wd = m->wd;
    P.input_bits = m->input_bits;
    P.last_live = m->last_live_mask();
    P.last_eval = m->last_eval_mask();
    P.class_base = 0;
    P.chunk_base = 0;
    L.rows = words;
    L.labels = dlabels;
    L.codes = dcodes;
    L.n = n;
    L.wd = m->wd;
    L.row_stride = stride;
    L.class_chunks = m->class_chunks;
    L.total_blocks = m->clause_chunks * m->classes;
    L.input_bits = m->input_bits;
    L.state_bits = m->state_bits;
    L.encoder = want_type == TM_ENCODER;
    L.last_live = m->last_live_mask();
    L.last_eval = m->last_eval_mask();
    L.all_ib1 = m->include_bits <= 1;
    train_tile(m, &L.lanes, &L.wpl);
    // An individualized machine trains as a table of one-class machines, one
    // per class, in one launch: class h's state, rows, width, masks and
    // random-literal count are its own (capi.c:1441-1454, s at capi.c:344);
    // its chunk c draws as chunk h * clause_chunks + c of the stream.
    std::vector<tm_train_trial> T;
    if (m->individualized) {
      T.assign(m->classes, P);
      std::vector<uint32_t> map;
      map.reserve((size_t) 2 * m->classes * m->clause_chunks);
      for (uint32_t h = 0; h < m->classes; h++) {
        const tm_ind_class &c = m->ind[h];
        tm_train_trial &Q = T[h];
        Q.counts = m->counts.as<uint32_t>() + c.coff;
        Q.actions = m->actions.as<uint32_t>() + c.aoff;
        Q.s_draws = (uint32_t)((double)(2 * c.features) / m->specificity);
        Q.rows = words + woff[h];
        Q.row_stride = strides[h];
        Q.wd = c.wd;
        Q.input_bits = c.input_bits;
        Q.last_live = live_mask(c.wd, c.input_chunks, c.tail_mask);
        Q.last_eval = eval_mask(c.wd, c.input_bits);
        Q.class_base = h;
        Q.chunk_base = h * m->clause_chunks;
        for (uint32_t ch = 0; ch < m->clause_chunks; ch++) { map.push_back(h); map.push_back(ch); }
      }
      if ((rc = m->t_trials.ensure(T.size() * sizeof(tm_train_trial))) || (rc = m->t_map.ensure(map.size() * 4)))
        return rc;
      HIP_TRY(hipMemcpyAsync(m->t_map.p, map.data(), map.size() * 4, hipMemcpyHostToDevice, m->stream));
      HIP_TRY(hipStreamSynchronize(m->stream));
      L.trials = m->t_trials.as<tm_train_trial>();
      L.block_map = m->t_map.as<uint32_t>();
    }
    L.all_full = 1;
    if (T.empty()) L.all_full = full_rows(P, L.wd);
    for (const tm_train_trial &Q : T) L.all_full &= full_rows(Q, L.wd);
    std::vector<uint32_t> ek(4 * group);
    bool stop = false;
    for (uint32_t e0 = 0; e0 < iterations && !stop; e0 += group) {
      const uint32_t cnt = (iterations - e0) < group ? (iterations - e0) : group;
      for (uint32_t e = 0; e < cnt; e++) {
        tm_u4 k = tm_philox(e0 + e, 0, 0, TM_TAG_EPOCH, P.seed_lo, P.seed_hi);
        ek[4 * e] = k.x; ek[4 * e + 1] = k.y; ek[4 * e + 2] = k.z; ek[4 * e + 3] = k.w;
      }
      HIP_TRY(hipMemcpyAsync(m->t_ekeys.p, ek.data(), (size_t) cnt * 16, hipMemcpyHostToDevice, m->stream));
      HIP_TRY(tm_launch_perm(m->seed, e0, cnt, n, m->t_perm.as<uint32_t>(), m->t_perm_scratch.p,
                             m->t_perm_scratch.cap, m->stream));
      P.perm_epoch0 = e0;
      if (!T.empty()) {
        for (tm_train_trial &Q : T) Q.perm_epoch0 = e0;
        HIP_TRY(hipMemcpyAsync(m->t_trials.p, T.data(), T.size() * sizeof(tm_train_trial), hipMemcpyHostToDevice,
                               m->stream));
      }
      L.epoch_first = e0;
      L.epochs = cnt;
      L.init = e0 == 0;
      bool queued = false;
      if (queue && (rc = build_queue(m, L, ek.data(), e0, &queued))) return rc;
      HIP_TRY(hipEventRecord(m->ev0, m->stream));
      HIP_TRY(tm_launch_train(L, m->stream));
      HIP_TRY(hipEventRecord(m->ev1, m->stream));
      HIP_TRY(hipStreamSynchronize(m->stream));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
      m->train_ms += ms;
      if (queued) {
        HIP_TRY(hipEventElapsedTime(&ms, m->evq0, m->evq1));
        m->queue_ms += ms;
      }
      m->predict_ready = false;
      if (each) {
        m->trained = true;
        int r = each(user, e0 + 1);
        if (r < 0) { fprintf(stderr, "Error in epoch callback (status %d)\n", r); stop = true; }
        else if (r == 0) stop = true;
        if ((rc = use_device(m))) return rc;
      }
    }
  }
  if (m->t_stats.p && iterations > 0 && n > 0) {
    HIP_TRY(hipMemcpyAsync(m->stats, m->t_stats.p, sizeof m->stats, hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
  }
  // the queue is sized by the data: give its memory back between calls
  m->q_id.release();
  m->q_draw.release();
  m->q_keys.release();
  m->q_cap = 0;
  // capi.c:1187-1189: a non-reusable machine releases its counters
  if (!m->reusable) {
    m->counts.release();
    m->has_state = false;
  }
  m->trained = true;
  return TM_OK;
}

void put_u32(std::vector<uint8_t> &v, uint32_t x) { v.insert(v.end(), (uint8_t *) &x, (uint8_t *) &x + 4); }

// The include plane as runs of clauses that share a row width: the whole
// machine, or one run per class of an individualized machine.
struct action_run {
  uint64_t clauses, bytes, wd, aoff, hoff;
};

std::vector<action_run> action_runs(const tm_machine *m)
{
  std::vector<action_run> r;
  if (!m->individualized) {
    r.push_back({ m->nclauses(), m->input_chunks, m->wd, 0, 0 });
  } else {
    for (const tm_ind_class &c : m->ind) r.push_back({ m->clauses, c.input_chunks, c.wd, c.aoff, c.hoff });
  }
  return r;
}

// Copy the include plane (planes = 1) or the count planes (planes = M)
// between the device layout, [clause][plane][wd] u32 words per run, and the
// reference's, [clause][plane][input_chunks] bytes per run (automata.h:7-31;
// an individualized machine's classes one after another, capi.c:576-593).
// `ref` is host or device memory; to_ref picks the direction.  Device bytes
// past a row's input_chunks are not part of the reference layout: a copy into
// the device leaves them zero.
int layout_copy(tm_machine *m, uint8_t *ref, bool ref_dev, bool to_ref, bool counts)
{
  const uint64_t planes = counts ? m->state_bits - 1 : 1;
  const dbuf &dev = counts ? m->counts : m->actions;
  const std::vector<action_run> runs = action_runs(m);
  const hipMemcpyKind kind = to_ref ? (ref_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost)
                                    : (ref_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
  bool padded = false;
  for (const action_run &r : runs) padded |= r.bytes != r.wd * 4;
  if (!to_ref && padded) HIP_TRY(tm_launch_zero(dev.p, m->action_words() * planes * 4, m->stream));
  for (const action_run &r : runs) {
    uint8_t *d = dev.as<uint8_t>() + r.aoff * planes * 4;
    uint8_t *h = ref + r.hoff * planes;
    const size_t height = r.clauses * planes, pitch = r.wd * 4;
    if (r.bytes == pitch)
      HIP_TRY(hipMemcpyAsync(to_ref ? h : d, to_ref ? d : h, height * pitch, kind, m->stream));
    else if (to_ref)
      HIP_TRY(hipMemcpy2DAsync(h, r.bytes, d, pitch, r.bytes, height, kind, m->stream));
    else
      HIP_TRY(hipMemcpy2DAsync(d, pitch, h, r.bytes, r.bytes, height, kind, m->stream));
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  return TM_OK;
}

int actions_to_host(tm_machine *m, uint8_t *dst) { return layout_copy(m, dst, false, true, false); }

int actions_from_host(tm_machine *m, const uint8_t *src)
{
  int rc = layout_copy(m, const_cast<uint8_t *>(src), false, false, false);
  m->predict_ready = false;
  return rc;
}

// capi.c:1509-1531 (+ the int32 type written first at capi.c:1543).
int serialize(tm_machine *m, std::vector<uint8_t> &v)
{
  v.clear();
  int32_t type = m->type;
  v.insert(v.end(), (uint8_t *) &type, (uint8_t *) &type + 4);
  put_u32(v, m->classes);
  put_u32(v, m->features);
  put_u32(v, m->clauses);
  put_u32(v, m->clause_tolerance);
  put_u32(v, m->clause_maximum);
  put_u32(v, m->target);
  put_u32(v, m->state_bits);
  put_u32(v, m->include_bits);
  put_u32(v, m->input_bits);
  put_u32(v, m->input_chunks);
  put_u32(v, m->clause_chunks);
  put_u32(v, m->state_chunks);
  put_u32(v, m->action_chunks);
  v.insert(v.end(), (uint8_t *) &m->specificity, (uint8_t *) &m->specificity + 8);
  v.push_back(m->tail_mask);
  size_t off = v.size();
  v.resize(off + m->action_chunks);
  int rc = actions_to_host(m, v.data() + off);
  if (rc) return rc;
  v.push_back(m->individualized ? 1 : 0);
  // capi.c:1528-1530: an individualized machine appends its feature sizes
  for (const tm_ind_class &c : m->ind) v.insert(v.end(), (const uint8_t *) &c.features, (const uint8_t *) &c.features + 8);
  return TM_OK;
}

struct reader {
  const uint8_t *p;
  size_t len, off = 0;
  bool take(void *dst, size_t n)
  {
    if (off + n > len) return false;
    memcpy(dst, p + off, n);
    off += n;
    return true;
  }
};

} // namespace

extern "C" {

void tm_options_default(tm_options *o)
{
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->state = 8;
  o->include_bits = 1;
  o->target = -1.0;
  o->negative = -1.0;
}

const char *tm_last_error(void) { return g_err; }

int tm_create_classifier(const tm_options *opts, tm_machine **out) { return create_common(opts, TM_CLASSIFIER, out); }
int tm_create_encoder(const tm_options *opts, tm_machine **out) { return create_common(opts, TM_ENCODER, out); }

int tm_train_classifier(tm_machine *m, uint32_t n, const uint8_t *problems, const int64_t *solutions,
                        uint32_t iterations, tm_epoch_fn each, void *user)
{
  return train_impl(m, n, problems, solutions, nullptr, iterations, each, user, false, TM_CLASSIFIER);
}

int tm_train_encoder(tm_machine *m, uint32_t n, const uint8_t *sentences, const uint8_t *codes,
                     uint32_t iterations, tm_epoch_fn each, void *user)
{
  return train_impl(m, n, sentences, nullptr, codes, iterations, each, user, false, TM_ENCODER);
}

int tm_train_classifier_dev(tm_machine *m, uint32_t n, const uint8_t *problems, const int64_t *solutions,
                            uint32_t iterations, tm_epoch_fn each, void *user)
{
  return train_impl(m, n, problems, solutions, nullptr, iterations, each, user, true, TM_CLASSIFIER);
}

int tm_train_encoder_dev(tm_machine *m, uint32_t n, const uint8_t *sentences, const uint8_t *codes,
                         uint32_t iterations, tm_epoch_fn each, void *user)
{
  return train_impl(m, n, sentences, nullptr, codes, iterations, each, user, true, TM_ENCODER);
}

int tm_predict_classifier(tm_machine *m, const uint8_t *problems, uint32_t n, int64_t *out)
{
  return predict_impl(m, problems, n, out, false, TM_CLASSIFIER);
}

int tm_predict_encoder(tm_machine *m, const uint8_t *problems, uint32_t n, uint8_t *out)
{
  return predict_impl(m, problems, n, out, false, TM_ENCODER);
}

int tm_predict_classifier_dev(tm_machine *m, const uint8_t *problems, uint32_t n, int64_t *out)
{
  return predict_impl(m, problems, n, out, true, TM_CLASSIFIER);
}

int tm_predict_encoder_dev(tm_machine *m, const uint8_t *problems, uint32_t n, uint8_t *out)
{
  return predict_impl(m, problems, n, out, true, TM_ENCODER);
}

int tm_create_classifier_ind(const tm_options *opts, const int64_t *feat_offsets, uint32_t n_offsets, tm_machine **out)
{
  return create_ind(opts, feat_offsets, n_offsets, TM_CLASSIFIER, out);
}

int tm_create_encoder_ind(const tm_options *opts, const int64_t *feat_offsets, uint32_t n_offsets, tm_machine **out)
{
  return create_ind(opts, feat_offsets, n_offsets, TM_ENCODER, out);
}

int tm_train_classifier_ind(tm_machine *m, uint32_t n, const uint8_t *problems, size_t problems_bytes,
                            const int64_t *solutions, const int64_t *dim_offsets, uint32_t iterations,
                            tm_epoch_fn each, void *user)
{
  return train_impl(m, n, problems, solutions, nullptr, iterations, each, user, false, TM_CLASSIFIER, true,
                    dim_offsets, problems_bytes);
}

int tm_train_encoder_ind(tm_machine *m, uint32_t n, const uint8_t *sentences, size_t sentences_bytes,
                         const uint8_t *codes, const int64_t *dim_offsets, uint32_t iterations, tm_epoch_fn each,
                         void *user)
{
  return train_impl(m, n, sentences, nullptr, codes, iterations, each, user, false, TM_ENCODER, true, dim_offsets,
                    sentences_bytes);
}

int tm_predict_classifier_ind(tm_machine *m, const uint8_t *problems, size_t problems_bytes,
                              const int64_t *dim_offsets, uint32_t n, int64_t *out)
{
  return predict_impl(m, problems, n, out, false, TM_CLASSIFIER, true, dim_offsets, problems_bytes);
}

int tm_predict_encoder_ind(tm_machine *m, const uint8_t *problems, size_t problems_bytes,
                           const int64_t *dim_offsets, uint32_t n, uint8_t *out)
{
  return predict_impl(m, problems, n, out, false, TM_ENCODER, true, dim_offsets, problems_bytes);
}

int tm_train_classifier_ind_dev(tm_machine *m, uint32_t n, const uint8_t *problems_dev, size_t problems_bytes,
                                const int64_t *solutions_dev, const int64_t *dim_offsets, uint32_t iterations,
                                tm_epoch_fn each, void *user)
{
  return train_impl(m, n, problems_dev, solutions_dev, nullptr, iterations, each, user, true, TM_CLASSIFIER, true,
                    dim_offsets, problems_bytes);
}

int tm_train_encoder_ind_dev(tm_machine *m, uint32_t n, const uint8_t *sentences_dev, size_t sentences_bytes,
                             const uint8_t *codes_dev, const int64_t *dim_offsets, uint32_t iterations,
                             tm_epoch_fn each, void *user)
{
  return train_impl(m, n, sentences_dev, nullptr, codes_dev, iterations, each, user, true, TM_ENCODER, true,
                    dim_offsets, sentences_bytes);
}

int tm_predict_classifier_ind_dev(tm_machine *m, const uint8_t *problems_dev, size_t problems_bytes,
                                  const int64_t *dim_offsets, uint32_t n, int64_t *out_dev)
{
  return predict_impl(m, problems_dev, n, out_dev, true, TM_CLASSIFIER, true, dim_offsets, problems_bytes);
}

int tm_predict_encoder_ind_dev(tm_machine *m, const uint8_t *problems_dev, size_t problems_bytes,
                               const int64_t *dim_offsets, uint32_t n, uint8_t *out_dev)
{
  return predict_impl(m, problems_dev, n, out_dev, true, TM_ENCODER, true, dim_offsets, problems_bytes);
}

// capi.c:752-900: keep the classes listed in `selected`, in that order, with
// their actions (and counts, while the machine has them).
int tm_restrict(tm_machine *m, const int64_t *selected, uint32_t n_selected)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (!m->individualized) return fail(TM_E_ARG, "restrict only supported for individualized TMs");
  if (n_selected == 0) return fail(TM_E_ARG, "selected_dims cannot be empty");
  if (!selected) return fail(TM_E_ARG, "restrict: selected_dims is NULL");
  for (uint32_t i = 0; i < n_selected; i++) {
    const int64_t dim = selected[i];
    if (dim < 0 || (uint64_t) dim >= m->classes)
      return fail(TM_E_ARG, "selected_dims[%d] = %d out of range [0, %d)", (int) i, (int) dim, (int) m->classes);
    for (uint32_t j = i + 1; j < n_selected; j++)
      if (selected[i] == selected[j]) return fail(TM_E_ARG, "duplicate dimension %d in selected_dims", (int) dim);
  }
  if ((rc = use_device(m))) return rc;
  const std::vector<tm_ind_class> old = m->ind;
  const uint32_t old_classes = m->classes;
  auto undo = [&]() {
    m->ind = old;
    m->classes = old_classes;
    ind_layout(m, "restrict");
  };
  m->ind.assign(n_selected, tm_ind_class());
  for (uint32_t i = 0; i < n_selected; i++) m->ind[i].features = old[(size_t) selected[i]].features;
  m->classes = n_selected;
  if ((rc = ind_layout(m, "restrict"))) { undo(); return rc; }
  const uint64_t M = m->state_bits - 1;
  dbuf na, nc;
  if ((rc = na.ensure(m->action_words() * 4)) || (m->has_state && (rc = nc.ensure(m->action_words() * M * 4)))) {
    na.release();
    nc.release();
    undo();
    return rc;
  }
  hipError_t e = hipSuccess;
  for (uint32_t i = 0; i < n_selected && e == hipSuccess; i++) {
    const tm_ind_class &src = old[(size_t) selected[i]], &dst = m->ind[i];
    const uint64_t words = (uint64_t) m->clauses * dst.wd;
    e = hipMemcpyAsync(na.as<uint32_t>() + dst.aoff, m->actions.as<uint32_t>() + src.aoff, words * 4,
                       hipMemcpyDeviceToDevice, m->stream);
    if (e == hipSuccess && m->has_state)
      e = hipMemcpyAsync(nc.as<uint32_t>() + dst.coff, m->counts.as<uint32_t>() + src.coff, words * M * 4,
                         hipMemcpyDeviceToDevice, m->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    na.release();
    nc.release();
    undo();
    return fail(TM_E_DEVICE, "restrict: device copy failed: %s", hipGetErrorString(e));
  }
  m->actions.release();
  m->actions = na;
  if (m->has_state) {
    m->counts.release();
    m->counts = nc;
  }
  m->class_chunks = (m->classes + 7) / 8;
  return alloc_predict_views(m);
}

int tm_get_feature_sizes(const tm_machine *m, uint64_t *out, uint32_t cap)
{
  if (!m) return fail(TM_E_ARG, "machine is NULL");
  if (!m->individualized) return fail(TM_E_ARG, "feature sizes: machine is not individualized");
  if (!out || cap < m->classes) return fail(TM_E_SIZE, "feature sizes: room for %u, need %u", cap, m->classes);
  for (uint32_t h = 0; h < m->classes; h++) out[h] = m->ind[h].features;
  return TM_OK;
}

int tm_persist(tm_machine *m, uint8_t *buf, size_t cap, size_t *len)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (!len) return fail(TM_E_ARG, "persist: len is NULL");
  if ((rc = use_device(m))) return rc;
  std::vector<uint8_t> v;
  if ((rc = serialize(m, v))) return rc;
  *len = v.size();
  if (!buf || cap < v.size()) return fail(TM_E_SIZE, "persist: buffer of %zu bytes, need %zu", cap, v.size());
  memcpy(buf, v.data(), v.size());
  return TM_OK;
}

int tm_persist_file(tm_machine *m, const char *path)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (!path) return fail(TM_E_ARG, "persist: path is NULL");
  if ((rc = use_device(m))) return rc;
  std::vector<uint8_t> v;
  if ((rc = serialize(m, v))) return rc;
  FILE *fh = fopen(path, "wb");
  if (!fh) return fail(TM_E_ARG, "persist: cannot open %s", path);
  size_t w = fwrite(v.data(), 1, v.size(), fh);
  int cr = fclose(fh);
  if (w != v.size() || cr != 0) return fail(TM_E_ARG, "persist: short write to %s", path);
  return TM_OK;
}

// capi.c:1675-1737 and :1751-1772.
int tm_load(const uint8_t *data, size_t len, tm_machine **out)
{
  if (!out) return fail(TM_E_ARG, "load: out is NULL");
  *out = nullptr;
  if (!data) return fail(TM_E_ARG, "load: data is NULL");
  tm_layout::header h;
  if (const char *e = tm_layout::read_header(data, len, &h)) return fail(TM_E_FORMAT, "%s", e);
  if (const char *e = tm_layout::check_common(h)) return fail(TM_E_FORMAT, "%s", e);
  reader r{ data, len, tm_layout::kHeaderBytes };
  tm_machine *m = new tm_machine();
  m->type = h.type;
  m->classes = h.classes; m->features = h.features; m->clauses = h.clauses; m->clause_tolerance = h.clause_tolerance;
  m->clause_maximum = h.clause_maximum; m->target = h.target; m->state_bits = h.state_bits;
  m->include_bits = h.include_bits ? h.include_bits : 1;
  m->input_bits = h.input_bits; m->input_chunks = h.input_chunks; m->clause_chunks = h.clause_chunks;
  m->state_chunks = h.state_chunks; m->action_chunks = h.action_chunks;
  m->specificity = h.specificity;
  m->tail_mask = h.tail_mask;
  m->class_chunks = (m->classes + 7) / 8;
  m->negative = 1.0 / (double) m->classes;   // not persisted (capi.c:1509-1531)
  int rc;
  if ((rc = check_tolerance(m->clause_tolerance, "load"))) { delete m; return rc; }
  // the actions and the individualized byte must be there before any buffer
  // is sized from the header
  if ((uint64_t) m->action_chunks + 1 > len - r.off) { delete m; return fail(TM_E_FORMAT, "load: truncated actions"); }
  std::vector<uint8_t> acts(m->action_chunks);
  uint8_t ind = 0;
  if (!r.take(acts.data(), acts.size()) || !r.take(&ind, 1)) { delete m; return fail(TM_E_FORMAT, "load: truncated actions"); }
  if (ind) {
    // capi.c:1698-1725: the feature sizes follow; the per-class layout is
    // derived from them
    m->individualized = true;
    if ((uint64_t) m->classes * 8 > len - r.off) { delete m; return fail(TM_E_FORMAT, "load: truncated feature sizes"); }
    m->ind.resize(m->classes);
    for (uint32_t h = 0; h < m->classes; h++)
      if (!r.take(&m->ind[h].features, 8)) { delete m; return fail(TM_E_FORMAT, "load: truncated feature sizes"); }
    const uint32_t persisted_actions = m->action_chunks;
    if (m->state_bits < 2) { delete m; return fail(TM_E_FORMAT, "load: inconsistent header"); }
    if ((rc = ind_layout(m, "load", false))) {
      char msg[sizeof g_err];
      snprintf(msg, sizeof msg, "%s", g_err);
      delete m;
      return rc == TM_E_ARG ? fail(TM_E_FORMAT, "%s", msg) : rc;
    }
    if (m->action_chunks != persisted_actions) { delete m; return fail(TM_E_FORMAT, "load: inconsistent header"); }
  } else {
    if (const char *e = tm_layout::check_plain(h)) { delete m; return fail(TM_E_FORMAT, "%s", e); }
    if ((rc = size_check(m->classes, m->clauses, m->input_chunks, m->state_bits < 2 ? 2 : m->state_bits, false))) { delete m; return rc; }
    m->features = m->input_bits / 2;
    m->wd = (m->input_chunks + 3) / 4;
  }
  if ((rc = init_runtime(m))) { tm_destroy(m); return rc; }
  if ((rc = m->actions.ensure(m->action_words() * 4)) || (rc = alloc_predict_views(m))) { tm_destroy(m); return rc; }
  if ((rc = actions_from_host(m, acts.data()))) { tm_destroy(m); return rc; }
  m->has_state = false;   // capi.c:1696: state = NULL after load
  m->trained = true;
  *out = m;
  return TM_OK;
}

int tm_load_file(const char *path, tm_machine **out)
{
  if (!path) return fail(TM_E_ARG, "load: path is NULL");
  FILE *fh = fopen(path, "rb");
  if (!fh) return fail(TM_E_ARG, "load: cannot open %s", path);
  std::vector<uint8_t> v;
  uint8_t tmp[65536];
  size_t k;
  while ((k = fread(tmp, 1, sizeof tmp, fh)) > 0) v.insert(v.end(), tmp, tmp + k);
  fclose(fh);
  return tm_load(v.data(), v.size(), out);
}

int tm_checkpoint(tm_machine *m, uint8_t *buf, size_t cap, size_t *len)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (!len) return fail(TM_E_ARG, "checkpoint: len is NULL");
  *len = m->action_chunks;
  if (!buf || cap < m->action_chunks) return fail(TM_E_SIZE, "checkpoint: buffer of %zu bytes, need %u", cap, m->action_chunks);
  if ((rc = use_device(m))) return rc;
  return actions_to_host(m, buf);
}

int tm_restore(tm_machine *m, const uint8_t *buf, size_t len)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (len != m->action_chunks)
    return fail(TM_E_SIZE, "checkpoint size mismatch: expected %u bytes, got %zu", m->action_chunks, len);
  if (!buf) return fail(TM_E_ARG, "restore: buffer is NULL");
  if ((rc = use_device(m))) return rc;
  return actions_from_host(m, buf);
}

int tm_checkpoint_dev(tm_machine *m, uint8_t *buf_dev, size_t cap, size_t *len)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (!len) return fail(TM_E_ARG, "checkpoint: len is NULL");
  *len = m->action_chunks;
  if (!buf_dev || cap < m->action_chunks)
    return fail(TM_E_SIZE, "checkpoint: buffer of %zu bytes, need %u", cap, m->action_chunks);
  if ((rc = use_device(m))) return rc;
  return layout_copy(m, buf_dev, true, true, false);
}

int tm_restore_dev(tm_machine *m, const uint8_t *buf_dev, size_t len)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (len != m->action_chunks)
    return fail(TM_E_SIZE, "checkpoint size mismatch: expected %u bytes, got %zu", m->action_chunks, len);
  if (!buf_dev) return fail(TM_E_ARG, "restore: buffer is NULL");
  if ((rc = use_device(m))) return rc;
  rc = layout_copy(m, const_cast<uint8_t *>(buf_dev), true, false, false);
  m->predict_ready = false;
  return rc;
}

int tm_counts(tm_machine *m, uint8_t *buf, size_t cap, size_t *len)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (!len) return fail(TM_E_ARG, "counts: len is NULL");
  *len = m->state_chunks;
  if (!m->has_state)
    return fail(TM_E_STATE, "counts: the machine holds no counters (loaded, or released after training; create "
                            "with reusable=true to keep them)");
  if (!buf || cap < m->state_chunks) return fail(TM_E_SIZE, "counts: buffer of %zu bytes, need %u", cap, m->state_chunks);
  if ((rc = use_device(m))) return rc;
  return layout_copy(m, buf, false, true, true);
}

// capi.c:1590-1673.
int tm_reconfigure(tm_machine *m, const tm_options *o)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (!m->reusable) return fail(TM_E_STATE, "reconfigure requires reusable=true at creation time");
  if (!o) return fail(TM_E_ARG, "reconfigure: options are NULL");
  if (!o->clauses) return fail(TM_E_ARG, "reconfigure: clauses must be greater than 0");
  if (o->clauses > 0x7FFFFFF8u) return fail(TM_E_ARG, "reconfigure: clauses too large for the 32-bit size fields");
  if (!o->clause_tolerance) return fail(TM_E_ARG, "reconfigure: clause_tolerance must be greater than 0");
  if (!o->clause_maximum) return fail(TM_E_ARG, "reconfigure: clause_maximum must be greater than 0");
  if (!(o->specificity > 0)) return fail(TM_E_ARG, "reconfigure: specificity must be a positive number");
  if ((rc = check_tolerance(o->clause_tolerance, "reconfigure"))) return rc;
  if ((rc = use_device(m))) return rc;
  const uint32_t nclauses = ((o->clauses + 7) / 8) * 8;
  uint32_t target;
  if ((rc = derive_target(o->target, nclauses, o->clause_tolerance, "reconfigure", &target))) return rc;
  if (m->individualized) {
    // capi.c:1609-1615, :1651-1663: the per-class layout follows the new
    // clause count
    const uint32_t old = m->clauses;
    m->clauses = nclauses;
    if ((rc = ind_layout(m, "reconfigure"))) {
      m->clauses = old;
      ind_layout(m, "reconfigure");
      return rc;
    }
  } else {
    if ((rc = size_check(m->classes, nclauses, m->input_chunks, m->state_bits))) return rc;
    m->clauses = nclauses;
    m->action_chunks = m->classes * nclauses * m->input_chunks;
    m->state_chunks = m->classes * nclauses * (m->state_bits - 1) * m->input_chunks;
  }
  m->clause_chunks = nclauses / 8;
  m->clause_tolerance = o->clause_tolerance;
  m->clause_maximum = o->clause_maximum;
  m->include_bits = o->include_bits ? o->include_bits : 1;
  m->specificity = o->specificity;
  m->target = target;
  m->seed = o->seed;
  if ((rc = alloc_device(m))) return rc;   // actions zeroed, counts (re)allocated
  m->predict_ready = false;
  m->trained = false;
  return TM_OK;
}

int tm_type(const tm_machine *m)
{
  if (!m) return fail(TM_E_ARG, "machine is NULL");
  return m->type;
}

void tm_destroy(tm_machine *m)
{
  if (!m) return;
  if (m->stream) hipStreamSynchronize(m->stream);
  dbuf *bufs[] = { &m->counts, &m->actions, &m->actT, &m->lit, &m->t_rows, &m->t_words, &m->t_labels,
                   &m->t_codes, &m->t_perm, &m->t_perm_scratch, &m->t_ekeys, &m->p_rows, &m->p_words,
                   &m->p_sums, &m->p_out, &m->t_stats, &m->q_counts, &m->q_offsets, &m->q_scratch,
                   &m->q_id, &m->q_draw, &m->q_keys, &m->t_trials, &m->t_map };
  for (dbuf *b : bufs) b->release();
  if (m->ev0) hipEventDestroy(m->ev0);
  if (m->ev1) hipEventDestroy(m->ev1);
  if (m->evq0) hipEventDestroy(m->evq0);
  if (m->evq1) hipEventDestroy(m->evq1);
  if (m->stream && m->owns_stream) hipStreamDestroy(m->stream);
  delete m;
}

int tm_get_info(const tm_machine *m, tm_info *o)
{
  if (!m || !o) return fail(TM_E_ARG, "get_info: NULL argument");
  o->type = (uint32_t) m->type;
  o->classes = m->classes; o->features = m->features; o->clauses = m->clauses;
  o->clause_tolerance = m->clause_tolerance; o->clause_maximum = m->clause_maximum; o->target = m->target;
  o->state_bits = m->state_bits; o->include_bits = m->include_bits; o->input_bits = m->input_bits;
  o->input_chunks = m->input_chunks; o->clause_chunks = m->clause_chunks; o->state_chunks = m->state_chunks;
  o->action_chunks = m->action_chunks; o->class_chunks = m->class_chunks;
  o->specificity = m->specificity; o->negative = m->negative; o->tail_mask = m->tail_mask;
  o->reusable = m->reusable; o->has_state = m->has_state; o->trained = m->trained;
  o->individualized = m->individualized;
  return TM_OK;
}

int tm_last_train_stats(const tm_machine *m, uint64_t out[TM_NSTATS])
{
  if (!m || !out) return fail(TM_E_ARG, "last_train_stats: NULL argument");
  memcpy(out, m->stats, sizeof m->stats);
  return TM_OK;
}

int tm_last_timing(const tm_machine *m, double *train_ms, double *predict_ms)
{
  if (!m) return fail(TM_E_ARG, "machine is NULL");
  if (train_ms) *train_ms = m->train_ms;
  if (predict_ms) *predict_ms = m->predict_ms;
  return TM_OK;
}

int tm_last_queue_timing(const tm_machine *m, double *queue_ms)
{
  if (!m) return fail(TM_E_ARG, "machine is NULL");
  if (queue_ms) *queue_ms = m->queue_ms;
  return TM_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Predict without allocation.  The class sums (and the rows, when they must be
// repacked into whole words) go to `scratch`, a file-scope device array of
// kNoAllocScratch bytes that every GPU holds for the life of the process.  Its
// address on a device comes from hipGetSymbolAddress, taken once per device
// with that device current; host code uses only that address, never the
// array's name.  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU race.

constexpr size_t kNoAllocScratch = (size_t) 32 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratch];

namespace {

std::mutex g_scratch_mu;
std::map<int, uint8_t *> g_scratch_ptr;

// The scratch array's address on the current device (looked up on the first
// call on that device).
int scratch_on_device(int device, uint8_t **out)
{
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) { *out = it->second; return TM_OK; }
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(TM_E_DEVICE, "predict (no allocation): hipGetSymbolAddress of the scratch array failed: %s",
                hipGetErrorString(e));
  g_scratch_ptr[device] = (uint8_t *) p;
  *out = (uint8_t *) p;
  return TM_OK;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

} // namespace

extern "C" {

int tm_predict_classifier_dev_noalloc(tm_machine *m, const uint8_t *problems_dev, uint32_t n, int64_t *out_dev)
{
  int rc;
  if ((rc = check_handle(m))) return rc;
  if (m->type != TM_CLASSIFIER) return fail(TM_E_ARG, "predict: machine is not a classifier");
  if (m->individualized)
    return fail(TM_E_ARG, "predict: individualized machine: predict with dim_offsets (the _ind entry points)");
  if (n == 0) return TM_OK;
  if (!problems_dev || !out_dev) return fail(TM_E_ARG, "predict: problems and out must not be NULL");
  const size_t B = m->input_chunks;
  const bool in_place = (B % 4 == 0) && ((uintptr_t) problems_dev % 16 == 0);
  const size_t sums_bytes = align_up((size_t) n * m->classes * 8, 256);
  const size_t rows_bytes = in_place ? 0 : (size_t) n * m->wd * 4;
  if (sums_bytes + rows_bytes > kNoAllocScratch)
    return fail(TM_E_SIZE, "predict (no allocation): %u rows need %zu bytes of scratch, the limit is %zu "
                "(class sums %zu bytes%s)", n, sums_bytes + rows_bytes, kNoAllocScratch, sums_bytes,
                in_place ? "" : " and rows repacked into whole words");
  if (!m->actT.p || m->actT.cap < m->action_words() * 4 || !m->lit.p || m->lit.cap < m->nclauses() * 4)
    return fail(TM_E_STATE, "predict (no allocation): the machine's predict views are not allocated");
  if ((rc = use_device(m))) return rc;
  uint8_t *base;
  if ((rc = scratch_on_device(m->device, &base))) return rc;
  if (!m->predict_ready) {   // the buffers exist: this only launches the transposes
    HIP_TRY(tm_launch_prepare_predict(m->actions.as<uint32_t>(), m->actT.as<uint32_t>(), m->lit.as<uint32_t>(),
                                      (uint32_t) m->nclauses(), m->wd, m->last_eval_mask(), m->stream));
    m->predict_ready = true;
  }
  int64_t *sums = (int64_t *) base;
  const uint32_t *words = (const uint32_t *) problems_dev;
  if (!in_place) {
    uint32_t *w = (uint32_t *)(base + sums_bytes);
    HIP_TRY(tm_launch_repack(problems_dev, n, (uint32_t) B, m->wd, w, m->stream));
    words = w;
  }
  // a fill kernel, not a memset: the sums start at zero on every call
  HIP_TRY(tm_launch_zero(sums, (size_t) n * m->classes * 8, m->stream));
  tm_predict_launch L;
  memset(&L, 0, sizeof L);
  L.actT = m->actT.as<uint32_t>();
  L.lit = m->lit.as<uint32_t>();
  L.rows = words;
  L.sums = sums;
  L.n = n;
  L.wd = m->wd;
  L.clauses_total = (uint32_t) m->nclauses();
  L.clauses = m->clauses;
  L.classes = m->classes;
  L.tolerance = m->clause_tolerance;
  L.last_eval = m->last_eval_mask();
  HIP_TRY(hipEventRecord(m->ev0, m->stream));
  HIP_TRY(tm_launch_predict(L, m->stream));
  HIP_TRY(hipEventRecord(m->ev1, m->stream));
  HIP_TRY(tm_launch_argmax(sums, n, m->classes, out_dev, m->stream));
  return predict_finish_timing(m);
}

} // extern "C"

// ---------------------------------------------------------------------------
// Batched search trials

struct tm_batch {
  std::vector<tm_machine *> m;
  int type = TM_CLASSIFIER;      // every trial's machine type
  bool ind = false;              // individualized trials (one feature set per class, shared by the trials)
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t wd = 0, input_chunks = 0;
  dbuf trials, block_map, perm, perm_scratch, ekeys, epoch, stats, rows, words, labels, codes, hits, pred;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipEvent_t> tev;   // individualized scoring: trial t's predictions are ready
  uint32_t lanes = 0, wpl = 0, order = 0;
  double tuned_ms = 0, train_ms = 0, eval_ms = 0;
};

namespace {

// Autotuner choices keyed by (row words, state bits, row stride, workgroup
// count rounded down to a power of two, capped at kTuneBlocksCap): the
// candidate tiles depend on the first two, a tile must divide the stride the
// rows were staged with, and the choice depends on the trial mix mostly
// through how many workgroups it makes, until there are enough to fill the
// GPU several times over.  A search round of a new trial mix then reuses an
// earlier tuning: before the cap, the first rounds of a search, whose
// batches grow from a few thousand chunks to tens of thousands, re-tuned
// at each new power of two (seven timed epochs each, about a third of a
// search's wall time).
constexpr uint64_t kTuneBlocksCap = 4096;
std::mutex g_tune_mu;
std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>, std::tuple<uint32_t, uint32_t, uint32_t, double>>
    g_tune_cache;

void batch_release_graph(tm_batch *b)
{
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Workgroup map (trial, chunk): order 0 is trial-major; order 1 starts with
// the trials of most random-literal draws per event (the heaviest chunks).
// (table entry, chunk) per workgroup; entry e has chunks[e] chunks (a plain
// trial: all of its classes' chunks; an individualized trial's class: that
// class's chunks)
std::vector<uint32_t> batch_map(const std::vector<tm_train_trial> &T, const std::vector<uint32_t> &chunks,
                                uint32_t order)
{
  std::vector<uint32_t> idx(T.size());
  for (uint32_t t = 0; t < idx.size(); t++) idx[t] = t;
  if (order == 1)
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t c) {
      return (uint64_t) T[a].s_draws * 8 + T[a].clause_chunks > (uint64_t) T[c].s_draws * 8 + T[c].clause_chunks;
    });
  std::vector<uint32_t> map;
  for (uint32_t t : idx) {
    if (!T[t].active) continue;
    for (uint32_t c = 0; c < chunks[t]; c++) { map.push_back(t); map.push_back(c); }
  }
  return map;
}

// labels (classifier trials) or codes (encoder trials, [n][class_chunks]
// bytes as tm_train_encoder takes them): the batch's type picks which.
// Individualized batches take the problems' byte size and dim_offsets (class
// h's rows start at byte dim_offsets[h], as tm_train_classifier_ind).
int batch_train(tm_batch *b, uint32_t n, const uint8_t *rows, size_t rows_bytes, const int64_t *dim_offsets,
                const int64_t *labels, const uint8_t *codes, uint32_t iterations, tm_batch_epoch_fn each, void *user,
                bool dev, int want_type)
{
  if (!b) return fail(TM_E_ARG, "batch is NULL");
  if (b->ind != (dim_offsets != nullptr))
    return fail(TM_E_ARG, b->ind ? "batch train: an individualized batch takes dim_offsets"
                                 : "batch train: dim_offsets given for a batch that is not individualized");
  if (b->type != want_type)
    return fail(TM_E_ARG, "batch train: the batch holds %s trials", b->type == TM_CLASSIFIER ? "classifier" : "encoder");
  if (n > 0 && (!rows || (want_type == TM_CLASSIFIER ? !labels : !codes)))
    return fail(TM_E_ARG, "batch train: problems and %s must not be NULL", want_type == TM_CLASSIFIER ? "solutions" : "codes");
  HIP_TRY(hipSetDevice(b->device));
  const uint32_t nt = (uint32_t) b->m.size();
  tm_machine *m0 = b->m[0];
  b->train_ms = 0;
  for (tm_machine *m : b->m) { m->predict_ready = false; m->trained = false; memset(m->stats, 0, sizeof m->stats); }
  if (iterations == 0 || n == 0) {
    // setup fills every word alike: the whole state in one call (also the
    // ragged per-class layout of an individualized machine)
    for (tm_machine *m : b->m)
      HIP_TRY(tm_launch_setup(m->counts.as<uint32_t>(), m->actions.as<uint32_t>(), m->action_words(),
                              m->state_bits - 1, 1, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (tm_machine *m : b->m) m->trained = true;
    return TM_OK;
  }
  // rows: a stride every candidate tile can read
  uint32_t shapes[16][2];
  const int nshapes = tm_train_shapes(m0->wd, m0->state_bits, shapes, 16);
  if (nshapes <= 0) return fail(TM_E_UNSUPPORTED, "batch train: no train tile covers %u words", m0->wd);
  uint32_t maxw = 1;
  for (int i = 0; i < nshapes; i++) maxw = std::max(maxw, shapes[i][1]);
  const uint32_t stride = (m0->wd + maxw - 1) / maxw * maxw;
  const uint32_t *words;
  int rc;
  std::vector<uint32_t> cstrides;     // individualized: per-class row strides and word offsets
  std::vector<uint64_t> woff;
  if (b->ind) {
    // every class's rows once for all trials, each stride a multiple of
    // every candidate tile's words per lane
    size_t extent;
    if ((rc = ind_extent(m0, n, dim_offsets, rows_bytes, "batch train", &extent))) return rc;
    cstrides.resize(m0->classes);
    for (uint32_t h = 0; h < m0->classes; h++) cstrides[h] = (m0->ind[h].wd + maxw - 1) / maxw * maxw;
    if ((rc = stage_rows_ind(m0, rows, n, extent, dim_offsets, cstrides, b->rows, b->words, woff, dev))) return rc;
    HIP_TRY(hipStreamSynchronize(m0->stream));
    words = b->words.as<uint32_t>();
  } else if ((rc = stage_rows(m0, rows, n, dev, stride, b->rows, b->words, &words))) {
    return rc;
  }
  const int64_t *dlabels = nullptr;
  const uint8_t *dcodes = nullptr;
  if (want_type == TM_CLASSIFIER) {
    dlabels = labels;
    if (!dev) {
      if ((rc = b->labels.ensure((size_t) n * 8))) return rc;
      HIP_TRY(hipMemcpyAsync(b->labels.p, labels, (size_t) n * 8, hipMemcpyHostToDevice, b->stream));
      dlabels = b->labels.as<int64_t>();
    }
  } else {
    dcodes = codes;
    if (!dev) {
      if ((rc = b->codes.ensure((size_t) n * m0->class_chunks))) return rc;
      HIP_TRY(hipMemcpyAsync(b->codes.p, codes, (size_t) n * m0->class_chunks, hipMemcpyHostToDevice, b->stream));
      dcodes = b->codes.as<uint8_t>();
    }
  }
  // per-trial sample orders and epoch keys for every epoch
  const size_t per = (size_t) iterations * n;
  if ((rc = b->perm.ensure(per * nt * 4))) return rc;
  if ((rc = b->perm_scratch.ensure(tm_perm_scratch_bytes(n)))) return rc;
  if ((rc = b->ekeys.ensure((size_t) nt * iterations * 16))) return rc;
  if ((rc = b->stats.ensure((size_t) nt * TM_NSTATS * 8))) return rc;
  if ((rc = b->epoch.ensure(256))) return rc;
  if ((rc = b->trials.ensure(nt * sizeof(tm_train_trial)))) return rc;
  std::vector<uint32_t> ek((size_t) nt * iterations * 4);
  std::vector<tm_train_trial> T(nt);
  for (uint32_t t = 0; t < nt; t++) {
    tm_machine *m = b->m[t];
    if (!m->has_state) return fail(TM_E_STATE, "batch train: trial %u has no automata state", t);
    uint32_t *pt = b->perm.as<uint32_t>() + per * t;
    HIP_TRY(tm_launch_perm(m->seed, 0, iterations, n, pt, b->perm_scratch.p, b->perm_scratch.cap, b->stream));
    for (uint32_t e = 0; e < iterations; e++) {
      tm_u4 k = tm_philox(e, 0, 0, TM_TAG_EPOCH, (uint32_t) m->seed, (uint32_t)(m->seed >> 32));
      uint32_t *o = &ek[((size_t) t * iterations + e) * 4];
      o[0] = k.x; o[1] = k.y; o[2] = k.z; o[3] = k.w;
    }
    tm_train_trial &P = T[t];
    memset(&P, 0, sizeof P);
    P.counts = m->counts.as<uint32_t>();
    P.actions = m->actions.as<uint32_t>();
    P.perm = pt;
    P.ekeys = b->ekeys.as<uint32_t>() + (size_t) t * iterations * 4;
    P.stats = nullptr;
    P.perm_epoch0 = 0;
    P.clause_chunks = m->clause_chunks;
    P.tolerance = m->clause_tolerance;
    P.maximum = m->clause_maximum;
    P.target = m->target;
    P.s_draws = (uint32_t)((double)(2u * m->features) / m->specificity);
    P.include_bits = m->include_bits;
    const double q = 1.0 - m->negative;
    P.skip_threshold = !(q > 0) ? 0 : (q >= 1 ? (1ull << 32) : (uint64_t) floor(q * 4294967296.0));
    P.seed_lo = (uint32_t) m->seed;
    P.seed_hi = (uint32_t)(m->seed >> 32);
    P.active = 1;
    P.rows = words;
    P.row_stride = stride;
    P.wd = m->wd;
    P.input_bits = m->input_bits;
    P.last_live = m->last_live_mask();
    P.last_eval = m->last_eval_mask();
    P.class_base = 0;
    P.chunk_base = 0;
  }
  // The launch's table: one entry per plain trial, or per (trial, class) of
  // individualized trials, each class entry as a solo individualized machine
  // trains it (its state offsets, rows, width, masks and random-literal
  // count; its chunks draw as chunks h * clause_chunks + c of the stream).
  std::vector<tm_train_trial> E;
  std::vector<uint32_t> etrial, echunks;
  for (uint32_t t = 0; t < nt; t++) {
    tm_machine *m = b->m[t];
    if (!b->ind) {
      E.push_back(T[t]);
      etrial.push_back(t);
      echunks.push_back(T[t].clause_chunks * m->classes);
      continue;
    }
    for (uint32_t h = 0; h < m->classes; h++) {
      const tm_ind_class &c = m->ind[h];
      tm_train_trial Q = T[t];
      Q.counts = m->counts.as<uint32_t>() + c.coff;
      Q.actions = m->actions.as<uint32_t>() + c.aoff;
      Q.s_draws = (uint32_t)((double)(2 * c.features) / m->specificity);
      Q.rows = words + woff[h];
      Q.row_stride = cstrides[h];
      Q.wd = c.wd;
      Q.input_bits = c.input_bits;
      Q.last_live = live_mask(c.wd, c.input_chunks, c.tail_mask);
      Q.last_eval = eval_mask(c.wd, c.input_bits);
      Q.class_base = h;
      Q.chunk_base = h * m->clause_chunks;
      E.push_back(Q);
      etrial.push_back(t);
      echunks.push_back(m->clause_chunks);
    }
  }
  const size_t ne = E.size();
  if ((rc = b->trials.ensure(ne * sizeof(tm_train_trial)))) return rc;
  HIP_TRY(hipMemcpyAsync(b->ekeys.p, ek.data(), ek.size() * 4, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(tm_launch_zero(b->stats.p, (size_t) nt * TM_NSTATS * 8, b->stream));
  HIP_TRY(tm_launch_zero(b->epoch.p, 256, b->stream));

  tm_train_launch L;
  memset(&L, 0, sizeof L);
  L.trials = b->trials.as<tm_train_trial>();
  L.epoch_dev = b->epoch.as<uint32_t>();
  L.rows = words;
  L.labels = dlabels;
  L.codes = dcodes;
  L.encoder = want_type == TM_ENCODER;
  L.n = n;
  L.wd = m0->wd;
  L.row_stride = stride;
  L.epochs = 1;
  L.class_chunks = m0->class_chunks;
  L.input_bits = m0->input_bits;
  L.state_bits = m0->state_bits;
  L.last_live = m0->last_live_mask();
  L.last_eval = m0->last_eval_mask();
  L.all_ib1 = 1;
  L.all_full = 1;
  for (auto &P : E) {
    L.all_ib1 &= P.include_bits <= 1;
    L.all_full &= full_rows(P, L.wd);
  }

  auto upload_map = [&](uint32_t order) -> int {
    std::vector<uint32_t> map = batch_map(E, echunks, order);
    int r = b->block_map.ensure(map.size() * 4 + 8);
    if (r) return r;
    HIP_TRY(hipMemcpyAsync(b->block_map.p, map.data(), map.size() * 4, hipMemcpyHostToDevice, b->stream));
    L.block_map = b->block_map.as<uint32_t>();
    L.total_blocks = (uint32_t)(map.size() / 2);
    return TM_OK;
  };

  // autotune on epoch 0 (every candidate computes the same epoch-0 state)
  uint64_t nblocks = 0;
  for (uint32_t c : echunks) nblocks += c;
  uint32_t nb_bucket = 1;
  while ((uint64_t) nb_bucket * 2 <= nblocks && nb_bucket < kTuneBlocksCap) nb_bucket *= 2;
  const auto key = std::make_tuple(m0->wd, m0->state_bits, stride, (uint32_t) b->type + (b->ind ? 2u : 0u), nb_bucket);
  bool cached = false;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      uint32_t lanes, wpl, order;
      double ms;
      std::tie(lanes, wpl, order, ms) = it->second;
      // only a tile among this batch's candidates is reused
      for (int i = 0; i < nshapes; i++)
        if (shapes[i][0] == lanes && shapes[i][1] == wpl && stride % wpl == 0) {
          b->lanes = lanes; b->wpl = wpl; b->order = order; b->tuned_ms = ms;
          cached = true;
        }
    }
  }
  if (!cached) {
    HIP_TRY(hipMemcpyAsync(b->trials.p, E.data(), ne * sizeof(tm_train_trial), hipMemcpyHostToDevice, b->stream));
    double best = 1e300;
    auto time_epoch = [&](uint32_t lanes, uint32_t wpl, uint32_t order) -> int {
      L.lanes = lanes;
      L.wpl = wpl;
      HIP_TRY(hipEventRecord(b->ev0, b->stream));
      HIP_TRY(tm_launch_train(L, b->stream));
      HIP_TRY(hipEventRecord(b->ev1, b->stream));
      HIP_TRY(hipStreamSynchronize(b->stream));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
      if (ms < best) { best = ms; b->lanes = lanes; b->wpl = wpl; b->order = order; }
      return TM_OK;
    };
    // every tile in heaviest-first order, then trial-major order for the
    // fastest tile only: the order moves the time by a few percent, the tile
    // by up to several times
    if ((rc = upload_map(1))) return rc;
    for (int i = 0; i < nshapes; i++)
      if ((rc = time_epoch(shapes[i][0], shapes[i][1], 1))) return rc;
    if ((rc = upload_map(0))) return rc;
    if ((rc = time_epoch(b->lanes, b->wpl, 0))) return rc;
    b->tuned_ms = best;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_cache[key] = std::make_tuple(b->lanes, b->wpl, b->order, b->tuned_ms);
  }
  for (size_t e = 0; e < ne; e++) E[e].stats = b->stats.as<unsigned long long>() + (size_t) etrial[e] * TM_NSTATS;
  HIP_TRY(hipMemcpyAsync(b->trials.p, E.data(), ne * sizeof(tm_train_trial), hipMemcpyHostToDevice, b->stream));
  if ((rc = upload_map(b->order))) return rc;
  L.lanes = b->lanes;
  L.wpl = b->wpl;

  // one epoch = [train kernel, epoch advance], captured once and replayed
  batch_release_graph(b);
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int lrc = tm_launch_train(L, b->stream);
  int arc = tm_launch_epoch_advance(b->epoch.as<uint32_t>(), b->stream);
  hipError_t cap = hipStreamEndCapture(b->stream, &b->graph);
  if (lrc || arc || cap != hipSuccess) return fail(TM_E_DEVICE, "batch train: graph capture failed (%d %d %d)", lrc, arc, (int) cap);
  HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));

  std::vector<uint8_t> active(nt, 1), prev(nt, 1);
  for (uint32_t e = 0; e < iterations; e++) {
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->train_ms += ms;
    for (uint32_t t = 0; t < nt; t++) if (active[t]) { b->m[t]->trained = true; b->m[t]->predict_ready = false; }
    if (!each) continue;
    int r = each(user, e + 1, active.data());
    HIP_TRY(hipSetDevice(b->device));
    if (r < 0) { fprintf(stderr, "Error in epoch callback (status %d)\n", r); break; }
    if (r == 0) break;
    bool any = false;
    for (uint32_t t = 0; t < nt; t++) {
      if (active[t] != prev[t]) {
        // every table entry of trial t (its classes, when individualized)
        uint32_t a = active[t] ? 1u : 0u;
        for (size_t e = 0; e < ne; e++)
          if (etrial[e] == t)
            HIP_TRY(hipMemcpyAsync((char *) b->trials.p + e * sizeof(tm_train_trial) +
                                       offsetof(tm_train_trial, active),
                                   &a, 4, hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        prev[t] = active[t];
      }
      any |= active[t] != 0;
    }
    if (!any) break;
  }
  std::vector<uint64_t> st((size_t) nt * TM_NSTATS);
  HIP_TRY(hipMemcpyAsync(st.data(), b->stats.p, st.size() * 8, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (uint32_t t = 0; t < nt; t++) memcpy(b->m[t]->stats, &st[(size_t) t * TM_NSTATS], TM_NSTATS * 8);
  return TM_OK;
}

// feat_offsets (individualized trials): classes + 1 offsets, one feature set
// per class shared by every trial, as tm_create_classifier_ind takes them.
int batch_create(const tm_options *opts, uint32_t ntrials, int type, const int64_t *feat_offsets, uint32_t n_offsets,
                 tm_batch **out)
{
  if (!out) return fail(TM_E_ARG, "batch create: out is NULL");
  *out = nullptr;
  if (!opts || ntrials == 0) return fail(TM_E_ARG, "batch create: need at least one trial");
  for (uint32_t t = 1; t < ntrials; t++)
    if (opts[t].classes != opts[0].classes || opts[t].features != opts[0].features ||
        opts[t].negative != opts[0].negative || opts[t].state != opts[0].state)
      return fail(TM_E_ARG, "batch create: trials must share classes, features, negative and state");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(TM_E_DEVICE, "no HIP device available");
  tm_batch *b = new tm_batch();
  int rc = TM_OK;
  if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    tm_batch_destroy(b);
    return fail(TM_E_DEVICE, "batch create: HIP stream/event creation failed");
  }
  for (uint32_t t = 0; t < ntrials && rc == TM_OK; t++) {
    tm_options o = opts[t];
    o.reusable = 1;
    tm_machine *m = nullptr;
    rc = feat_offsets ? create_ind(&o, feat_offsets, n_offsets, type, &m) : create_common(&o, type, &m, b->stream);
    if (rc == TM_OK) b->m.push_back(m);
  }
  if (rc) { tm_batch_destroy(b); return rc; }
  b->type = type;
  b->ind = feat_offsets != nullptr;
  b->wd = b->m[0]->wd;
  b->input_chunks = b->m[0]->input_chunks;
  *out = b;
  return TM_OK;
}

} // namespace

extern "C" {

int tm_batch_create(const tm_options *opts, uint32_t ntrials, tm_batch **out)
{
  return batch_create(opts, ntrials, TM_CLASSIFIER, nullptr, 0, out);
}

int tm_batch_create_encoder(const tm_options *opts, uint32_t ntrials, tm_batch **out)
{
  return batch_create(opts, ntrials, TM_ENCODER, nullptr, 0, out);
}

int tm_batch_create_classifier_ind(const tm_options *opts, uint32_t ntrials, const int64_t *feat_offsets,
                                   uint32_t n_offsets, tm_batch **out)
{
  if (!feat_offsets) return fail(TM_E_ARG, "batch create: feat_offsets must not be NULL");
  return batch_create(opts, ntrials, TM_CLASSIFIER, feat_offsets, n_offsets, out);
}

int tm_batch_create_encoder_ind(const tm_options *opts, uint32_t ntrials, const int64_t *feat_offsets,
                                uint32_t n_offsets, tm_batch **out)
{
  if (!feat_offsets) return fail(TM_E_ARG, "batch create: feat_offsets must not be NULL");
  return batch_create(opts, ntrials, TM_ENCODER, feat_offsets, n_offsets, out);
}

int tm_batch_machine(tm_batch *b, uint32_t trial, tm_machine **out)
{
  if (!b || !out) return fail(TM_E_ARG, "batch machine: NULL argument");
  if (trial >= b->m.size()) return fail(TM_E_ARG, "batch machine: trial %u out of range", trial);
  *out = b->m[trial];
  return TM_OK;
}

int tm_batch_train_classifier(tm_batch *b, uint32_t n, const uint8_t *problems, const int64_t *solutions,
                              uint32_t iterations, tm_batch_epoch_fn each, void *user)
{
  return batch_train(b, n, problems, 0, nullptr, solutions, nullptr, iterations, each, user, false, TM_CLASSIFIER);
}

int tm_batch_train_classifier_dev(tm_batch *b, uint32_t n, const uint8_t *problems, const int64_t *solutions,
                                  uint32_t iterations, tm_batch_epoch_fn each, void *user)
{
  return batch_train(b, n, problems, 0, nullptr, solutions, nullptr, iterations, each, user, true, TM_CLASSIFIER);
}

int tm_batch_train_encoder(tm_batch *b, uint32_t n, const uint8_t *sentences, const uint8_t *codes,
                           uint32_t iterations, tm_batch_epoch_fn each, void *user)
{
  return batch_train(b, n, sentences, 0, nullptr, nullptr, codes, iterations, each, user, false, TM_ENCODER);
}

int tm_batch_train_encoder_dev(tm_batch *b, uint32_t n, const uint8_t *sentences_dev, const uint8_t *codes_dev,
                               uint32_t iterations, tm_batch_epoch_fn each, void *user)
{
  return batch_train(b, n, sentences_dev, 0, nullptr, nullptr, codes_dev, iterations, each, user, true, TM_ENCODER);
}

int tm_batch_train_classifier_ind(tm_batch *b, uint32_t n, const uint8_t *problems, size_t problems_bytes,
                                  const int64_t *solutions, const int64_t *dim_offsets, uint32_t iterations,
                                  tm_batch_epoch_fn each, void *user)
{
  if (!dim_offsets) return fail(TM_E_ARG, "batch train: dim_offsets must not be NULL");
  return batch_train(b, n, problems, problems_bytes, dim_offsets, solutions, nullptr, iterations, each, user, false,
                     TM_CLASSIFIER);
}

int tm_batch_train_classifier_ind_dev(tm_batch *b, uint32_t n, const uint8_t *problems_dev, size_t problems_bytes,
                                      const int64_t *solutions_dev, const int64_t *dim_offsets, uint32_t iterations,
                                      tm_batch_epoch_fn each, void *user)
{
  if (!dim_offsets) return fail(TM_E_ARG, "batch train: dim_offsets must not be NULL");
  return batch_train(b, n, problems_dev, problems_bytes, dim_offsets, solutions_dev, nullptr, iterations, each, user,
                     true, TM_CLASSIFIER);
}

int tm_batch_train_encoder_ind(tm_batch *b, uint32_t n, const uint8_t *sentences, size_t sentences_bytes,
                               const uint8_t *codes, const int64_t *dim_offsets, uint32_t iterations,
                               tm_batch_epoch_fn each, void *user)
{
  if (!dim_offsets) return fail(TM_E_ARG, "batch train: dim_offsets must not be NULL");
  return batch_train(b, n, sentences, sentences_bytes, dim_offsets, nullptr, codes, iterations, each, user, false,
                     TM_ENCODER);
}

int tm_batch_train_encoder_ind_dev(tm_batch *b, uint32_t n, const uint8_t *sentences_dev, size_t sentences_bytes,
                                   const uint8_t *codes_dev, const int64_t *dim_offsets, uint32_t iterations,
                                   tm_batch_epoch_fn each, void *user)
{
  if (!dim_offsets) return fail(TM_E_ARG, "batch train: dim_offsets must not be NULL");
  return batch_train(b, n, sentences_dev, sentences_bytes, dim_offsets, nullptr, codes_dev, iterations, each, user,
                     true, TM_ENCODER);
}

int tm_batch_accuracy_dev(tm_batch *b, const uint8_t *problems, const int64_t *solutions, uint32_t n,
                          double *accuracy)
{
  if (!b || !accuracy) return fail(TM_E_ARG, "batch accuracy: NULL argument");
  if (b->type != TM_CLASSIFIER) return fail(TM_E_ARG, "batch accuracy: the batch holds encoder trials");
  if (b->ind) return fail(TM_E_ARG, "batch accuracy: individualized trials are scored through tm_batch_machine");
  if (n == 0) { for (size_t t = 0; t < b->m.size(); t++) accuracy[t] = 0; return TM_OK; }
  if (!problems || !solutions) return fail(TM_E_ARG, "batch accuracy: problems and solutions must not be NULL");
  HIP_TRY(hipSetDevice(b->device));
  const uint32_t nt = (uint32_t) b->m.size();
  int rc;
  if ((rc = b->hits.ensure((size_t) nt * 8))) return rc;
  if ((rc = b->pred.ensure((size_t) n * 8))) return rc;
  HIP_TRY(tm_launch_zero(b->hits.p, (size_t) nt * 8, b->stream));
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  for (uint32_t t = 0; t < nt; t++) {
    if ((rc = predict_enqueue_dev(b->m[t], problems, n, b->pred.as<int64_t>()))) return rc;
    HIP_TRY(tm_launch_count_hits(b->pred.as<int64_t>(), solutions, n, b->hits.as<unsigned long long>() + t, b->stream));
  }
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  std::vector<unsigned long long> h(nt);
  HIP_TRY(hipMemcpyAsync(h.data(), b->hits.p, nt * 8, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->eval_ms = ms;
  for (uint32_t t = 0; t < nt; t++) accuracy[t] = (double) h[t] / (double) n;
  return TM_OK;
}

int tm_batch_class_accuracy_dev(tm_batch *b, const uint8_t *problems, const int64_t *solutions, uint32_t n,
                                double *macro)
{
  if (!b || !macro) return fail(TM_E_ARG, "batch class accuracy: NULL argument");
  if (b->type != TM_CLASSIFIER) return fail(TM_E_ARG, "batch class accuracy: the batch holds encoder trials");
  if (b->ind) return fail(TM_E_ARG, "batch class accuracy: individualized trials are scored by tm_batch_class_accuracy_ind_dev");
  const uint32_t nt = (uint32_t) b->m.size(), K = b->m[0]->classes;
  if (n == 0) {
    for (size_t i = 0; i < (size_t) 3 * nt; i++) macro[i] = 0;
    return TM_OK;
  }
  if (!problems || !solutions) return fail(TM_E_ARG, "batch class accuracy: problems and solutions must not be NULL");
  HIP_TRY(hipSetDevice(b->device));
  int rc;
  const size_t per = (size_t) 3 * K;
  if ((rc = b->hits.ensure((size_t) nt * per * 8))) return rc;
  if ((rc = b->pred.ensure((size_t) n * 8))) return rc;
  HIP_TRY(tm_launch_zero(b->hits.p, (size_t) nt * per * 8, b->stream));
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  for (uint32_t t = 0; t < nt; t++) {
    if ((rc = predict_enqueue_dev(b->m[t], problems, n, b->pred.as<int64_t>()))) return rc;
    HIP_TRY(tm_launch_class_counts(b->pred.as<int64_t>(), solutions, n, K,
                                   b->hits.as<unsigned long long>() + t * per, b->stream));
  }
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  std::vector<uint64_t> h((size_t) nt * per);
  HIP_TRY(hipMemcpyAsync(h.data(), b->hits.p, h.size() * 8, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->eval_ms = ms;
  for (uint32_t t = 0; t < nt; t++) tm_finish_class_accuracy(&h[(size_t) t * per], K, nullptr, macro + 3 * t);
  return TM_OK;
}

int tm_batch_class_accuracy_ind_dev(tm_batch *b, const uint8_t *problems, size_t problems_bytes,
                                    const int64_t *solutions, const int64_t *dim_offsets, uint32_t n, double *macro)
{
  if (!b || !macro) return fail(TM_E_ARG, "batch class accuracy: NULL argument");
  if (b->type != TM_CLASSIFIER) return fail(TM_E_ARG, "batch class accuracy: the batch holds encoder trials");
  if (!b->ind) return fail(TM_E_ARG, "batch class accuracy: the batch is not individualized");
  const uint32_t nt = (uint32_t) b->m.size(), K = b->m[0]->classes;
  if (n == 0) {
    for (size_t i = 0; i < (size_t) 3 * nt; i++) macro[i] = 0;
    return TM_OK;
  }
  if (!problems || !solutions || !dim_offsets)
    return fail(TM_E_ARG, "batch class accuracy: problems, solutions and dim_offsets must not be NULL");
  HIP_TRY(hipSetDevice(b->device));
  int rc;
  const size_t per = (size_t) 3 * K;
  if ((rc = b->hits.ensure((size_t) nt * per * 8))) return rc;
  if ((rc = b->pred.ensure((size_t) nt * n * 8))) return rc;
  while (b->tev.size() < nt) {
    hipEvent_t e;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->tev.push_back(e);
  }
  HIP_TRY(tm_launch_zero(b->hits.p, (size_t) nt * per * 8, b->stream));
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  // each trial predicts on its own stream into its own slot; the counts on
  // the batch's stream wait for it: one host synchronisation for all trials
  for (uint32_t t = 0; t < nt; t++) {
    tm_machine *m = b->m[t];
    int64_t *pred = b->pred.as<int64_t>() + (size_t) t * n;
    if ((rc = predict_sums_ind(m, problems, problems_bytes, dim_offsets, n, true))) return rc;
    HIP_TRY(tm_launch_argmax(m->p_sums.as<int64_t>(), n, K, pred, m->stream));
    HIP_TRY(hipEventRecord(b->tev[t], m->stream));
    HIP_TRY(hipStreamWaitEvent(b->stream, b->tev[t], 0));
    HIP_TRY(tm_launch_class_counts(pred, solutions, n, K, b->hits.as<unsigned long long>() + t * per, b->stream));
  }
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  std::vector<uint64_t> h((size_t) nt * per);
  HIP_TRY(hipMemcpyAsync(h.data(), b->hits.p, h.size() * 8, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->eval_ms = ms;
  for (uint32_t t = 0; t < nt; t++) tm_finish_class_accuracy(&h[(size_t) t * per], K, nullptr, macro + 3 * t);
  return TM_OK;
}

int tm_batch_tuning(const tm_batch *b, uint32_t *lanes, uint32_t *wpl, uint32_t *order, double *epoch_ms)
{
  if (!b) return fail(TM_E_ARG, "batch is NULL");
  if (lanes) *lanes = b->lanes;
  if (wpl) *wpl = b->wpl;
  if (order) *order = b->order;
  if (epoch_ms) *epoch_ms = b->tuned_ms;
  return TM_OK;
}

int tm_batch_timing(const tm_batch *b, double *train_ms, double *eval_ms)
{
  if (!b) return fail(TM_E_ARG, "batch is NULL");
  if (train_ms) *train_ms = b->train_ms;
  if (eval_ms) *eval_ms = b->eval_ms;
  return TM_OK;
}

void tm_batch_destroy(tm_batch *b)
{
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  batch_release_graph(b);
  for (tm_machine *m : b->m) tm_destroy(m);
  for (hipEvent_t e : b->tev) hipEventDestroy(e);
  dbuf *bufs[] = { &b->trials, &b->block_map, &b->perm, &b->perm_scratch, &b->ekeys, &b->epoch, &b->stats,
                   &b->rows, &b->words, &b->labels, &b->hits, &b->pred };
  for (dbuf *x : bufs) x->release();
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

} // extern "C"
