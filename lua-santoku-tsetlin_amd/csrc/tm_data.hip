// tm_data.hip -- device-resident data preparation and dev metrics.
//
// Packers produce the engine's input rows on the device: flip-interleaved,
// ceil(2F/8) bytes per sample, bits 0..F-1 = x, F..2F-1 = not x, LSB first,
// padding bits zero (the layout of santoku-matrix's ivec:bits_to_cvec(n, F,
// true), doc/sth.md:85-91).  The metrics count on the device and finish on the
// host in the reference's arithmetic order (evaluator/capi.c:119-203,
// :497-574), so their doubles are the reference's.
//
// Every entry point runs on the legacy default stream and returns after its
// work is done; inputs produced on other non-blocking streams must be complete
// before the call.
#include "tsetlin_amd.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "tm_internal.h"

namespace {

int fail(int code, const char *fmt, ...)
{
  char msg[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  return tm_set_error(code, msg);
}

#define HIP_TRY(expr)                                                                \
  do {                                                                               \
    hipError_t e_ = (hipError_t)(expr);                                              \
    if (e_ != hipSuccess) return fail(TM_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// A device buffer freed at scope exit.
struct scratch {
  void *p = nullptr;
  ~scratch() { if (p) hipFree(p); }
  hipError_t get(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 4); }
};

uint32_t grid_for(uint64_t items, uint32_t block)
{
  const uint64_t g = (items + block - 1) / block;
  return (uint32_t)(g < 65536 ? (g ? g : 1) : 65536);
}

// Bits of output byte j (of a row) that lie in [lo, hi).
__device__ __forceinline__ uint32_t byte_range_mask(uint64_t j, uint64_t lo, uint64_t hi)
{
  const uint64_t b0 = 8 * j;
  const uint64_t a = lo > b0 ? lo - b0 : 0, b = hi > b0 ? hi - b0 : 0;
  const uint32_t aa = a > 8 ? 8u : (uint32_t) a, bb = b > 8 ? 8u : (uint32_t) b;
  return bb > aa ? (((1u << bb) - 1u) & ~((1u << aa) - 1u)) : 0u;
}

// Rows with no feature present: the not-x half all ones.  One u32 word per
// thread over the word-aligned buffer of `words` words (the row bytes, then
// up to 3 padding bytes that the caller never reads).
__global__ void absent_rows_kernel(uint32_t *out, uint64_t words, uint64_t row_bytes, uint64_t features,
                                   uint64_t total_bytes)
{
  const uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
  for (uint64_t w = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
    uint32_t v = 0;
    for (uint32_t k = 0; k < 4; k++) {
      const uint64_t i = 4 * w + k;
      if (i >= total_bytes) break;
      v |= byte_range_mask(i % row_bytes, features, 2 * features) << (8 * k);
    }
    out[w] = v;
  }
}

// Feature f of sample s present: set x bit f, clear not-x bit F + f.
// Duplicates are idempotent; indices outside [0, n*F) are counted.
__global__ void scatter_bits_kernel(const int64_t *set_bits, uint64_t n_set, uint64_t n, uint64_t features,
                                    uint64_t row_bytes, uint32_t *out, unsigned long long *invalid)
{
  const uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
  uint32_t bad = 0;
  for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n_set; i += stride) {
    const int64_t v = set_bits[i];
    if (v < 0 || (uint64_t) v >= n * features) { bad++; continue; }
    const uint64_t s = (uint64_t) v / features, f = (uint64_t) v % features;
    const uint64_t base = s * row_bytes * 8;
    const uint64_t px = base + f, pn = base + features + f;
    atomicOr(&out[px >> 5], 1u << (px & 31));
    atomicAnd(&out[pn >> 5], ~(1u << (pn & 31)));
  }
  if (bad) atomicAdd(invalid, (unsigned long long) bad);
}

// Flip-interleave packed rows: raw row s is ceil(F/8) bytes, bit f = feature f.
__global__ void flip_interleave_kernel(const uint8_t *raw, uint64_t n, uint64_t features, uint64_t raw_bytes,
                                       uint64_t row_bytes, uint8_t *out)
{
  const uint64_t total = n * row_bytes, stride = (uint64_t) gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint64_t s = i / row_bytes, j = i % row_bytes;
    const uint8_t *r = raw + s * raw_bytes;
    uint32_t v = 0;
    for (uint32_t b = 0; b < 8; b++) {
      const uint64_t pos = 8 * j + b;
      uint32_t bit = 0;
      if (pos < features) bit = (r[pos >> 3] >> (pos & 7)) & 1u;
      else if (pos < 2 * features) bit = ((r[(pos - features) >> 3] >> ((pos - features) & 7)) & 1u) ^ 1u;
      v |= bit << b;
    }
    out[i] = (uint8_t) v;
  }
}

// Per-class TP / FP / FN (evaluator/capi.c:142-154), counted in LDS per
// workgroup when the classes fit, else straight into global memory.
constexpr uint32_t kLdsClasses = 2048;

__global__ void class_counts_kernel(const int64_t *pred, const int64_t *expect, uint64_t n, uint32_t classes,
                                    unsigned long long *tp, unsigned long long *fp, unsigned long long *fn)
{
  __shared__ uint32_t l_tp[kLdsClasses], l_fp[kLdsClasses], l_fn[kLdsClasses];
  const bool lds = classes <= kLdsClasses;
  if (lds)
    for (uint32_t c = threadIdx.x; c < classes; c += blockDim.x) l_tp[c] = l_fp[c] = l_fn[c] = 0;
  __syncthreads();
  const uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    // the reference's unsigned int conversion of each int64
    const uint32_t yp = (uint32_t) pred[i], yt = (uint32_t) expect[i];
    if (yp >= classes || yt >= classes) continue;
    if (yp == yt) {
      if (lds) atomicAdd(&l_tp[yt], 1u); else atomicAdd(&tp[yt], 1ull);
    } else if (lds) {
      atomicAdd(&l_fp[yp], 1u);
      atomicAdd(&l_fn[yt], 1u);
    } else {
      atomicAdd(&fp[yp], 1ull);
      atomicAdd(&fn[yt], 1ull);
    }
  }
  if (!lds) return;
  __syncthreads();
  for (uint32_t c = threadIdx.x; c < classes; c += blockDim.x) {
    if (l_tp[c]) atomicAdd(&tp[c], (unsigned long long) l_tp[c]);
    if (l_fp[c]) atomicAdd(&fp[c], (unsigned long long) l_fp[c]);
    if (l_fn[c]) atomicAdd(&fn[c], (unsigned long long) l_fn[c]);
  }
}

// Per-bit disagreement counts (evaluator/capi.c:521-536): one code byte per
// thread, counts in LDS per workgroup when the bits fit.
constexpr uint32_t kLdsBits = 8192;

__global__ void bit_diff_kernel(const uint8_t *pred, const uint8_t *expect, uint64_t n, uint32_t hidden,
                                unsigned long long *bdiff)
{
  __shared__ uint32_t l_b[kLdsBits];
  const bool lds = hidden <= kLdsBits;
  if (lds)
    for (uint32_t c = threadIdx.x; c < hidden; c += blockDim.x) l_b[c] = 0;
  __syncthreads();
  const uint64_t chunks = (hidden + 7) / 8, total = n * chunks;
  const uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint32_t byte = (uint32_t)(i % chunks);
    uint32_t d = (uint32_t)(pred[i] ^ expect[i]);
    while (d) {
      const uint32_t b = __builtin_ctz(d);
      d &= d - 1;
      const uint32_t j = byte * 8 + b;
      if (j >= hidden) break;
      if (lds) atomicAdd(&l_b[j], 1u); else atomicAdd(&bdiff[j], 1ull);
    }
  }
  if (!lds) return;
  __syncthreads();
  for (uint32_t c = threadIdx.x; c < hidden; c += blockDim.x)
    if (l_b[c]) atomicAdd(&bdiff[c], (unsigned long long) l_b[c]);
}

} // namespace

// evaluator/capi.c:156-173, in its order: counts [tp][fp][fn] per class.
void tm_finish_class_accuracy(const uint64_t *h, uint32_t classes, double *per_class, double *macro)
{
  double pa = 0.0, ra = 0.0, fa = 0.0;
  for (uint32_t c = 0; c < classes; c++) {
    const uint64_t t = h[c], f = h[classes + c], m = h[2 * (size_t) classes + c];
    const double p = (t + f) > 0 ? (double) t / (t + f) : 0.0;
    const double r = (t + m) > 0 ? (double) t / (t + m) : 0.0;
    const double f1 = (p + r) > 0 ? 2.0 * p * r / (p + r) : 0.0;
    if (per_class) { per_class[3 * c] = p; per_class[3 * c + 1] = r; per_class[3 * c + 2] = f1; }
    pa += p;
    ra += r;
    fa += f1;
  }
  if (macro) { macro[0] = pa / classes; macro[1] = ra / classes; macro[2] = fa / classes; }
}

int tm_launch_class_counts(const int64_t *pred, const int64_t *expect, uint32_t n, uint32_t classes,
                           unsigned long long *counts, hipStream_t st)
{
  if (n == 0) return (int) hipSuccess;
  unsigned long long *tp = counts, *fp = tp + classes, *fn = fp + classes;
  hipLaunchKernelGGL(class_counts_kernel, dim3(grid_for(n, 256) < 1024 ? grid_for(n, 256) : 1024), dim3(256), 0, st,
                     pred, expect, (uint64_t) n, classes, tp, fp, fn);
  return (int) hipGetLastError();
}

extern "C" {

int tm_bits_to_cvec_dev(const int64_t *set_bits_dev, uint64_t n_set, uint32_t n, uint32_t features, uint8_t *rows_dev,
                        uint64_t *n_invalid)
{
  if (n_invalid) *n_invalid = 0;
  if (features == 0) return fail(TM_E_ARG, "bits_to_cvec: features must be greater than 0");
  if (n == 0) return TM_OK;
  if (!rows_dev || (n_set && !set_bits_dev)) return fail(TM_E_ARG, "bits_to_cvec: NULL buffer");
  const uint64_t row_bytes = (2ull * features + 7) / 8, total = (uint64_t) n * row_bytes;
  const uint64_t words = (total + 3) / 4;
  // the word atomics need a 4-byte aligned buffer that covers whole words
  const bool direct = ((uintptr_t) rows_dev % 4 == 0) && (total % 4 == 0);
  scratch stage, inv;
  uint32_t *out = (uint32_t *) rows_dev;
  if (!direct) {
    HIP_TRY(stage.get(words * 4));
    out = (uint32_t *) stage.p;
  }
  HIP_TRY(inv.get(8));
  HIP_TRY(hipMemsetAsync(inv.p, 0, 8, 0));
  hipLaunchKernelGGL(absent_rows_kernel, dim3(grid_for(words, 256)), dim3(256), 0, 0, out, words, row_bytes,
                     (uint64_t) features, total);
  HIP_TRY(hipGetLastError());
  if (n_set) {
    hipLaunchKernelGGL(scatter_bits_kernel, dim3(grid_for(n_set, 256)), dim3(256), 0, 0, set_bits_dev, n_set,
                       (uint64_t) n, (uint64_t) features, row_bytes, out, (unsigned long long *) inv.p);
    HIP_TRY(hipGetLastError());
  }
  if (!direct) HIP_TRY(hipMemcpyAsync(rows_dev, out, total, hipMemcpyDeviceToDevice, 0));
  uint64_t bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, inv.p, 8, hipMemcpyDeviceToHost, 0));
  HIP_TRY(hipStreamSynchronize(0));
  if (n_invalid) *n_invalid = bad;
  return TM_OK;
}

int tm_bits_to_cvec_ind_dev(const int64_t *ind_dev, const int64_t *ind_offsets, const int64_t *feat_offsets,
                            uint32_t classes, uint32_t n, uint8_t *bitmap_dev, uint64_t bitmap_bytes,
                            int64_t *dim_offsets, uint64_t *n_invalid)
{
  if (n_invalid) *n_invalid = 0;
  if (classes == 0) return fail(TM_E_ARG, "bits_to_cvec_ind: classes must be greater than 0");
  if (!ind_offsets || !feat_offsets || !dim_offsets) return fail(TM_E_ARG, "bits_to_cvec_ind: NULL offsets");
  if (ind_offsets[0] != 0 || feat_offsets[0] != 0)
    return fail(TM_E_ARG, "bits_to_cvec_ind: offsets must start at 0");
  // class h's block: n rows of ceil(2 k_h / 8) bytes at dim_offsets[h]; in
  // the word-aligned staging buffer it starts at word woff[h]
  std::vector<uint64_t> rb(classes), woff(classes);
  uint64_t total = 0, words = 0;
  for (uint32_t h = 0; h < classes; h++) {
    if (feat_offsets[h + 1] <= feat_offsets[h] || feat_offsets[h + 1] - feat_offsets[h] > 0x7FFFFFFF)
      return fail(TM_E_ARG, "bits_to_cvec_ind: class %u has %lld features", h,
                  (long long) (feat_offsets[h + 1] - feat_offsets[h]));
    if (ind_offsets[h + 1] < ind_offsets[h])
      return fail(TM_E_ARG, "bits_to_cvec_ind: ind_offsets decrease at class %u", h);
    rb[h] = (2ull * (uint64_t) (feat_offsets[h + 1] - feat_offsets[h]) + 7) / 8;
    dim_offsets[h] = (int64_t) total;
    woff[h] = words;
    total += (uint64_t) n * rb[h];
    words += ((uint64_t) n * rb[h] + 3) / 4;
  }
  dim_offsets[classes] = (int64_t) total;
  if (total > bitmap_bytes)
    return fail(TM_E_SIZE, "bits_to_cvec_ind: the bitmap holds %llu bytes, the classes need %llu",
                (unsigned long long) bitmap_bytes, (unsigned long long) total);
  if (n == 0) return TM_OK;
  if (!bitmap_dev || (ind_offsets[classes] && !ind_dev)) return fail(TM_E_ARG, "bits_to_cvec_ind: NULL buffer");
  scratch stage, inv;
  HIP_TRY(stage.get(words * 4));
  HIP_TRY(inv.get(8));
  HIP_TRY(hipMemsetAsync(inv.p, 0, 8, 0));
  uint32_t *out = (uint32_t *) stage.p;
  for (uint32_t h = 0; h < classes; h++) {
    const uint64_t k = (uint64_t) (feat_offsets[h + 1] - feat_offsets[h]), bytes = (uint64_t) n * rb[h];
    const uint64_t bw = (bytes + 3) / 4, cnt = (uint64_t) (ind_offsets[h + 1] - ind_offsets[h]);
    hipLaunchKernelGGL(absent_rows_kernel, dim3(grid_for(bw, 256)), dim3(256), 0, 0, out + woff[h], bw, rb[h], k,
                       bytes);
    HIP_TRY(hipGetLastError());
    if (cnt) {
      hipLaunchKernelGGL(scatter_bits_kernel, dim3(grid_for(cnt, 256)), dim3(256), 0, 0, ind_dev + ind_offsets[h],
                         cnt, (uint64_t) n, k, rb[h], out + woff[h], (unsigned long long *) inv.p);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(bitmap_dev + dim_offsets[h], out + woff[h], bytes, hipMemcpyDeviceToDevice, 0));
  }
  uint64_t bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, inv.p, 8, hipMemcpyDeviceToHost, 0));
  HIP_TRY(hipStreamSynchronize(0));
  if (n_invalid) *n_invalid = bad;
  return TM_OK;
}

int tm_flip_interleave_dev(const uint8_t *raw_dev, uint32_t n, uint32_t features, uint8_t *rows_dev)
{
  if (features == 0) return fail(TM_E_ARG, "flip_interleave: features must be greater than 0");
  if (n == 0) return TM_OK;
  if (!raw_dev || !rows_dev) return fail(TM_E_ARG, "flip_interleave: NULL buffer");
  const uint64_t raw_bytes = ((uint64_t) features + 7) / 8, row_bytes = (2ull * features + 7) / 8;
  hipLaunchKernelGGL(flip_interleave_kernel, dim3(grid_for((uint64_t) n * row_bytes, 256)), dim3(256), 0, 0, raw_dev,
                     (uint64_t) n, (uint64_t) features, raw_bytes, row_bytes, rows_dev);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(0));
  return TM_OK;
}

int tm_class_accuracy_dev(const int64_t *predicted_dev, const int64_t *expected_dev, uint32_t n, uint32_t classes,
                          double *per_class, double *macro)
{
  if (classes == 0) return fail(TM_E_ARG, "class_accuracy: n_classes must be > 0");
  if (n && (!predicted_dev || !expected_dev)) return fail(TM_E_ARG, "class_accuracy: NULL buffer");
  scratch cnt;
  HIP_TRY(cnt.get((size_t) 3 * classes * 8));
  HIP_TRY(hipMemsetAsync(cnt.p, 0, (size_t) 3 * classes * 8, 0));
  unsigned long long *tp = (unsigned long long *) cnt.p, *fp = tp + classes, *fn = fp + classes;
  if (n) {
    hipLaunchKernelGGL(class_counts_kernel, dim3(grid_for(n, 256) < 1024 ? grid_for(n, 256) : 1024), dim3(256), 0, 0,
                       predicted_dev, expected_dev, (uint64_t) n, classes, tp, fp, fn);
    HIP_TRY(hipGetLastError());
  }
  std::vector<uint64_t> h((size_t) 3 * classes);
  HIP_TRY(hipMemcpyAsync(h.data(), cnt.p, h.size() * 8, hipMemcpyDeviceToHost, 0));
  HIP_TRY(hipStreamSynchronize(0));
  tm_finish_class_accuracy(h.data(), classes, per_class, macro);
  return TM_OK;
}

int tm_encoding_accuracy_dev(const uint8_t *predicted_dev, const uint8_t *expected_dev, uint32_t n, uint32_t hidden,
                             double *bits, double *summary)
{
  if (hidden == 0) return fail(TM_E_ARG, "encoding_accuracy: n_hidden must be > 0");
  if (n && (!predicted_dev || !expected_dev)) return fail(TM_E_ARG, "encoding_accuracy: NULL buffer");
  scratch cnt;
  HIP_TRY(cnt.get((size_t) hidden * 8));
  HIP_TRY(hipMemsetAsync(cnt.p, 0, (size_t) hidden * 8, 0));
  if (n) {
    const uint64_t items = (uint64_t) n * ((hidden + 7) / 8);
    hipLaunchKernelGGL(bit_diff_kernel, dim3(grid_for(items, 256) < 1024 ? grid_for(items, 256) : 1024), dim3(256), 0,
                       0, predicted_dev, expected_dev, (uint64_t) n, hidden, (unsigned long long *) cnt.p);
    HIP_TRY(hipGetLastError());
  }
  std::vector<uint64_t> b(hidden);
  HIP_TRY(hipMemcpyAsync(b.data(), cnt.p, b.size() * 8, hipMemcpyDeviceToHost, 0));
  HIP_TRY(hipStreamSynchronize(0));
  // evaluator/capi.c:542-570, in its order; n_samples * n_dims is the
  // reference's unsigned int product
  uint64_t diff_total = 0;
  for (uint32_t j = 0; j < hidden; j++) diff_total += b[j];
  double min_b = 1.0, max_b = 0.0;
  for (uint32_t j = 0; j < hidden; j++) {
    const double t = (double) b[j] / (double) n;
    if (t < min_b) min_b = t;
    if (t > max_b) max_b = t;
    if (bits) bits[j] = t;
  }
  const double mean = (double) diff_total / (double) (uint32_t)(n * hidden);
  double var = 0.0;
  for (uint32_t j = 0; j < hidden; j++) {
    const double ber = (double) b[j] / n;
    var += (ber - mean) * (ber - mean);
  }
  const double sd = hidden > 1 ? sqrt(var / (hidden - 1)) : 0.0;
  if (summary) { summary[0] = 1.0 - mean; summary[1] = 1.0 - min_b; summary[2] = 1.0 - max_b; summary[3] = sd; }
  return TM_OK;
}

} // extern "C"
