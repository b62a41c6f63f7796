// tm_kernels.hip -- gfx950 kernels of the Tsetlin-machine engine.
//
// Train: one workgroup per 8-clause chunk (the unit of independence of the
// reference, capi.c:201-243: feedback reads only the chunk's own 8 votes).
// The chunk's bit-sliced automata stay in VGPRs for the whole launch; lane g of
// clause group j owns words [g*WPL, g*WPL+WPL) of clause j's literal row in
// every plane.  Per sample: AND-NOT + popcount per word, a group reduction of
// the failed-literal count, an LDS exchange of the 8 signed votes, then the
// Type I / Type II feedback on the owning lanes' registers.
//
// Predict: a tiled boolean contraction over (samples x clauses x words) with
// the include plane transposed to [word][clause] so lanes (= clauses) read
// coalesced, and the complemented literal rows staged in LDS and read as
// wave-wide broadcasts.
#include "tm_internal.h"
#include "tm_layout.h"
#include "tm_rng.h"
#include <hipcub/hipcub.hpp>
#include <algorithm>

#include "tm_train_kernel.h"

namespace {

__global__ void setup_kernel(uint32_t *counts, uint32_t *actions, uint64_t ncount_words, uint64_t naction_words)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
  for (uint64_t t = i; t < ncount_words; t += stride) counts[t] = 0xFFFFFFFFu;
  for (uint64_t t = i; t < naction_words; t += stride) actions[t] = 0;
}

__global__ void perm_keys_kernel(uint32_t seed_lo, uint32_t seed_hi, uint32_t epoch, uint32_t n, uint32_t *keys, uint32_t *vals)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = tm_philox(i, epoch, 0, TM_TAG_PERM, seed_lo, seed_hi).x;
  vals[i] = i;
}

// ---------------------------------------------------------------------------
// predict

// actT[w][c] = actions[c][w]; 32x32 tiles through LDS.
__global__ void transpose_kernel(const uint32_t *actions, uint32_t *actT, uint32_t nclauses, uint32_t wd)
{
  __shared__ uint32_t tile[32][33];
  const uint32_t c0 = blockIdx.x * 32, w0 = blockIdx.y * 32;
  const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 256 threads: 32 x 8
  for (uint32_t r = ty; r < 32; r += 8) {
    uint32_t c = c0 + r, w = w0 + tx;
    tile[r][tx] = (c < nclauses && w < wd) ? actions[(size_t) c * wd + w] : 0u;
  }
  __syncthreads();
  for (uint32_t r = ty; r < 32; r += 8) {
    uint32_t w = w0 + r, c = c0 + tx;
    if (c < nclauses && w < wd) actT[(size_t) w * nclauses + c] = tile[tx][r];
  }
}

__global__ void literal_count_kernel(const uint32_t *actions, uint32_t *lit, uint32_t nclauses, uint32_t wd, uint32_t last_eval)
{
  const uint32_t c = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (c >= nclauses) return;
  uint32_t n = 0;
  for (uint32_t w = lane; w < wd; w += 64) n += __popc(actions[(size_t) c * wd + w] & word_mask(w, wd, last_eval));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  if (lane == 0) lit[c] = n;
}

constexpr int PRED_WAVES = 4;    // waves per workgroup
constexpr int PRED_SPW = 16;     // samples per wave
constexpr int PRED_WC = 32;      // words per staged slice
constexpr int PRED_TPG = 4;      // 64-clause tiles per workgroup (a tile group)

// A workgroup evaluates one tile group (up to PRED_TPG consecutive 64-clause
// tiles) against 64 samples.  Each tile's votes are summed per class over its
// 64 lanes (a DPP reduction to a scalar), and the sums of consecutive tiles of
// one class are kept per sample (lane s of the wave holds sample s's), so a
// (sample, class) gets one 64-bit atomic add per tile group rather than one
// per tile.  Four tiles per group: eight halve the atomics again but keep
// twice as many sample tiles' rows in flight per XCD, more than its L2 holds
// beside a 2 MiB clause pass (measured: fetch 1.66 GB against 0.56 GB per
// 20,000-row call at 2048 clauses/class, the same time).
__global__ __launch_bounds__(64 * PRED_WAVES) void predict_kernel(tm_predict_launch L)
{
  __shared__ uint4 xs[PRED_WAVES * PRED_SPW][PRED_WC / 4];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // logical order: passes of `pass_groups` tile groups (about 2 MiB of the
  // include plane); within a pass every sample tile, and within that the
  // pass's groups (their include plane stays in the XCD's L2 while the rows
  // stream past)
  const uint32_t nst = (L.n + PRED_WAVES * PRED_SPW - 1) / (PRED_WAVES * PRED_SPW);
  const uint32_t nct = (L.clauses_total + 63) / 64;
  const uint32_t ngr = (nct + PRED_TPG - 1) / PRED_TPG;
  const uint32_t lb = xcd_logical(blockIdx.x, nst * ngr);
  const uint32_t gp = L.pass_groups, full = ngr / gp, in_full = full * gp * nst;
  uint32_t st_i, gr_i;
  if (lb < in_full) {
    const uint32_t pass = lb / (gp * nst), r = lb % (gp * nst);
    st_i = r / gp;
    gr_i = pass * gp + r % gp;
  } else {
    const uint32_t r = lb - in_full, w = ngr - full * gp;
    st_i = r / w;
    gr_i = full * gp + r % w;
  }
  const uint32_t ct0 = gr_i * PRED_TPG, ct1 = min(nct, ct0 + PRED_TPG);
  const uint32_t s0 = st_i * (PRED_WAVES * PRED_SPW);

  // each thread stages STG words of a slice; the next slice's words (or the
  // next tile's first slice: the rows are the same for every tile) are loaded
  // while this one is evaluated.  Thread tid stages column tid % PRED_WC of
  // rows s0 + tid / PRED_WC + k * RS.  Rows and the include plane are read
  // with buffer loads: a descriptor with a scalar base, a 32-bit lane offset
  // and a scalar offset, where 64-bit per-lane addresses would take two
  // VGPRs per load in flight.  Every offset is clamped into its buffer: a row
  // past the tile's last sample re-reads the last one (its sums are never
  // written), and a word past the row's end is masked out when staged.
  constexpr uint32_t STG = PRED_SPW * PRED_WC / 64, RS = 64 * PRED_WAVES / PRED_WC;
  static_assert(PRED_WAVES * PRED_SPW * PRED_WC == STG * 64 * PRED_WAVES && (64 * PRED_WAVES) % PRED_WC == 0,
                "slice staging");
  const uint32_t wcol = tid % PRED_WC;
  const uint32_t rv = min(L.n - s0, (uint32_t)(PRED_WAVES * PRED_SPW));   // rows of this sample tile
  const __amdgpu_buffer_rsrc_t rrs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(L.rows + (size_t) s0 * L.wd), (short) 0, (int)(rv * L.wd * 4), 0x00020000);
  uint32_t pre[STG];
  auto fetch_slice = [&](uint32_t w0) {
    // the row offsets are recomputed per slice (an empty asm keeps the
    // compiler from holding all of them across the loops)
    uint32_t r0 = tid / PRED_WC;
    asm volatile("" : "+v"(r0));
    const uint32_t wc = min(w0 + wcol, L.wd - 1) * 4;
#pragma unroll
    for (uint32_t k = 0; k < STG; k++)
      pre[k] = __builtin_amdgcn_raw_buffer_load_b32(rrs, min(r0 + k * RS, rv - 1) * (L.wd * 4) + wc, 0, 0);
  };
  fetch_slice(0);

  // running sums for class rc (wave-uniform) of this wave's samples: lane si
  // holds sample si's
  int64_t run = 0;
  uint32_t rc = 0xFFFFFFFFu;
  auto flush = [&]() {
    const uint32_t s = s0 + wave * PRED_SPW + lane;
    if (rc != 0xFFFFFFFFu && lane < PRED_SPW && s < L.n)
      atomicAdd(reinterpret_cast<unsigned long long *>(&L.sums[(size_t) s * L.sums_stride + L.class_offset + rc]),
                (unsigned long long) run);
  };

  for (uint32_t ct = ct0; ct < ct1; ct++) {
    const uint32_t c = ct * 64 + lane;
    const bool cvalid = c < L.clauses_total;
    uint32_t acc[PRED_SPW];
#pragma unroll
    for (int i = 0; i < PRED_SPW; i++) acc[i] = 0;
    for (uint32_t w0 = 0; w0 < L.wd; w0 += PRED_WC) {
      __syncthreads();
      uint32_t *xw = reinterpret_cast<uint32_t *>(&xs[0][0]);
      const uint32_t wm = word_mask(w0 + wcol, L.wd, L.last_eval);
#pragma unroll
      for (uint32_t k = 0; k < STG; k++) xw[tid + k * 64 * PRED_WAVES] = ~pre[k] & wm;
      __syncthreads();
      if (w0 + PRED_WC < L.wd) fetch_slice(w0 + PRED_WC);
      else if (ct + 1 < ct1) fetch_slice(0);
      // the slice's include words: a word past the row re-reads the last
      // one and meets a zero staged word; a clause past the last one reads
      // the last clause's words and votes nothing
      const uint32_t sw = min((uint32_t) PRED_WC, L.wd - w0);   // words of this slice
      const __amdgpu_buffer_rsrc_t ars = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(L.actT + (size_t) w0 * L.clauses_total), (short) 0, (int)(sw * L.clauses_total * 4), 0x00020000);
      const uint32_t cl4 = min(c, L.clauses_total - 1) * 4;
      uint32_t a[PRED_WC];
#pragma unroll
      for (int ww = 0; ww < PRED_WC; ww++)
        a[ww] = __builtin_amdgcn_raw_buffer_load_b32(ars, cl4, min((uint32_t) ww, sw - 1) * L.clauses_total * 4, 0);
#pragma unroll
      for (int si = 0; si < PRED_SPW; si++) {
        const uint32_t sl = wave * PRED_SPW + si;
        // two accumulators keep two independent popcount chains in flight
        uint32_t t0 = acc[si], t1 = 0;
#pragma unroll
        for (int q = 0; q < PRED_WC / 4; q++) {
          const uint4 xv = xs[sl][q];
          t0 = bcnt_acc(a[4 * q + 0] & xv.x, t0);
          t1 = bcnt_acc(a[4 * q + 1] & xv.y, t1);
          t0 = bcnt_acc(a[4 * q + 2] & xv.z, t0);
          t1 = bcnt_acc(a[4 * q + 3] & xv.w, t1);
        }
        acc[si] = t0 + t1;
      }
    }

    // votes (capi.c:110-119) and their class sums (capi.c:928-933); a vote is
    // at most the tolerance (< 2^25, checked on the host), so a tile's 64
    // votes sum exactly in 32 bits
    const uint32_t lit = cvalid ? L.lit[c] : 0u;
    const uint32_t cbase = ct * 64;
    const uint32_t cls_first = cbase / L.clauses;
    const uint32_t cls_last = min(cbase + 63, L.clauses_total - 1) / L.clauses;
    const uint32_t my_cls = c / L.clauses;
    // cap - failed, clamped at 0, signed by polarity; a clause without
    // literals votes the tolerance (its failed count is 0)
    const int32_t cap = !cvalid ? 0 : (int32_t)(lit == 0 ? L.tolerance : (lit < L.tolerance ? lit : L.tolerance));
    const int32_t neg = -(int32_t)(c & 1);
    for (uint32_t k = cls_first; k <= cls_last; k++) {
      if (k != rc) {
        flush();
        rc = k;
        run = 0;
      }
      const bool mine_k = my_cls == k;
      int32_t tile = 0;   // lane si: the tile's sum for sample si
#pragma unroll
      for (int si = 0; si < PRED_SPW; si++) {
        int32_t v = cap - (int32_t) acc[si];
        v = v < 0 || !mine_k ? 0 : v;
        const uint32_t t = group_sum<64>((uint32_t)((v ^ neg) - neg));
        tile = lane == (uint32_t) si ? (int32_t) t : tile;
      }
      run += tile;
    }
  }
  flush();
}

// argmax with strict '>' from -inf: ties go to the lowest class (capi.c:934-942).
__global__ void argmax_kernel(const int64_t *sums, uint32_t n, uint32_t classes, int64_t *out)
{
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t *r = sums + (size_t) s * classes;
  int64_t best = -INT64_MAX;
  uint32_t bc = 0;
  for (uint32_t k = 0; k < classes; k++)
    if (r[k] > best) { best = r[k]; bc = k; }
  out[s] = bc;
}

// bit h = (sum_h > 0), LSB first; padding bits zero (capi.c:984-990).
__global__ void encode_kernel(const int64_t *sums, uint32_t n, uint32_t classes, uint8_t *out)
{
  const uint32_t cc = (classes + 7) / 8;
  const uint64_t total = (uint64_t) n * cc, stride = (uint64_t) gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint32_t s = (uint32_t)(i / cc), b = (uint32_t)(i % cc);
    uint8_t v = 0;
    for (uint32_t t = 0; t < 8; t++) {
      uint32_t k = b * 8 + t;
      if (k < classes && sums[(size_t) s * classes + k] > 0) v |= (uint8_t)(1u << t);
    }
    out[i] = v;
  }
}

__global__ void repack_kernel(const uint8_t *src, uint32_t n, uint32_t bytes, uint32_t wd, uint32_t *dst)
{
  const uint64_t total = (uint64_t) n * wd, stride = (uint64_t) gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint64_t s = i / wd, w = i % wd;
    uint32_t v = 0;
    for (uint32_t b = 0; b < 4; b++) {
      uint64_t o = w * 4 + b;
      if (o < bytes) v |= (uint32_t) src[s * bytes + o] << (8 * b);
    }
    dst[i] = v;
  }
}

__global__ void zero_kernel(uint32_t *p, uint64_t nwords)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
  for (; i < nwords; i += stride) p[i] = 0;
}

__global__ void count_hits_kernel(const int64_t *pred, const int64_t *labels, uint32_t n, unsigned long long *hits)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool hit = i < n && pred[i] == labels[i];
  const uint64_t b = __ballot(hit);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(hits, (unsigned long long) __popcll(b));
}

__global__ void epoch_advance_kernel(uint32_t *epoch)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *epoch += 1;
}

// ---------------------------------------------------------------------------
// step queue of a classifier epoch

constexpr uint32_t kQueueSeg = 1024;   // epoch positions per (chunk, segment) wave

// Whether chunk ch trains on epoch position pos: the sample is of the chunk's
// class, or a negative that the skip draw keeps (the in-kernel scan's test).
// Per epoch position: the inner hash of the skip draw (it depends on the
// sample only) and the sample's label, so the passes over (chunk, position)
// read both in order instead of gathering labels per chunk.
__global__ __launch_bounds__(256) void queue_keys_kernel(tm_queue_launch Q)
{
  const uint32_t pos = blockIdx.x * 256 + threadIdx.x;
  if (pos >= Q.n) return;
  const uint32_t s = Q.perm[pos];
  Q.pos_keys[pos] = tm_lowbias32(s ^ Q.ek0);
  Q.pos_keys[Q.n + pos] = (uint32_t) Q.labels[s];
}

// tm_skip_draw(ek0, ek1, s, ch) from the sample's precomputed inner hash
// key; sc is the sample's class.  An entry: a sample of the chunk's class, or
// a negative that the skip draw keeps.
__device__ __forceinline__ bool queue_keep(const tm_queue_launch &Q, bool valid, uint32_t key, uint32_t sc,
                                           uint32_t ch, bool &tgt)
{
  tgt = valid && sc == ch / Q.clause_chunks;
  return valid && (tgt || (uint64_t) tm_lowbias32(key ^ (ch * TM_GOLDEN + Q.ek1)) >= Q.skip_threshold);
}

// Chunks per wave of the two queue passes: the wave reads a position's key
// and class once for all of them.
constexpr uint32_t kQueueCpw = 8;

// grid (ceil(chunks / kQueueCpw), ceil(nseg / 4)), 4 waves per block, each
// wave one segment of up to kQueueCpw consecutive chunks
__global__ __launch_bounds__(256) void queue_count_kernel(tm_queue_launch Q)
{
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t seg = blockIdx.y * 4 + (threadIdx.x >> 6);
  const uint32_t c0 = blockIdx.x * kQueueCpw;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) Q.counts[(size_t) Q.chunks * Q.nseg] = 0;
  if (seg >= Q.nseg) return;
  const uint32_t p0 = seg * kQueueSeg;
  const uint32_t p1 = (uint32_t) min((uint64_t) Q.n, (uint64_t) p0 + kQueueSeg);
  uint64_t cnt[kQueueCpw] = {};
  for (uint32_t b = p0; b < p1; b += 64) {
    const uint32_t pos = b + lane;
    const bool valid = pos < Q.n;
    const uint32_t key = valid ? Q.pos_keys[pos] : 0u;
    const uint32_t sc = valid ? Q.pos_keys[Q.n + pos] : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t k = 0; k < kQueueCpw; k++) {
      if (c0 + k >= Q.chunks) break;
      bool tgt;
      cnt[k] += __popcll(__ballot(queue_keep(Q, valid, key, sc, c0 + k, tgt)));
    }
  }
  if (lane == 0) {
#pragma unroll
    for (uint32_t k = 0; k < kQueueCpw; k++)
      if (c0 + k < Q.chunks) Q.counts[(size_t)(c0 + k) * Q.nseg + seg] = cnt[k];
  }
}

// Same grid: each wave writes each of its chunks' entries of its segment
// from their offsets on, in epoch order, with the step draws of
// tm_philox(sample, chunk, epoch).  About one position in ten is an entry, so
// the wave compacts each chunk's entries into LDS and computes the draws 64
// at a time with every lane busy (Philox in the branch of the wanted lanes
// would run for nearly every 64 positions).
__device__ __forceinline__ void queue_emit(const tm_queue_launch &Q, uint32_t e, uint64_t idx, uint32_t ch)
{
  Q.qid[idx] = e;
  // the eight clauses' draws: words 0-3 and 4-7 of the entry, plane-major
  const tm_u4 d = tm_philox(e & 0x7FFFFFFFu, ch, Q.epoch, TM_TAG_STEP, Q.seed_lo, Q.seed_hi);
  Q.qdraw[idx] = d.x;
  Q.qdraw[Q.qcap + idx] = d.y;
  Q.qdraw[2 * Q.qcap + idx] = d.z;
  Q.qdraw[3 * Q.qcap + idx] = d.w;
  const tm_u4 d2 = tm_philox(e & 0x7FFFFFFFu, ch, Q.epoch, TM_TAG_STEP + 1, Q.seed_lo, Q.seed_hi);
  Q.qdraw[4 * Q.qcap + idx] = d2.x;
  Q.qdraw[5 * Q.qcap + idx] = d2.y;
  Q.qdraw[6 * Q.qcap + idx] = d2.z;
  Q.qdraw[7 * Q.qcap + idx] = d2.w;
}

__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__global__ __launch_bounds__(256) void queue_write_kernel(tm_queue_launch Q)
{
  __shared__ uint32_t pend_lds[4][kQueueCpw][128];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t seg = blockIdx.y * 4 + wv;
  const uint32_t c0 = blockIdx.x * kQueueCpw;
  if (seg >= Q.nseg) return;
  const uint32_t p0 = seg * kQueueSeg;
  const uint32_t p1 = (uint32_t) min((uint64_t) Q.n, (uint64_t) p0 + kQueueSeg);
  const uint64_t below = (1ull << lane) - 1;
  uint64_t base[kQueueCpw];   // output index of each chunk's pend[0]
  uint32_t np[kQueueCpw];     // entries waiting per chunk (< 64 between rounds)
#pragma unroll
  for (uint32_t k = 0; k < kQueueCpw; k++) {
    base[k] = c0 + k < Q.chunks ? Q.offsets[(size_t)(c0 + k) * Q.nseg + seg] : 0;
    np[k] = 0;
  }
  for (uint32_t b = p0; b < p1; b += 64) {
    const uint32_t pos = b + lane;
    const bool valid = pos < Q.n;
    const uint32_t s = valid ? Q.perm[pos] : 0u;
    const uint32_t key = valid ? Q.pos_keys[pos] : 0u;
    const uint32_t sc = valid ? Q.pos_keys[Q.n + pos] : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t k = 0; k < kQueueCpw; k++) {
      if (c0 + k >= Q.chunks) break;
      uint32_t *pend = pend_lds[wv][k];
      bool tgt;
      const bool w = queue_keep(Q, valid, key, sc, c0 + k, tgt);
      const uint64_t mask = __ballot(w);
      if (w) pend[np[k] + __popcll(mask & below)] = s | (tgt ? 0x80000000u : 0u);
      np[k] += __popcll(mask);
      if (np[k] >= 64) {
        wave_lds_sync();
        const uint32_t e = pend[lane];
        const uint32_t rest = lane + 64 < np[k] ? pend[lane + 64] : 0u;
        wave_lds_sync();
        if (lane + 64 < np[k]) pend[lane] = rest;
        queue_emit(Q, e, base[k] + lane, c0 + k);
        base[k] += 64;
        np[k] -= 64;
      }
    }
  }
  wave_lds_sync();
#pragma unroll
  for (uint32_t k = 0; k < kQueueCpw; k++)
    if (c0 + k < Q.chunks && lane < np[k]) queue_emit(Q, pend_lds[wv][k][lane], base[k] + lane, c0 + k);
}

template <int WPL, int G>
int launch_train_shape(const tm_train_launch &L, hipStream_t st)
{
  const dim3 grid(L.total_blocks), block(8 * G);
  // rows that fill the full-wave tile exactly take the variant without tail
  // masks (include_bits 1 only); up to 4 words per lane it runs at four
  // workgroups per CU
  const bool full = G == 64 && L.all_ib1 && L.all_full && (uint64_t) G * WPL == L.wd;
  constexpr bool R4 = G == 64 && WPL <= 4;   // step-queue mode only
  if (L.qid) {
    if (full && R4) hipLaunchKernelGGL((train_kernel_r4<R4 ? WPL : 1, true>), grid, block, 0, st, L);
    else if (full) hipLaunchKernelGGL((train_kernel<WPL, G, 7, true, true, G == 64>), grid, block, 0, st, L);
    else if (L.all_ib1) hipLaunchKernelGGL((train_kernel<WPL, G, 7, true, true>), grid, block, 0, st, L);
    else hipLaunchKernelGGL((train_kernel<WPL, G, 7, false, true>), grid, block, 0, st, L);
  } else {
    if (full) hipLaunchKernelGGL((train_kernel<WPL, G, 7, true, false, G == 64>), grid, block, 0, st, L);
    else if (L.all_ib1) hipLaunchKernelGGL((train_kernel<WPL, G, 7, true, false>), grid, block, 0, st, L);
    else hipLaunchKernelGGL((train_kernel<WPL, G, 7, false, false>), grid, block, 0, st, L);
  }
  return (int) hipGetLastError();
}

} // namespace

// Train kernels of the other state bits (tm_train_states.hip, one object per
// plane count): scan mode and the general include_bits path.
#define TM_PLANES_DECL(M) int tm_launch_train_planes_##M(const tm_train_launch &L, uint32_t G, uint32_t W, hipStream_t st);
TM_PLANES_DECL(1) TM_PLANES_DECL(2) TM_PLANES_DECL(3) TM_PLANES_DECL(4) TM_PLANES_DECL(5) TM_PLANES_DECL(6)
TM_PLANES_DECL(8) TM_PLANES_DECL(9) TM_PLANES_DECL(10) TM_PLANES_DECL(11) TM_PLANES_DECL(12) TM_PLANES_DECL(13)
TM_PLANES_DECL(14) TM_PLANES_DECL(15)
#undef TM_PLANES_DECL

// ---------------------------------------------------------------------------
// launchers

// Default tile and the autotuner's candidates (tm_layout.h).
int tm_train_shape(uint32_t wd, uint32_t state_bits, uint32_t *lanes, uint32_t *wpl)
{
  return tm_layout::train_shape(wd, state_bits, lanes, wpl);
}

int tm_train_shapes(uint32_t wd, uint32_t state_bits, uint32_t (*out)[2], int cap)
{
  return tm_layout::train_shapes(wd, state_bits, out, cap);
}

int tm_launch_train(const tm_train_launch &L, hipStream_t st)
{
  uint32_t G = L.lanes, W = L.wpl;
  if (!G && tm_train_shape(L.wd, L.state_bits, &G, &W)) return (int) hipErrorInvalidValue;
  if (L.state_bits < 2 || L.state_bits > TM_MAX_STATE_BITS || !tm_layout::tile_fits(W, L.state_bits))
    return (int) hipErrorInvalidValue;
  // L.wd is the widest trial's row; the host checks each trial's stride
  if (L.row_stride % W != 0 || L.row_stride < L.wd || (uint64_t) G * W < L.wd) return (int) hipErrorInvalidValue;
  // one workgroup per chunk: the grid must stay below 2^32 work-items
  if ((uint64_t) L.total_blocks * 8 * G > 0xFFFFFFFFull) return (int) hipErrorInvalidValue;
  // the queue serves one classifier machine of 8 state bits for one epoch
  if (L.qid && (L.trials || L.block_map || L.epoch_dev || L.encoder || L.epochs != 1 || !L.qoff || !L.qdraw ||
                L.state_bits != 8))
    return (int) hipErrorInvalidValue;
  switch (L.state_bits - 1) {
    case 1: return tm_launch_train_planes_1(L, G, W, st);
    case 2: return tm_launch_train_planes_2(L, G, W, st);
    case 3: return tm_launch_train_planes_3(L, G, W, st);
    case 4: return tm_launch_train_planes_4(L, G, W, st);
    case 5: return tm_launch_train_planes_5(L, G, W, st);
    case 6: return tm_launch_train_planes_6(L, G, W, st);
    case 8: return tm_launch_train_planes_8(L, G, W, st);
    case 9: return tm_launch_train_planes_9(L, G, W, st);
    case 10: return tm_launch_train_planes_10(L, G, W, st);
    case 11: return tm_launch_train_planes_11(L, G, W, st);
    case 12: return tm_launch_train_planes_12(L, G, W, st);
    case 13: return tm_launch_train_planes_13(L, G, W, st);
    case 14: return tm_launch_train_planes_14(L, G, W, st);
    case 15: return tm_launch_train_planes_15(L, G, W, st);
    default: break;
  }
  switch (G * 100 + W) {
    case 801: return launch_train_shape<1, 8>(L, st);
    case 802: return launch_train_shape<2, 8>(L, st);
    case 1601: return launch_train_shape<1, 16>(L, st);
    case 1602: return launch_train_shape<2, 16>(L, st);
    case 3201: return launch_train_shape<1, 32>(L, st);
    case 3202: return launch_train_shape<2, 32>(L, st);
    case 6401: return launch_train_shape<1, 64>(L, st);
    case 6402: return launch_train_shape<2, 64>(L, st);
    case 6404: return launch_train_shape<4, 64>(L, st);
    case 6408: return launch_train_shape<8, 64>(L, st);
    case 6416: return launch_train_shape<16, 64>(L, st);
  }
  return (int) hipErrorInvalidValue;
}

int tm_launch_count_hits(const int64_t *pred, const int64_t *labels, uint32_t n, unsigned long long *hits,
                         hipStream_t st)
{
  hipLaunchKernelGGL(count_hits_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pred, labels, n, hits);
  return (int) hipGetLastError();
}

int tm_launch_epoch_advance(uint32_t *epoch_dev, hipStream_t st)
{
  hipLaunchKernelGGL(epoch_advance_kernel, dim3(1), dim3(64), 0, st, epoch_dev);
  return (int) hipGetLastError();
}

int tm_launch_setup(uint32_t *counts, uint32_t *actions, uint64_t nclauses, uint32_t m, uint32_t wd, hipStream_t st)
{
  hipLaunchKernelGGL(setup_kernel, dim3(1024), dim3(256), 0, st, counts, actions,
                     nclauses * m * wd, nclauses * wd);
  return (int) hipGetLastError();
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t) 255; }

static size_t perm_sort_temp_bytes(uint32_t n)
{
  size_t temp = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (uint32_t *) nullptr, (uint32_t *) nullptr,
                                     (uint32_t *) nullptr, (uint32_t *) nullptr, (int) n);
  return align256(temp);
}

// scratch layout: [sort temp storage][keys n][sorted keys n][values n], each 256-aligned
size_t tm_perm_scratch_bytes(uint32_t n)
{
  return perm_sort_temp_bytes(n) + 3 * align256((size_t) n * sizeof(uint32_t));
}

int tm_launch_perm(uint64_t seed, uint32_t epoch_first, uint32_t epochs, uint32_t n, uint32_t *perm_out,
                   void *scratch, size_t scratch_bytes, hipStream_t st)
{
  const size_t temp_bytes = perm_sort_temp_bytes(n);
  const size_t arr = align256((size_t) n * sizeof(uint32_t));
  if (scratch_bytes < temp_bytes + 3 * arr) return (int) hipErrorInvalidValue;
  char *temp = (char *) scratch;
  uint32_t *keys = (uint32_t *)(temp + temp_bytes);
  uint32_t *keys_sorted = (uint32_t *)(temp + temp_bytes + arr);
  uint32_t *vals = (uint32_t *)(temp + temp_bytes + 2 * arr);
  for (uint32_t e = 0; e < epochs; e++) {
    hipLaunchKernelGGL(perm_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, st,
                       (uint32_t) seed, (uint32_t)(seed >> 32), epoch_first + e, n, keys, vals);
    size_t tb = temp_bytes;
    hipError_t err = hipcub::DeviceRadixSort::SortPairs(temp, tb, keys, keys_sorted, vals,
                                                        perm_out + (size_t) e * n, (int) n, 0, 32, st);
    if (err != hipSuccess) return (int) err;
  }
  return (int) hipGetLastError();
}

uint32_t tm_queue_segments(uint32_t n) { return (uint32_t)(((uint64_t) n + kQueueSeg - 1) / kQueueSeg); }

size_t tm_queue_scratch_bytes(uint64_t items)
{
  size_t temp = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, temp, (uint64_t *) nullptr, (uint64_t *) nullptr, (int) items);
  return std::max<size_t>(temp, 256);
}

int tm_launch_queue_count(const tm_queue_launch &Q, void *scratch, size_t scratch_bytes, hipStream_t st)
{
  const uint64_t items = (uint64_t) Q.chunks * Q.nseg + 1;
  if (Q.n == 0 || !Q.pos_keys || Q.nseg != tm_queue_segments(Q.n) || items > 0x7FFFFFFFull ||
      (Q.nseg + 3) / 4 > 65535)
    return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(queue_keys_kernel, dim3((Q.n + 255) / 256), dim3(256), 0, st, Q);
  hipLaunchKernelGGL(queue_count_kernel, dim3((Q.chunks + kQueueCpw - 1) / kQueueCpw, (Q.nseg + 3) / 4), dim3(256), 0,
                     st, Q);
  size_t tb = scratch_bytes;
  hipError_t err = hipcub::DeviceScan::ExclusiveSum(scratch, tb, Q.counts, Q.offsets, (int) items, st);
  if (err != hipSuccess) return (int) err;
  return (int) hipGetLastError();
}

int tm_launch_queue_write(const tm_queue_launch &Q, hipStream_t st)
{
  if (!Q.qid || !Q.qdraw || !Q.pos_keys || Q.nseg != tm_queue_segments(Q.n) || (Q.nseg + 3) / 4 > 65535)
    return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(queue_write_kernel, dim3((Q.chunks + kQueueCpw - 1) / kQueueCpw, (Q.nseg + 3) / 4), dim3(256), 0,
                     st, Q);
  return (int) hipGetLastError();
}

int tm_launch_prepare_predict(const uint32_t *actions, uint32_t *actT, uint32_t *lit, uint32_t nclauses,
                              uint32_t wd, uint32_t last_eval, hipStream_t st)
{
  hipLaunchKernelGGL(transpose_kernel, dim3((nclauses + 31) / 32, (wd + 31) / 32), dim3(256), 0, st,
                     actions, actT, nclauses, wd);
  hipLaunchKernelGGL(literal_count_kernel, dim3((nclauses + 3) / 4), dim3(256), 0, st,
                     actions, lit, nclauses, wd, last_eval);
  return (int) hipGetLastError();
}

int tm_launch_predict(const tm_predict_launch &L, hipStream_t st)
{
  const uint64_t spt = PRED_WAVES * PRED_SPW;            // samples per tile
  const uint64_t nct = (L.clauses_total + 63) / 64;
  const uint64_t ngr = (nct + PRED_TPG - 1) / PRED_TPG;  // tile groups
  // A dispatch's grid is at most 2^32 - 1 work-items: larger sample counts
  // go out as several launches over consecutive sample ranges.
  const uint64_t max_blocks = 0xFFFFFFFFull / (64 * PRED_WAVES);
  // buffer descriptors: a slice of the include plane (32 words of every
  // clause) and a tile's rows must each stay below 2^31 bytes
  if (ngr > max_blocks || L.n == 0 || (uint64_t) L.clauses_total * PRED_WC * 4 > 0x7FFFFFFFull ||
      spt * L.wd * 4 > 0x7FFFFFFFull)
    return (int) hipErrorInvalidValue;
  const uint64_t st_per_launch = max_blocks / ngr;
  tm_predict_launch P = L;
  P.sums_stride = L.sums_stride ? L.sums_stride : L.classes;
  if (P.class_offset + L.classes > P.sums_stride || L.tolerance >= (1u << 25)) return (int) hipErrorInvalidValue;
  // tile groups per pass: about 2 MiB of include plane
  const uint64_t group_bytes = 64ull * PRED_TPG * L.wd * 4;
  P.pass_groups = (uint32_t) std::max<uint64_t>(1, std::min<uint64_t>(ngr, (2ull << 20) / group_bytes));
  for (uint64_t s0 = 0; s0 < L.n; s0 += st_per_launch * spt) {
    const uint64_t n = std::min<uint64_t>(L.n - s0, st_per_launch * spt);
    P.rows = L.rows + s0 * L.wd;
    P.sums = L.sums + s0 * P.sums_stride;
    P.n = (uint32_t) n;
    const uint64_t nst = (n + spt - 1) / spt;
    hipLaunchKernelGGL(predict_kernel, dim3((uint32_t)(nst * ngr)), dim3(64 * PRED_WAVES), 0, st, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int) e;
  }
  return (int) hipSuccess;
}

int tm_launch_argmax(const int64_t *sums, uint32_t n, uint32_t classes, int64_t *out, hipStream_t st)
{
  hipLaunchKernelGGL(argmax_kernel, dim3((n + 255) / 256), dim3(256), 0, st, sums, n, classes, out);
  return (int) hipGetLastError();
}

int tm_launch_encode(const int64_t *sums, uint32_t n, uint32_t classes, uint8_t *out, hipStream_t st)
{
  const uint64_t total = (uint64_t) n * ((classes + 7) / 8);
  const uint32_t blocks = (uint32_t) std::min<uint64_t>((total + 255) / 256, 1u << 16);
  hipLaunchKernelGGL(encode_kernel, dim3(std::max(blocks, 1u)), dim3(256), 0, st, sums, n, classes, out);
  return (int) hipGetLastError();
}

int tm_launch_repack(const uint8_t *src, uint32_t n, uint32_t bytes, uint32_t wd, uint32_t *dst, hipStream_t st)
{
  const uint64_t total = (uint64_t) n * wd;
  const uint32_t blocks = (uint32_t) std::min<uint64_t>((total + 255) / 256, 1u << 16);
  hipLaunchKernelGGL(repack_kernel, dim3(std::max(blocks, 1u)), dim3(256), 0, st, src, n, bytes, wd, dst);
  return (int) hipGetLastError();
}

int tm_launch_zero(void *p, size_t bytes, hipStream_t st)
{
  hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(256), 0, st, (uint32_t *) p, (uint64_t)(bytes / 4));
  return (int) hipGetLastError();
}
