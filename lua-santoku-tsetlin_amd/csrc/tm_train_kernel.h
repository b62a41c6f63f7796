// tm_train_kernel.h -- the train kernel template and its device helpers.
//
// Included by tm_kernels.hip (8 state bits, every mode) and by
// tm_train_states.hip (other state bits, one object per plane count).
#pragma once
#include <type_traits>

#include "tm_internal.h"
#include "tm_rng.h"

namespace {

// ---------------------------------------------------------------------------
// helpers

// Sum over the G consecutive lanes of a clause group, returned to every lane
// of the group.  Full waves reduce with DPP row shifts and row broadcasts
// (lane 63 ends with the total); half waves the same, with the broadcast
// kept inside each half (lanes 31 and 63 end with the totals); groups of 16
// and 8 lanes with a DPP butterfly (the other lane of a pair, the other pair
// of a quad, the other quad of a half-row by its mirror, and for 16 the
// other half-row by the row's mirror), which stays in the VALU where a
// shuffle goes through the LDS crossbar and back.
template <int G>
__device__ __forceinline__ uint32_t group_sum(uint32_t v)
{
  if (G == 64) {
    v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xF, 0xF, true);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xF, 0xF, true);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, true);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xF, 0xF, true);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xF, 0xF, true);   // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xF, 0xF, true);   // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
  }
  if (G == 32) {
    v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xF, 0xF, true);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xF, 0xF, true);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, true);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xF, 0xF, true);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xA, 0xF, true);   // row_bcast:15 into rows 1, 3
    const uint32_t lo = __builtin_amdgcn_readlane(v, 31), hi = __builtin_amdgcn_readlane(v, 63);
    return (threadIdx.x & 32) ? hi : lo;
  }
  if (G == 16 || G == 8) {
    v += __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xF, 0xF, true);    // quad_perm [1,0,3,2]: lane ^ 1
    v += __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xF, 0xF, true);    // quad_perm [2,3,0,1]: lane ^ 2
    v += __builtin_amdgcn_update_dpp(0u, v, 0x141, 0xF, 0xF, true);   // row_half_mirror: the other quad
    if (G == 16) v += __builtin_amdgcn_update_dpp(0u, v, 0x140, 0xF, 0xF, true);   // row_mirror: the other 8
    return v;
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// popcount(v) + acc as one v_bcnt_u32_b32.  Left to itself the compiler
// turns a chain of popcount sums into popcounts of zero plus v_add3 trees, an
// extra instruction for every two words.
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t v, uint32_t acc)
{
  uint32_t r;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(v), "v"(acc));
  return r;
}

// Workgroup barrier that orders LDS only: outstanding global loads (the
// prefetched literal row) stay in flight across it.
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int G>
__device__ __forceinline__ bool group_any(bool pred)
{
  uint64_t b = __ballot(pred);
  if (G == 64) return b != 0;
  uint32_t lane = threadIdx.x & 63;
  uint32_t base = lane & ~(uint32_t)(G - 1);
  uint64_t gm = (G == 64) ? ~0ull : ((1ull << G) - 1);
  return ((b >> base) & gm) != 0;
}

// Workgroups are dealt to the XCDs round-robin (workgroup b runs on XCD b % 8),
// each XCD with its own L2.  This bijection on [0, nb) gives XCD x a contiguous
// range of logical indices, so neighbouring tiles (the clause tiles of one
// predict pass) share an L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_logical(uint32_t b, uint32_t nb)
{
  const uint32_t x = b % kXcds, per = nb / kXcds, extra = nb % kXcds;
  return x * per + min(x, extra) + b / kXcds;
}

__device__ __forceinline__ uint32_t word_mask(uint32_t w, uint32_t wd, uint32_t last)
{
  return w + 1 < wd ? 0xFFFFFFFFu : (w + 1 == wd ? last : 0u);
}

// Per literal: every plane one (the +1 bound) / some plane 0..top one (not
// at the -1 bound).
template <int M, int WPL>
__device__ __forceinline__ uint32_t all_ones(const uint32_t (&st)[M + 1][WPL], int k)
{
  uint32_t v = st[0][k];
#pragma unroll
  for (int p = 1; p <= M; p++) v &= st[p][k];
  return v;
}

// some plane FROM..M one: the counter is at least 2^FROM
template <int M, int WPL, int FROM>
__device__ __forceinline__ uint32_t any_one_from(const uint32_t (&st)[M + 1][WPL], int k)
{
  uint32_t v = 0;
#pragma unroll
  for (int p = FROM; p <= M; p++) v |= st[p][k];
  return v;
}

template <int M, int WPL, int TOP>
__device__ __forceinline__ uint32_t any_one(const uint32_t (&st)[M + 1][WPL], int k)
{
  uint32_t v = st[0][k];
#pragma unroll
  for (int p = 1; p <= TOP; p++) v |= st[p][k];
  return v;
}

// +1 (sgn all ones) or -1 (sgn zero) on the literals in m: the carry / borrow
// ripple of automata.h:88-146 / :164-194 / :556-594.  The caller has already
// dropped the literals at the bound (the saturation of those routines), so the
// ripple never leaves plane M.  Branch-free, so groups of one wave may step in
// different directions: two logic ops per plane.
template <int M, int WPL>
__device__ __forceinline__ void ripple(uint32_t (&st)[M + 1][WPL], int k, uint32_t m, uint32_t sgn)
{
  const uint32_t ns = ~sgn;
#pragma unroll
  for (int p = 0; p <= M; p++) {
    const uint32_t s0 = st[p][k];
    // m & (s0 ^ ns) as one three-input op (truth table over m, s0, ns:
    // 0xF0 & (0xCC ^ 0xAA) = 0x60); left to itself the compiler spends an
    // extra xor per plane here.  The carry is taken before the plane is
    // written, so the new plane can take the old one's register: written
    // the other way round, the old plane outlives the new one and the
    // compiler moves every plane back into place after each pass.
    const uint32_t carry = p < M ? __builtin_amdgcn_bitop3_b32(m, s0, ns, 0x60) : 0u;
    st[p][k] = s0 ^ m;
    m = carry;
  }
}

// The same with a two-bit amount: +-(m0 + 2 m1) on each literal, where the
// caller has clamped the amount so that no counter leaves its range.  Plane 1
// takes the amount's second bit in a full adder (borrow: majority of the
// complemented plane, m1 and the incoming borrow).
template <int M, int WPL>
__device__ __forceinline__ void ripple2(uint32_t (&st)[M + 1][WPL], int k, uint32_t m0, uint32_t m1, uint32_t sgn)
{
  const uint32_t ns = ~sgn;
  const uint32_t s0 = st[0][k];
  uint32_t m = __builtin_amdgcn_bitop3_b32(m0, s0, ns, 0x60);
  st[0][k] = s0 ^ m0;
  const uint32_t s1 = st[1][k];
  const uint32_t carry = M > 1 ? __builtin_amdgcn_bitop3_b32(s1 ^ ns, m1, m, 0xE8) : 0u;   // majority
  st[1][k] = __builtin_amdgcn_bitop3_b32(s1, m1, m, 0x96);                                 // xor3
  m = carry;
#pragma unroll
  for (int p = 2; p <= M; p++) {
    const uint32_t sp = st[p][k];
    const uint32_t c = p < M ? __builtin_amdgcn_bitop3_b32(m, sp, ns, 0x60) : 0u;
    st[p][k] = sp ^ m;
    m = c;
  }
}

// include plane AND the top include_bits-1 count planes (automata.h:499-512).
// IB1: include_bits == 1, where "included" is the include plane itself.
template <int M, int WPL, bool IB1>
__device__ __forceinline__ uint32_t included(const uint32_t (&st)[M + 1][WPL], int k, uint32_t ib)
{
  uint32_t v = st[M][k];
  if (IB1) return v;
#pragma unroll
  for (int i = 1; i <= M; i++)
    if ((uint32_t) i < ib) v &= st[M - i][k];
  return v;
}

// Global-memory pointers.  Pointers that reach the kernel inside a struct
// lose their address space, and loads through them become FLAT loads, which
// count against lgkmcnt as well as vmcnt: the LDS-only barrier would then wait
// for them.
template <typename T> using gptr = const __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ gptr<T> global_ptr(const T *p) { return (gptr<T>) p; }
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));

// Load WPL consecutive words of a literal row from word wbase.  Callers
// clamp wbase into the row (the row stride is a multiple of WPL) so the load
// is unconditional; words past the row's end are masked by every use.
template <int WPL>
__device__ __forceinline__ void load_words(const uint32_t *row_generic, uint32_t wbase, uint32_t (&x)[WPL])
{
  const gptr<uint32_t> row = global_ptr(row_generic);
  if constexpr (WPL % 4 == 0) {
#pragma unroll
    for (int k = 0; k < WPL; k += 4) {
      const u32x4 v = *(gptr<u32x4>)(row + wbase + k);
      x[k] = v.x; x[k + 1] = v.y; x[k + 2] = v.z; x[k + 3] = v.w;
    }
  } else if constexpr (WPL == 2) {
    const u32x2 v = *(gptr<u32x2>)(row + wbase);
    x[0] = v.x; x[1] = v.y;
  } else {
#pragma unroll
    for (int k = 0; k < WPL; k++) x[k] = row[wbase + k];
  }
}

// A workgroup-uniform struct copied word by word through readfirstlane: the
// trial's parameters then live in scalar registers (loaded from either the
// trial table or the kernel arguments, the compiler would keep them in VGPRs).
template <typename T>
__device__ __forceinline__ T to_sgpr(const T &v)
{
  static_assert(sizeof(T) % 4 == 0, "32-bit words");
  T out;
  const uint32_t *src = reinterpret_cast<const uint32_t *>(&v);
  uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
  for (int i = 0; i < (int) (sizeof(T) / 4); i++) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return out;
}

// Event counters of a clause group (Type I on a true clause, of which grew,
// Type II, Type I on a false clause, random-literal draws).  Smaller groups
// differ per lane and keep one register per counter.  A full-wave group's
// counters are wave-uniform: in scalar registers (said explicitly; the
// compiler otherwise keeps some in VGPRs), or with IN_LDS in LDS, where every
// lane of the wave writes the same sum to the same slot and no register is
// held.
template <int G, bool IN_LDS>
struct event_counters {
  uint32_t c[5] = { 0, 0, 0, 0, 0 };
  __device__ __forceinline__ void init(unsigned long long *) {}
  __device__ __forceinline__ void add(int i, uint32_t d)
  {
    c[i] = G == 64 ? __builtin_amdgcn_readfirstlane(c[i] + d) : c[i] + d;
  }
  __device__ __forceinline__ unsigned long long get(int i) const { return c[i]; }
};

template <>
struct event_counters<64, true> {
  unsigned long long *slot;     // this clause's 5 counters in LDS, zeroed by init
  __device__ __forceinline__ void init(unsigned long long *s)
  {
    slot = s;
#pragma unroll
    for (int i = 0; i < 5; i++) slot[i] = 0;
  }
  __device__ __forceinline__ void add(int i, uint32_t d) { slot[i] += d; }
  __device__ __forceinline__ unsigned long long get(int i) const { return slot[i]; }
};

// ---------------------------------------------------------------------------
// training

template <int WPL, int G, int M, bool IB1, bool QUEUE, bool FULL, bool CNT_LDS, bool MULTI>
__device__ __forceinline__ void train_body(const tm_train_launch &L)
{
  constexpr int TW = G * WPL;                  // words per clause tile
  const uint32_t tid = threadIdx.x;
  // clause within the chunk; a full-wave group makes it wave-uniform, and
  // saying so keeps every per-clause decision on the scalar unit
  const uint32_t j = G == 64 ? __builtin_amdgcn_readfirstlane(tid / G) : tid / G;
  const uint32_t g = tid % G;                  // lane within the clause group
  const uint32_t lane = tid & 63;
  // Chunks stay dealt round-robin over the XCDs: each XCD then holds a share
  // of every class, which balances the per-class work (a chunk's sample count
  // depends on its class).  Rows are not shared through L2 anyway: every chunk
  // walks the epoch order at its own pace.
  uint32_t trial = 0, ch = blockIdx.x;
  if (L.block_map) {   // workgroup-uniform: held in scalar registers
    trial = __builtin_amdgcn_readfirstlane(L.block_map[2 * blockIdx.x]);
    ch = __builtin_amdgcn_readfirstlane(L.block_map[2 * blockIdx.x + 1]);
  }
  const tm_train_trial P = to_sgpr(L.trials ? L.trials[trial] : L.one);
  if (!P.active) return;
  const uint32_t epoch_first = L.epoch_dev ? *L.epoch_dev : L.epoch_first;
  const bool init = L.epoch_dev ? epoch_first == 0 : L.init != 0;
  const uint32_t cls = P.class_base + ch / P.clause_chunks;
  const uint32_t clause = ch * 8 + j;                 // index into this trial's state
  const uint32_t chr = P.chunk_base + ch;             // chunk and clause in the random stream
  const uint32_t clause_r = chr * 8 + j;
  const uint32_t wbase = g * WPL;
  const uint32_t wd = P.wd;
  // Mask of word k of this lane.  FULL: every trial's row fills the tile
  // exactly (G * WPL words, also the stride between rows) and both last-word
  // masks are all ones, so no word is masked and the masks take no registers.
  auto wmask = [&](int k, uint32_t last) -> uint32_t {
    return FULL ? 0xFFFFFFFFu : word_mask(wbase + k, wd, last);
  };
  uint32_t *counts = P.counts;
  uint32_t *actions = P.actions;

  __shared__ __attribute__((aligned(16))) int32_t votes_lds[2][8];
  // bits of the literals hit by random-literal draws, per clause; with
  // MULTI, on three levels: hit once (0), a second time (1), a third (2)
  constexpr int HL = MULTI ? 3 : 1;
  __shared__ uint32_t hits[HL][8][TW];

  for (uint32_t i = tid; i < HL * 8 * TW; i += 8 * G) (&hits[0][0][0])[i] = 0;

  uint32_t st[M + 1][WPL];
#pragma unroll
  for (int k = 0; k < WPL; k++) {
    uint32_t w = wbase + k;
    if (init) {
#pragma unroll
      for (int p = 0; p < M; p++) st[p][k] = 0xFFFFFFFFu;
      st[M][k] = 0;
    } else if (FULL || w < wd) {
#pragma unroll
      for (int p = 0; p < M; p++) st[p][k] = counts[((size_t) clause * M + p) * wd + w];
      st[M][k] = actions[(size_t) clause * wd + w];
    } else {
#pragma unroll
      for (int p = 0; p <= M; p++) st[p][k] = 0;
    }
  }

  auto recount = [&]() -> uint32_t {
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < WPL; k++) c = bcnt_acc(st[M][k] & wmask(k, P.last_eval), c);
    return group_sum<G>(c);
  };
  uint32_t lit = recount();
  lds_barrier();

  const uint32_t T = P.target;
  const uint32_t tol = P.tolerance;
  // votes = max(0, cap - failed): a clause without literals fails nothing and
  // votes tol, else min(literals, tol) - failed (capi.c:110-119); cap only
  // changes with the literal count
  auto cap_of = [&](uint32_t l) -> int32_t { return (int32_t)(l == 0 ? tol : (l < tol ? l : tol)); };
  int32_t cap = cap_of(lit);
  const int32_t vneg = -(int32_t)(j & 1);      // odd clauses vote against
  const uint32_t T2 = 2 * T;
  // lanes past the row's end read its last tile instead (masked by every use)
  const uint32_t wload = FULL ? wbase : min(wbase, P.row_stride - WPL);
  // byte offsets into votes_lds of this step's slot for clause j and of the
  // step's first slot; both alternate between the two halves
  uint32_t vw = j * 4, vr = 0;
  uint32_t n_steps = 0;
  __shared__ unsigned long long ev_lds[CNT_LDS ? 8 : 1][5];
  event_counters<G, CNT_LDS> cnt;
  cnt.init(ev_lds[CNT_LDS ? j : 0]);

  for (uint32_t e = 0; e < L.epochs; e++) {
    const uint32_t epoch = epoch_first + e;
    const uint32_t er = epoch - P.perm_epoch0;
    const gptr<uint32_t> perm = global_ptr(P.perm + (size_t) er * L.n);
    const uint32_t ek0 = P.ekeys[4 * er + 0], ek1 = P.ekeys[4 * er + 1];
    const uint32_t ek2 = P.ekeys[4 * er + 2], ek3 = P.ekeys[4 * er + 3];

    // Samples come in blocks of 64, one per lane: lane l holds the sample id,
    // its target flag and its step-draw words.  `want` marks the lanes this
    // chunk trains on.  Without a queue the chunk scans the epoch order itself
    // (negative sampling applied here); with one, a block is the next 64 queue
    // entries, loaded one block ahead.
    uint32_t b0 = 0, bs = 0;
    uint64_t want = 0, tmask = 0;
    tm_u4 bd = { 0, 0, 0, 0 }, bd2 = { 0, 0, 0, 0 };   // smaller groups: all 8 clauses' draws
    uint32_t bsel = 0;                 // full-wave group: this clause's draw word
    uint64_t qb = 0, qe = 0;           // queue entries not yet loaded
    uint32_t qi_n = 0, qd_n = 0;       // prefetched block
    tm_u4 qd4_n = { 0, 0, 0, 0 }, qd4b_n = { 0, 0, 0, 0 };
    auto qload = [&]() {              // called with qb < qe
      // lanes past the chunk's last entry re-read it: the load stays
      // unconditional, so it can land in place without a wait
      const uint64_t idx = min(qb + lane, qe - 1);
      qi_n = L.qid[idx];
      if constexpr (G == 64) {
        qd_n = L.qdraw[(uint64_t) j * L.qcap + idx];
      } else {
        qd4_n.x = L.qdraw[idx];
        qd4_n.y = L.qdraw[L.qcap + idx];
        qd4_n.z = L.qdraw[2 * L.qcap + idx];
        qd4_n.w = L.qdraw[3 * L.qcap + idx];
        qd4b_n.x = L.qdraw[4 * L.qcap + idx];
        qd4b_n.y = L.qdraw[5 * L.qcap + idx];
        qd4b_n.z = L.qdraw[6 * L.qcap + idx];
        qd4b_n.w = L.qdraw[7 * L.qcap + idx];
      }
    };
    if constexpr (QUEUE) {
      qb = L.qoff[(size_t) ch * L.qseg];
      qe = L.qoff[(size_t)(ch + 1) * L.qseg];
      if (qb < qe) qload();
    }
    auto scan = [&]() -> bool {       // advance to the next block with work
      if constexpr (QUEUE) {
        if (qb >= qe) return false;
        const uint64_t left = qe - qb;
        want = left >= 64 ? ~0ull : ((1ull << left) - 1);
        bs = qi_n;
        asm volatile("" : "+v"(bs));
        if constexpr (G == 64) {
          bsel = qd_n;
          asm volatile("" : "+v"(bsel));
        } else {
          bd = qd4_n;
          bd2 = qd4b_n;
        }
        qb += 64;
        if (qb < qe) qload();
        return true;
      } else {
        while (b0 < L.n) {
          const uint32_t pos = b0 + lane;
          const bool valid = pos < L.n;
          const uint32_t s = valid ? perm[pos] : 0u;
          bool tgt, w;
          if (L.encoder) {
            tgt = valid && ((L.codes[(size_t) s * L.class_chunks + (cls >> 3)] >> (cls & 7)) & 1);
            w = valid;
          } else {
            const uint32_t sc = valid ? (uint32_t) L.labels[s] : 0xFFFFFFFFu;
            tgt = valid && sc == cls;
            w = valid && (tgt || (uint64_t) tm_skip_draw(ek0, ek1, s, chr) >= P.skip_threshold);
          }
          b0 += 64;
          want = __ballot(w);
          if (want) {
            tmask = __ballot(tgt);
            bs = s;
            if constexpr (G == 64) {
              // a full-wave group draws only its own clause's Philox block
              bd = tm_philox(s, chr, epoch, TM_TAG_STEP + (j >> 2), P.seed_lo, P.seed_hi);
              bsel = (j & 3) == 0 ? bd.x : (j & 3) == 1 ? bd.y : (j & 3) == 2 ? bd.z : bd.w;
            } else {
              bd = tm_philox(s, chr, epoch, TM_TAG_STEP, P.seed_lo, P.seed_hi);
              bd2 = tm_philox(s, chr, epoch, TM_TAG_STEP + 1, P.seed_lo, P.seed_hi);
            }
            return true;
          }
        }
        return false;
      }
    };
    // take the next wanted sample of the current block (want != 0)
    uint32_t nq_s = 0, nq_d = 0;
    bool nq_t = false;
    auto take = [&]() {
      const int q = __builtin_ctzll(want);
      want &= ~(1ull << q);
      if constexpr (QUEUE) {
        const uint32_t v = __builtin_amdgcn_readlane(bs, q);
        nq_s = v & 0x7FFFFFFFu;
        nq_t = v >> 31;
      } else {
        nq_s = __builtin_amdgcn_readlane(bs, q);
        nq_t = (tmask >> q) & 1;
      }
      if constexpr (G == 64) {
        nq_d = __builtin_amdgcn_readlane(bsel, q);
      } else {
        // lane q's eight draws, then this lane's clause's (j differs per lane)
        const uint32_t d0 = __builtin_amdgcn_readlane(bd.x, q), d1 = __builtin_amdgcn_readlane(bd.y, q);
        const uint32_t d2 = __builtin_amdgcn_readlane(bd.z, q), d3 = __builtin_amdgcn_readlane(bd.w, q);
        const uint32_t d4 = __builtin_amdgcn_readlane(bd2.x, q), d5 = __builtin_amdgcn_readlane(bd2.y, q);
        const uint32_t d6 = __builtin_amdgcn_readlane(bd2.z, q), d7 = __builtin_amdgcn_readlane(bd2.w, q);
        const uint32_t lo = (j & 3) == 0 ? d0 : (j & 3) == 1 ? d1 : (j & 3) == 2 ? d2 : d3;
        const uint32_t hi = (j & 3) == 0 ? d4 : (j & 3) == 1 ? d5 : (j & 3) == 2 ? d6 : d7;
        nq_d = (j >> 2) ? hi : lo;
      }
    };
    uint32_t xn[WPL];
    // FULL rows are G * WPL words apart (a power of two): a shift, not a multiply
    auto fetch = [&](uint32_t sample) {
      const size_t off = FULL ? (size_t) sample * (G * WPL) : (size_t) sample * P.row_stride;
      load_words<WPL>(P.rows + off, wload, xn);
    };

    if (!scan()) continue;
    take();
    fetch(nq_s);
    for (;;) {
      const uint32_t sq = nq_s, u32 = nq_d;
      const bool tq = nq_t;
      // x is defined here, by an empty asm on the copy: without it the
      // compiler renames x onto the prefetch registers across its loop
      // rotation and then waits for the next row's load (vmcnt(0)) before
      // this step's evaluation, exposing the prefetch latency every step
      uint32_t x[WPL];
#pragma unroll
      for (int k = 0; k < WPL; k++) {
        x[k] = xn[k];
        asm volatile("" : "+v"(x[k]));
      }
      // Wave priorities (measured orderings): on the four-workgroup tile
      // (CNT_LDS) a step's evaluation, up to the vote barrier, issues first
      // (2), since every chunk's next step waits on all eight evaluations;
      // then a clause applying random-literal decrements (1), usually the
      // one its chunk waits for; then the rest (0).  With two workgroups per
      // CU (the wider tiles) a prioritised evaluation measured slower, and
      // only the random-literal clause is raised (2).
      if constexpr (CNT_LDS) __builtin_amdgcn_s_setprio(2);
      // find the next sample and start loading its row before this one's work
      // (the common case, a block with samples left, takes the next one
      // without a branch on the scan's result)
      bool more = true;
      if (want) {
        take();
        fetch(nq_s);
      } else if ((more = scan())) {
        take();
        fetch(nq_s);
      }

      uint32_t f = 0;
#pragma unroll
      for (int k = 0; k < WPL; k++) f = bcnt_acc(st[M][k] & ~x[k] & wmask(k, P.last_eval), f);
      f = group_sum<G>(f);
      int32_t votes = cap - (int32_t) f;
      if (votes < 0) votes = 0;
      // a full-wave group's vote is the same on every lane: all lanes store
      // it (no exec-mask branch); smaller groups store from their lane 0
      char *const vbase = reinterpret_cast<char *>(&votes_lds[0][0]);
      if (G == 64 || g == 0) *reinterpret_cast<int32_t *>(vbase + vw) = (votes ^ vneg) - vneg;
      lds_barrier();
      if constexpr (CNT_LDS) __builtin_amdgcn_s_setprio(0);
      // the step's eight votes as two 16-byte reads (vr is a multiple of 32)
      const i32x4 va = *reinterpret_cast<const i32x4 *>(vbase + vr);
      const i32x4 vc = *reinterpret_cast<const i32x4 *>(vbase + vr + 16);
      int32_t v = __builtin_amdgcn_readfirstlane(va.x + va.y + va.z + va.w + vc.x + vc.y + vc.z + vc.w);
      vw ^= 32;
      vr ^= 32;
      if (v > (int32_t) T) v = (int32_t) T;
      if (v < -(int32_t) T) v = -(int32_t) T;
      // chance((T -+ v) / 2T): fire iff u32 * 2T < (T -+ v) * 2^32, which for
      // integers holds exactly when the product's high word is below T -+ v
      // (2T < 2^32: the host keeps T below 2^30)
      const uint32_t num = (uint32_t)(tq ? (int32_t) T - v : (int32_t) T + v);
      const bool fire = __umulhi(u32, T2) < num;
      n_steps++;
      if (fire) {
        const bool type1 = tq ^ (j & 1);
        const bool out = votes > 0;
        // kind 1: Type I on a true clause (capi.c:179-183): +1 on true
        //   literals when literals <= clause_maximum, -1 on false literals
        //   that are not included (the sets are disjoint, so "included" read
        //   before the +1 is the same as after it);
        // kind 2: Type I on a false clause (capi.c:184-192): s_draws random
        //   single-literal decrements;
        // kind 3: Type II on a true clause (capi.c:194-197): +1 on false
        //   literals that are not included.
        const int kind = type1 ? (out ? 1 : 2) : (out ? 3 : 0);
        if (kind != 0) {
          const bool grow = kind == 1 && lit <= P.maximum;
          uint32_t kev = 0, nd = 0, base = 0, pend = 0;
          if (kind == 1) {
            cnt.add(0, 1);
            if (grow) cnt.add(1, 1);
          } else if (kind == 3) cnt.add(2, 1);
          else {
            // The draws are spread over the group's lanes (draw r on lane r % G).
            // Saturating decrements commute, so each round applies one
            // decrement to every distinct literal hit by the draws still
            // pending; a draw whose literal was already hit this round waits
            // for the next round.
            kev = tm_literal_key(ek2, ek3, sq, clause_r);
            nd = (P.s_draws + G - 1) / G;
            // raised until its decrements are applied (see the loop head)
            __builtin_amdgcn_s_setprio(CNT_LDS ? 1 : 2);
            cnt.add(3, 1);
            cnt.add(4, P.s_draws);
          }
          bool again, load = true;   // load: the group starts the batch of draws at `base`
          uint32_t phase = 0;
          do {
            uint32_t mk[WPL], m1[MULTI ? WPL : 1], sgn;
            if constexpr (MULTI) {
#pragma unroll
              for (int k = 0; k < WPL; k++) m1[k] = 0u;
            }
            if (kind == 2) {
              uint32_t levels = 1;     // hit levels this pass applies
              if (load) {
                // first pass over the batch: every lane's draws in order
                // (a trip count uniform over the group), the hash input
                // stepping by G*phi; a draw whose literal is already hit
                // this round stays pending
                const uint32_t cnt = min(32u, nd - base);
                uint32_t r = base * G + g, hin = kev + r * TM_GOLDEN;
                const uint32_t hstep = G * TM_GOLDEN;
                // every lane's draw exists in the first `full` iterations
                // (base * G < s_draws); only the batch's last may be short
                const uint32_t full = min(cnt, (P.s_draws - base * G) / G);
                uint32_t i = 0;
                // four draws at a time: their hashes first, then their four
                // atomics in flight together, then the hits they found.  With
                // input_bits = 2^b the draw mulhi(h, 2^b) is h >> (32 - b),
                // a full-rate shift instead of a quarter-rate multiply, and
                // the repeat test is one bit extract.  The four-workgroup
                // tile (CNT_LDS) keeps the multiply and the compare: a second
                // copy of the loop or the extract take it past 64 VGPRs.
                constexpr bool LEAN = !CNT_LDS;
                auto batch = [&](auto pow2) {
                  constexpr bool POW2 = decltype(pow2)::value;
                  const uint32_t sh = POW2 ? __builtin_clz(P.input_bits) + 1 : 0u;
                  for (; i + 4 <= full; i += 4) {
                    uint32_t lb[4], old[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                      const uint32_t h = tm_lowbias32(hin + u * hstep);
                      lb[u] = POW2 ? h >> sh : tm_mulhi(h, P.input_bits);
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) old[u] = atomicOr(&hits[0][j][lb[u] >> 5], 1u << (lb[u] & 31));
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                      if constexpr (LEAN) pend |= __builtin_amdgcn_ubfe(old[u], lb[u] & 31, 1) << (i + u);
                      else if (old[u] & (1u << (lb[u] & 31))) pend |= 1u << (i + u);
                    }
                    hin += 4 * hstep;
                  }
                };
                if (LEAN && (P.input_bits & (P.input_bits - 1)) == 0) batch(std::integral_constant<bool, LEAN>());
                else batch(std::false_type());
                for (; i < full; i++) {
                  const uint32_t lbit = tm_mulhi(tm_lowbias32(hin), P.input_bits);
                  const uint32_t hb = 1u << (lbit & 31);
                  const uint32_t old = atomicOr(&hits[0][j][lbit >> 5], hb);
                  if (old & hb) pend |= 1u << i;
                  hin += hstep;
                }
                r += full * G;
                for (; i < cnt; i++) {
                  const bool valid = r < P.s_draws;
                  const uint32_t lbit = tm_mulhi(tm_lowbias32(hin), P.input_bits);
                  const uint32_t hb = valid ? 1u << (lbit & 31) : 0u;
                  const uint32_t old = atomicOr(&hits[0][j][lbit >> 5], hb);
                  if (old & hb) pend |= 1u << i;
                  r += G;
                  hin += hstep;
                }
                load = false;
                if constexpr (MULTI) {
                  // a draw whose literal was already hit is drawn again onto
                  // the next level; one that finds all three levels hit
                  // stays pending for a round of its own
                  while (levels < HL && group_any<G>(pend != 0)) {
                    uint32_t pp = pend;
                    while (pp) {
                      const uint32_t i = __builtin_ctz(pp);
                      pp &= pp - 1;
                      const uint32_t lbit = tm_literal_draw(kev, (base + i) * G + g, P.input_bits);
                      const uint32_t hb = 1u << (lbit & 31);
                      if (!(atomicOr(&hits[levels][j][lbit >> 5], hb) & hb)) pend &= ~(1u << i);
                    }
                    levels++;
                  }
                }
              } else {
                uint32_t pp = pend;
                while (pp) {
                  const uint32_t i = __builtin_ctz(pp);
                  pp &= pp - 1;
                  const uint32_t lbit = tm_literal_draw(kev, (base + i) * G + g, P.input_bits);
                  const uint32_t hb = 1u << (lbit & 31);
                  const uint32_t old = atomicOr(&hits[0][j][lbit >> 5], hb);
                  if (!(old & hb)) pend &= ~(1u << i);
                }
              }
              __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup", "local");
              if constexpr (MULTI) {
                // levels nest, so a literal was hit h0 + h1 + h2 times: a
                // two-bit count (k1 k0), clamped to the counter (a counter
                // below 4 is the two bits c1 c0) so that the subtraction
                // saturates at zero.  Each level is read and cleared in one
                // exchange per word; the level count is uniform, so the
                // branch on it stays outside the words.
                auto count = [&](auto nlev) {
                  constexpr uint32_t NL = decltype(nlev)::value;
                  // every level's words are read and cleared first, so the
                  // exchanges are in flight together
                  uint32_t hl[NL][WPL];
#pragma unroll
                  for (uint32_t l = 0; l < NL; l++) {
#pragma unroll
                    for (int k = 0; k < WPL; k++) hl[l][k] = atomicExch(&hits[l][j][wbase + k], 0u);
                  }
#pragma unroll
                  for (int k = 0; k < WPL; k++) {
                    const uint32_t h0 = hl[0][k];
                    const uint32_t h1 = NL > 1 ? hl[NL > 1 ? 1 : 0][k] : 0u;
                    const uint32_t h2 = NL > 2 ? hl[NL > 2 ? 2 : 0][k] : 0u;
                    const uint32_t k0 = h0 ^ h1 ^ h2, k1 = h1;
                    const uint32_t c0 = st[0][k], c1 = st[1][k];
                    const uint32_t hi = M > 1 ? any_one_from<M, WPL, 2>(st, k) : 0u;
                    // count above counter: k1 > c1, or k1 == c1 and k0 > c0
                    const uint32_t gt = ~hi & ((k1 & ~c1) | (~(k1 ^ c1) & k0 & ~c0));
                    mk[k] = (gt & c0) | (~gt & k0);
                    m1[k] = (gt & c1) | (~gt & k1);
                  }
                };
                if (levels > 2) count(std::integral_constant<uint32_t, 3>());
                else if (levels > 1) count(std::integral_constant<uint32_t, 2>());
                else count(std::integral_constant<uint32_t, 1>());
              } else {
                // read and clear every word first, so the exchanges are in
                // flight together
#pragma unroll
                for (int k = 0; k < WPL; k++) mk[k] = atomicExch(&hits[0][j][wbase + k], 0u);
#pragma unroll
                for (int k = 0; k < WPL; k++) mk[k] &= any_one<M, WPL, M>(st, k);
              }
              __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup", "local");
              sgn = 0u;
            } else if (kind == 1 && grow && phase == 0) {
              // +1 on the true literals first, as tk_automata_inc precedes
              // tk_automata_dec_not_excluded (capi.c:181-182)
#pragma unroll
              for (int k = 0; k < WPL; k++)
                mk[k] = x[k] & ~all_ones<M, WPL>(st, k) & wmask(k, P.last_live);
              sgn = ~0u;
            } else if (kind == 3) {
              // +1 on false literals that are not included; none is at the
              // bound, since a counter with every plane one is included
#pragma unroll
              for (int k = 0; k < WPL; k++)
                mk[k] = ~x[k] & ~included<M, WPL, IB1>(st, k, P.include_bits) & wmask(k, P.last_live);
              sgn = ~0u;
            } else {
              // -1 on false literals that are not included and not at zero
              // (with include_bits 1 their action plane is zero already)
#pragma unroll
              for (int k = 0; k < WPL; k++)
                mk[k] = ~x[k] & ~included<M, WPL, IB1>(st, k, P.include_bits) & wmask(k, P.last_live) &
                        any_one<M, WPL, IB1 ? M - 1 : M>(st, k);
              sgn = 0u;
            }
            // one update site for every kind keeps the state in place (branch
            // arms that each rewrite it make the compiler copy all of it)
#pragma unroll
            for (int k = 0; k < WPL; k++) {
              if constexpr (MULTI) ripple2<M, WPL>(st, k, mk[k], m1[k], sgn);
              else ripple<M, WPL>(st, k, mk[k], sgn);
            }
            again = false;
            if (kind == 2) {
              // every lane of the group leaves a batch together, so no lane
              // reloads draws that were already applied
              const bool pending = group_any<G>(pend != 0);
              if (!pending) {
                base += 32;
                load = true;
              }
              again = pending || base < nd;
            } else if (kind == 1 && grow && phase == 0) {
              again = true;
            }
            phase++;
          } while (again);
          // with include_bits 1, a Type I step without growth only lowers
          // counters whose action bit is zero: the literal count stands
          if (!(IB1 && kind == 1 && !grow)) {
            lit = recount();
            cap = cap_of(lit);
          }
          if (kind == 2) __builtin_amdgcn_s_setprio(0);
        }
      }
      if (!more) break;
    }
  }

  if (P.stats && g == 0) {
    if (j == 0) atomicAdd(&P.stats[0], (unsigned long long) n_steps);
#pragma unroll
    for (int i = 0; i < 5; i++) atomicAdd(&P.stats[1 + i], cnt.get(i));
  }

#pragma unroll
  for (int k = 0; k < WPL; k++) {
    const uint32_t w = wbase + k;
    if (FULL || w < wd) {
#pragma unroll
      for (int p = 0; p < M; p++) counts[((size_t) clause * M + p) * wd + w] = st[p][k];
      actions[(size_t) clause * wd + w] = st[M][k];
    }
  }
}

template <int WPL, int G, int M, bool IB1, bool QUEUE, bool FULL = false>
__global__ __launch_bounds__(8 * G, (WPL == 8 && IB1 ? 4 : 1)) void train_kernel(tm_train_launch L)
{
  // the multi-level draws' second amount register per word fits where the
  // state takes at most 32 VGPRs (64 without tail masks); larger tiles spill
  constexpr bool MULTI = (M + 1) * WPL <= (FULL ? 64 : 32);
  train_body<WPL, G, M, IB1, QUEUE, FULL, false, MULTI>(L);
}

// The full-wave tile of up to 4 words per lane, include_bits 1, rows without
// tail masks, step-queue mode, at four workgroups (32 waves) per CU.  That
// takes at most 64 VGPRs and few enough SGPRs: a CU of gfx950 holds four
// 512-thread workgroups of 64 VGPRs at 70 SGPRs each but only three at 86, so
// the kernel is held to 76 SGPRs (the compiler keeps the rest in the lanes of
// a VGPR) and its event counters live in LDS.  The 4-word tile then takes 64
// VGPRs by itself; an occupancy bound of 8 waves per SIMD instead makes the
// compiler spill whenever the body needs more.
template <int WPL, bool QUEUE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_num_sgpr(76))) void train_kernel_r4(tm_train_launch L)
{
  train_body<WPL, 64, 7, true, QUEUE, true, true, false>(L);
}

} // namespace
