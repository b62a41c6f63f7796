// tm_rng.h -- the engine's counter-based random stream.
//
// Every random decision of training is a pure function of
// (seed, epoch, sample, chunk, clause, draw index), so the GPU kernels can
// evaluate any of them in any order and the result does not depend on launch
// shape, scheduling or device count.  The same functions compile for host and
// device.
//
//   key                = (seed lo32, seed hi32)
//   epoch keys ek[0..3] = Philox4x32-10(key, ctr = (epoch, 0, 0, TAG_EPOCH))
//   sample order        = positions sorted by (Philox(key, (i, epoch, 0, TAG_PERM))[0], i)
//   skip draw (s, ch)   = lowbias32(lowbias32(s ^ ek0) ^ (ch*GOLDEN + ek1))
//                         (negative sampling, capi.c:215)
//   step draw (s, ch, j) = word j&3 of Philox(key, (s, ch, epoch, TAG_STEP + (j>>2))):
//                         one 32-bit draw per clause j of the chunk; clause j
//                         fires iff mulhi(draw, 2T) < T -+ v, i.e. with
//                         probability (T -+ v) / 2T to within 2^-32
//                         (the chance(p) pair draws, capi.c:221-239)
//   literal key (s, cl) = lowbias32(lowbias32(s ^ ek2) ^ (cl*GOLDEN + ek3))
//   literal draw r      = (lowbias32(kev + r*GOLDEN) * input_bits) >> 32
//                         (the "random % input_bits" of capi.c:188)
//
// Philox4x32-10: Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
// SC'11.  lowbias32: C. Wellons' hash-prospector integer mixer.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define TM_HD __host__ __device__ __forceinline__
#else
#define TM_HD inline
#endif

#define TM_TAG_EPOCH 0x45504f43u
#define TM_TAG_PERM 0x5045524du
#define TM_TAG_STEP 0x53544550u     // clauses 0-3 of a chunk; TM_TAG_STEP + 1: clauses 4-7
#define TM_GOLDEN 0x9E3779B9u

struct tm_u4 { uint32_t x, y, z, w; };

TM_HD uint32_t tm_mulhi(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t) a * b) >> 32);
#endif
}

TM_HD tm_u4 tm_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t hi0 = tm_mulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    uint32_t hi1 = tm_mulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  tm_u4 o = { c0, c1, c2, c3 };
  return o;
}

TM_HD uint32_t tm_lowbias32(uint32_t x)
{
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

TM_HD uint32_t tm_skip_draw(uint32_t ek0, uint32_t ek1, uint32_t sample, uint32_t chunk)
{
  return tm_lowbias32(tm_lowbias32(sample ^ ek0) ^ (chunk * TM_GOLDEN + ek1));
}

TM_HD uint32_t tm_literal_key(uint32_t ek2, uint32_t ek3, uint32_t sample, uint32_t clause)
{
  return tm_lowbias32(tm_lowbias32(sample ^ ek2) ^ (clause * TM_GOLDEN + ek3));
}

TM_HD uint32_t tm_literal_draw(uint32_t kev, uint32_t r, uint32_t input_bits)
{
  return tm_mulhi(tm_lowbias32(kev + r * TM_GOLDEN), input_bits);
}
