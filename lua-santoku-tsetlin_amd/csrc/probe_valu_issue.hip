// Integer VALU issue rate of the GPU: wave64 instructions per second for the
// full-rate logic ops the train kernel is made of (xor, three-input logic,
// add) and for popcount.  16 independent chains per lane, one output
// word per thread (out has grid*block words).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <vector>

template <int OP>
__global__ __launch_bounds__(256) void probe(uint32_t *out, uint32_t seed, int iters)
{
  uint32_t a[16];
#pragma unroll
  for (int i = 0; i < 16; i++) a[i] = seed * (threadIdx.x + 1 + i) ^ (blockIdx.x << i);
  const uint32_t b = seed ^ threadIdx.x, c = ~seed + blockIdx.x;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if (OP == 0) a[i] = a[i] ^ (a[(i + 1) & 15]);
      if (OP == 1) a[i] = __builtin_amdgcn_bitop3_b32(a[i], a[(i + 1) & 15], b, 0x96);
      if (OP == 2) a[i] = __builtin_popcount(a[i]) + a[(i + 1) & 15];
      if (OP == 3) a[i] = a[i] + a[(i + 1) & 15] + c;
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) r ^= a[i];
  out[(size_t) blockIdx.x * 256 + threadIdx.x] = r;
}

template <int OP>
void run(const char *name, uint32_t *out, int blocks, int iters, int per_iter)
{
  hipEvent_t e0, e1;
  hipEventCreate(&e0); hipEventCreate(&e1);
  probe<OP><<<blocks, 256>>>(out, 7, iters);      // warm
  hipEventRecord(e0);
  const int reps = 5;
  for (int r = 0; r < reps; r++) probe<OP><<<blocks, 256>>>(out, 7 + r, iters);
  hipEventRecord(e1);
  hipEventSynchronize(e1);
  float ms; hipEventElapsedTime(&ms, e0, e1);
  double winstr = (double) reps * blocks * 4 * iters * per_iter;   // 4 waves per block
  printf("%-8s %8.3f ms/launch  %.1f G wave-instr/s  %.2f T lane-ops/s\n", name, ms / reps, winstr / (ms / 1e3) / 1e9,
         winstr * 64 / (ms / 1e3) / 1e12);
}

int main()
{
  const int blocks = 65536, iters = 1024;
  uint32_t *out;
  if (hipMalloc(&out, (size_t) blocks * 256 * 4) != hipSuccess) { printf("alloc failed\n"); return 1; }
  hipDeviceProp_t p; hipGetDeviceProperties(&p, 0);
  printf("%s CUs %d clock %d kHz\n", p.gcnArchName, p.multiProcessorCount, p.clockRate);
  run<0>("xor", out, blocks, iters, 16);
  run<1>("bitop3", out, blocks, iters, 16);
  run<2>("bcnt", out, blocks, iters, 16);
  run<3>("add", out, blocks, iters, 24);
  hipError_t e = hipDeviceSynchronize();
  printf("status %s\n", hipGetErrorString(e));
  hipFree(out);
  return e == hipSuccess ? 0 : 1;
}
