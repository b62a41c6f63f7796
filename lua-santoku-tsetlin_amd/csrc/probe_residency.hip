// Workgroup residency of the GPU: how many 512-thread workgroups (the train
// kernel's shape: 8 waves, one per clause of a chunk) one CU holds at once,
// for a given VGPR, SGPR and LDS footprint.  Every workgroup sleeps a fixed
// time, so a grid of k workgroups per CU takes one sleep when k fit at once
// and two when they do not.  One output word per workgroup (out has grid
// words).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>

// Footprint: 64 VGPRs, LDS words of LDS, and SGPRs up to s<S-1> (S 0: only
// what the kernel needs).
template <int S, int LDSW>
__global__ __launch_bounds__(512) void sleeper(uint32_t *out, int iters)
{
  __shared__ uint32_t lds[LDSW];
  asm volatile("" ::: "v63");
  if constexpr (S == 64) asm volatile("" ::: "s63");
  if constexpr (S == 80) asm volatile("" ::: "s79");
  if constexpr (S == 88) asm volatile("" ::: "s87");
  if constexpr (S == 96) asm volatile("" ::: "s95");
  if constexpr (S == 102) asm volatile("" ::: "s101");
  if (threadIdx.x == 0) lds[0] = blockIdx.x;
  __syncthreads();
  for (int i = 0; i < iters; i++) __builtin_amdgcn_s_sleep(127);
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = lds[0];
}

template <typename K>
static float time_ms(K kernel, uint32_t *out, int grid, int iters)
{
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1.f;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(512), 0, 0, out, iters);   // warm
  hipEventRecord(a, 0);
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(512), 0, 0, out, iters);
  hipEventRecord(b, 0);
  hipEventSynchronize(b);
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  return ms;
}

int main()
{
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus <= 0) return 1;
  const int max_per_cu = 5, iters = 200;
  uint32_t *out = nullptr;
  if (hipMalloc(&out, (size_t) max_per_cu * cus * 4) != hipSuccess) return 1;
  printf("CUs %d; ms per launch of k workgroups per CU (512 threads, 64 VGPRs, sleeping)\n", cus);
  printf("k | 8.7KB LDS, %-14s | 1KB LDS | SGPRs to s63 | s79 | s87 | s95 | s101\n", "kernel's SGPRs");
  for (int k = 1; k <= max_per_cu; k++) {
    const int grid = k * cus;
    printf("%d | %.3f | %.3f | %.3f | %.3f | %.3f | %.3f | %.3f\n", k, time_ms(sleeper<0, 2176>, out, grid, iters),
           time_ms(sleeper<0, 256>, out, grid, iters), time_ms(sleeper<64, 2176>, out, grid, iters),
           time_ms(sleeper<80, 2176>, out, grid, iters), time_ms(sleeper<88, 2176>, out, grid, iters),
           time_ms(sleeper<96, 2176>, out, grid, iters), time_ms(sleeper<102, 2176>, out, grid, iters));
  }
  const bool ok = hipGetLastError() == hipSuccess;
  hipFree(out);
  return ok ? 0 : 2;
}
