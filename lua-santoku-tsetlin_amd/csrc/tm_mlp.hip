// tm_mlp.hip -- the search surrogate: an FP32 MLP 12 -> 64 -> 64 -> 64 -> 1
// (ReLU hidden layers, linear output) with hand-written forward, backward and
// Adam kernels.  No counterpart in the reference: optimize.lua's search
// proposes trials from its samplers only (optimize.lua:187-516).
//
// Parameters are one flat float array (9217 values):
//   W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
// Training splits the samples into slices of 32, one workgroup each, with the
// weights (also transposed), activations and deltas of its slice in LDS; the
// slices' gradient sums meet in a second kernel that applies Adam.  Several
// networks' fits (a search refits an ensemble and its cross-validation folds
// every round) share those two launches, one grid row per network.
#include "tsetlin_amd.h"
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <vector>


namespace {

constexpr int IN = TM_MLP_INPUTS, H = TM_MLP_HIDDEN;
constexpr int OW1 = 0, OB1 = OW1 + H * IN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * H,
              OW4 = OB3 + H, OB4 = OW4 + H, NP = OB4 + 1;
static_assert(NP == TM_MLP_PARAMS, "parameter layout");
constexpr int MLP_THREADS = 256;

thread_local char g_mlp_err[256] = "";

int mlp_fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_mlp_err, sizeof g_mlp_err, fmt, ap);
  va_end(ap);
  return code;
}

// forward of one sample; h1/h2/h3 are the post-ReLU activations
__device__ __forceinline__ float forward_one(const float *__restrict__ w, const float *x, float (&h1)[H], float (&h2)[H],
                                             float (&h3)[H])
{
#pragma unroll
  for (int i = 0; i < H; i++) {
    float a = w[OB1 + i];
#pragma unroll
    for (int k = 0; k < IN; k++) a = fmaf(w[OW1 + i * IN + k], x[k], a);
    h1[i] = fmaxf(a, 0.f);
  }
#pragma unroll
  for (int i = 0; i < H; i++) {
    float a = w[OB2 + i];
#pragma unroll
    for (int k = 0; k < H; k++) a = fmaf(w[OW2 + i * H + k], h1[k], a);
    h2[i] = fmaxf(a, 0.f);
  }
#pragma unroll
  for (int i = 0; i < H; i++) {
    float a = w[OB3 + i];
#pragma unroll
    for (int k = 0; k < H; k++) a = fmaf(w[OW3 + i * H + k], h2[k], a);
    h3[i] = fmaxf(a, 0.f);
  }
  float o = w[OB4];
#pragma unroll
  for (int k = 0; k < H; k++) o = fmaf(w[OW4 + k], h3[k], o);
  return o;
}

__global__ __launch_bounds__(MLP_THREADS) void mlp_predict_kernel(const float *__restrict__ params,
                                                                  const float *__restrict__ X, float *__restrict__ Y,
                                                                  uint32_t n)
{
  __shared__ float w[NP];
  for (int i = threadIdx.x; i < NP; i += MLP_THREADS) w[i] = params[i];
  __syncthreads();
  const uint32_t s = blockIdx.x * MLP_THREADS + threadIdx.x;
  if (s >= n) return;
  float x[IN], h1[H], h2[H], h3[H];
#pragma unroll
  for (int k = 0; k < IN; k++) x[k] = X[(size_t) s * IN + k];
  Y[s] = forward_one(w, x, h1, h2, h3);
}

// Training is full-batch mean-squared error with Adam, one step at a time:
//   mlp_grad_kernel: one workgroup per slice of MLP_SLICE samples runs the
//     forward and backward passes of its slice in LDS and writes the slice's
//     gradient sums and squared-error sum;
//   mlp_adam_kernel: one thread per parameter adds the slices' sums and
//     applies the Adam update (or, for a gradient evaluation, stores the sum).
// A step is the two launches, so every workgroup of the gradient pass reads
// the weights the previous step wrote.
constexpr int MLP_SLICE = 32;
constexpr int HP = H + 1;     // activation row stride: rows of consecutive samples fall in different LDS banks

struct mlp_grad_args {
  const float *params;
  const float *X, *Y;
  float *part;                 // [slices][NP] gradient sums
  float *lpart;                // [slices] squared-error sums
  uint32_t n;
};

// out[s][i] = relu(b[i] + sum_k W[i][k] in[s][k]) with the weights transposed
// (wt[k][i], rows of HP floats so that the transposing writes spread over the
// banks): the lanes of a wave take consecutive units i of one sample, so the
// weight reads are consecutive and the input read is a broadcast
__device__ __forceinline__ void fwd_layer(const float *wt, const float *b, int nin, const float *in, int in_stride,
                                          float *out, int tid)
{
  for (int idx = tid; idx < MLP_SLICE * H; idx += MLP_THREADS) {
    const int s = idx / H, i = idx % H;
    float a = b[i];
    for (int k = 0; k < nin; k++) a = fmaf(wt[k * HP + i], in[s * in_stride + k], a);
    out[s * HP + i] = fmaxf(a, 0.f);
  }
}

// d[s][i] = (a[s][i] > 0) * sum_k W[k][i] dn[s][k], W in its own layout [k][i]
__device__ __forceinline__ void back_layer(const float *w, const float *dn, const float *a, float *d, int tid)
{
  for (int idx = tid; idx < MLP_SLICE * H; idx += MLP_THREADS) {
    const int s = idx / H, i = idx % H;
    float acc = 0.f;
    for (int k = 0; k < H; k++) acc = fmaf(w[k * H + i], dn[s * HP + k], acc);
    d[s * HP + i] = a[s * HP + i] > 0.f ? acc : 0.f;
  }
}

// the gradient pass of one slice of one network
__device__ __forceinline__ void grad_slice(const mlp_grad_args &A, const uint32_t slice)
{
  __shared__ float w[NP];
  __shared__ float wt[IN * HP + 2 * H * HP];        // W1, W2, W3 transposed, rows of HP floats
  __shared__ float xs[MLP_SLICE * IN];
  __shared__ float act[3][MLP_SLICE * HP];          // post-ReLU activations of the three hidden layers
  __shared__ float del[3][MLP_SLICE * HP];          // their deltas
  __shared__ float dout[MLP_SLICE];
  const int tid = threadIdx.x;
  const uint32_t s0 = slice * MLP_SLICE;
  const uint32_t ns = min((uint32_t) MLP_SLICE, A.n - s0);
  // the parameters as 16-byte loads, all of a thread's in flight together
  // (one at a time, each waiting for the last, they took most of a step)
  auto put = [&](int p, float v) {
    w[p] = v;
    if (p < OB1) wt[(p % IN) * HP + p / IN] = v;
    else if (p >= OW2 && p < OB2) wt[IN * HP + ((p - OW2) % H) * HP + (p - OW2) / H] = v;
    else if (p >= OW3 && p < OB3) wt[IN * HP + H * HP + ((p - OW3) % H) * HP + (p - OW3) / H] = v;
  };
  constexpr int NQ = NP / 4, QPT = (NQ + MLP_THREADS - 1) / MLP_THREADS;
  const float4 *pq = reinterpret_cast<const float4 *>(A.params);
  float4 q[QPT];
#pragma unroll
  for (int u = 0; u < QPT; u++) {
    const int qi = tid + u * MLP_THREADS;
    q[u] = qi < NQ ? pq[qi] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int u = 0; u < QPT; u++) {
    const int qi = tid + u * MLP_THREADS;
    if (qi < NQ) {
      put(4 * qi, q[u].x);
      put(4 * qi + 1, q[u].y);
      put(4 * qi + 2, q[u].z);
      put(4 * qi + 3, q[u].w);
    }
  }
  for (int p = 4 * NQ + tid; p < NP; p += MLP_THREADS) put(p, A.params[p]);
  // samples past the slice's end enter as zeros and leave with zero deltas
  for (int idx = tid; idx < MLP_SLICE * IN; idx += MLP_THREADS)
    xs[idx] = (uint32_t)(idx / IN) < ns ? A.X[(size_t) s0 * IN + idx] : 0.f;
  __syncthreads();
  fwd_layer(wt, w + OB1, IN, xs, IN, act[0], tid);
  __syncthreads();
  fwd_layer(wt + IN * HP, w + OB2, H, act[0], HP, act[1], tid);
  __syncthreads();
  fwd_layer(wt + IN * HP + H * HP, w + OB3, H, act[1], HP, act[2], tid);
  __syncthreads();
  float l = 0.f;
  if (tid < MLP_SLICE) {     // wave 0: one sample per lane
    float o = w[OB4];
    for (int k = 0; k < H; k++) o = fmaf(w[OW4 + k], act[2][tid * HP + k], o);
    const float err = (uint32_t) tid < ns ? o - A.Y[s0 + tid] : 0.f;
    l = err * err;
    dout[tid] = 2.f * err / (float) A.n;
  }
  if (tid < 64) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if (tid == 0) A.lpart[slice] = l;
  }
  __syncthreads();
  for (int idx = tid; idx < MLP_SLICE * H; idx += MLP_THREADS) {
    const int s = idx / H, i = idx % H;
    del[2][s * HP + i] = act[2][s * HP + i] > 0.f ? dout[s] * w[OW4 + i] : 0.f;
  }
  __syncthreads();
  back_layer(w + OW3, del[2], act[1], del[1], tid);
  __syncthreads();
  back_layer(w + OW2, del[1], act[0], del[0], tid);
  __syncthreads();
  // the slice's gradient sums, one parameter per thread at a time
  float *part = A.part + (size_t) slice * NP;
  for (int p = tid; p < NP; p += MLP_THREADS) {
    float acc = 0.f;
    if (p < OB1) {
      const int i = p / IN, k = p % IN;
      for (int s = 0; s < MLP_SLICE; s++) acc = fmaf(del[0][s * HP + i], xs[s * IN + k], acc);
    } else if (p < OW2) {
      for (int s = 0; s < MLP_SLICE; s++) acc += del[0][s * HP + p - OB1];
    } else if (p < OB2) {
      const int i = (p - OW2) / H, k = (p - OW2) % H;
      for (int s = 0; s < MLP_SLICE; s++) acc = fmaf(del[1][s * HP + i], act[0][s * HP + k], acc);
    } else if (p < OW3) {
      for (int s = 0; s < MLP_SLICE; s++) acc += del[1][s * HP + p - OB2];
    } else if (p < OB3) {
      const int i = (p - OW3) / H, k = (p - OW3) % H;
      for (int s = 0; s < MLP_SLICE; s++) acc = fmaf(del[2][s * HP + i], act[1][s * HP + k], acc);
    } else if (p < OW4) {
      for (int s = 0; s < MLP_SLICE; s++) acc += del[2][s * HP + p - OB3];
    } else if (p < OB4) {
      for (int s = 0; s < MLP_SLICE; s++) acc = fmaf(dout[s], act[2][s * HP + p - OW4], acc);
    } else {
      for (int s = 0; s < MLP_SLICE; s++) acc += dout[s];
    }
    part[p] = acc;
  }
}

__global__ __launch_bounds__(MLP_THREADS) void mlp_grad_kernel(mlp_grad_args A) { grad_slice(A, blockIdx.x); }

// Several networks' fits in one launch: network blockIdx.y, slice blockIdx.x
// (the grid spans the largest network's slices).
constexpr uint32_t MLP_MANY = 16;
struct mlp_grad_many_args {
  mlp_grad_args a[MLP_MANY];
};

__global__ __launch_bounds__(MLP_THREADS) void mlp_grad_many_kernel(mlp_grad_many_args M)
{
  const mlp_grad_args &A = M.a[blockIdx.y];
  if (blockIdx.x * MLP_SLICE >= A.n) return;
  grad_slice(A, blockIdx.x);
}

struct mlp_adam_args {
  float *params, *m, *v;       // weights and Adam moments
  const float *part, *lpart;   // the gradient pass's slice sums
  float *grad_out;             // gradient evaluation: the summed gradient goes here, nothing is updated
  float *loss_out;             // mean squared error of the step's forward pass
  uint32_t slices, n, t;       // t: Adam step number, from 1
  float lr, beta1, beta2, eps;
};

__device__ __forceinline__ void adam_param(const mlp_adam_args &A, const uint32_t p)
{
  if (p == 0 && A.loss_out) {
    float l = 0.f;
    for (uint32_t b = 0; b < A.slices; b++) l += A.lpart[b];
    *A.loss_out = l / (float) A.n;
  }
  if (p >= (uint32_t) NP) return;
  float g = 0.f;
  for (uint32_t b = 0; b < A.slices; b++) g += A.part[(size_t) b * NP + p];
  if (A.grad_out) {
    A.grad_out[p] = g;
    return;
  }
  // Adam (Kingma & Ba, 2015) with bias correction
  const float t = (float) A.t;
  const float c1 = 1.f - powf(A.beta1, t), c2 = 1.f - powf(A.beta2, t);
  const float mp = A.beta1 * A.m[p] + (1.f - A.beta1) * g;
  const float vp = A.beta2 * A.v[p] + (1.f - A.beta2) * g * g;
  A.m[p] = mp;
  A.v[p] = vp;
  A.params[p] -= A.lr * (mp / c1) / (sqrtf(vp / c2) + A.eps);
}

__global__ __launch_bounds__(MLP_THREADS) void mlp_adam_kernel(mlp_adam_args A)
{
  adam_param(A, blockIdx.x * MLP_THREADS + threadIdx.x);
}

struct mlp_adam_many_args {
  mlp_adam_args a[MLP_MANY];
};

__global__ __launch_bounds__(MLP_THREADS) void mlp_adam_many_kernel(mlp_adam_many_args M)
{
  adam_param(M.a[blockIdx.y], blockIdx.x * MLP_THREADS + threadIdx.x);
}

} // namespace

struct tm_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *params = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr;
  float *x = nullptr, *y = nullptr, *part = nullptr, *lpart = nullptr, *out = nullptr;
  size_t cap = 0;
  uint32_t step = 0;
};

namespace {

#define MLP_HIP(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return mlp_fail(TM_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

int ensure(tm_mlp *m, size_t n)
{
  if (n <= m->cap) return TM_OK;
  hipFree(m->x); hipFree(m->y); hipFree(m->part); hipFree(m->lpart); hipFree(m->out);
  m->x = m->y = m->part = m->lpart = m->out = nullptr;
  m->cap = 0;
  const size_t slices = (n + MLP_SLICE - 1) / MLP_SLICE;
  MLP_HIP(hipMalloc(&m->x, n * IN * sizeof(float)));
  MLP_HIP(hipMalloc(&m->y, n * sizeof(float)));
  MLP_HIP(hipMalloc(&m->part, slices * NP * sizeof(float)));
  MLP_HIP(hipMalloc(&m->lpart, slices * sizeof(float)));
  MLP_HIP(hipMalloc(&m->out, n * sizeof(float)));
  m->cap = n;
  return TM_OK;
}

// The samples onto the network's buffers and the kernel arguments of its
// steps; the launches go on the network's own stream.
int stage(tm_mlp *m, const float *X, const float *Y, uint32_t n, float lr, float *gdev, mlp_grad_args &G,
          mlp_adam_args &A, hipStream_t stream)
{
  MLP_HIP(hipSetDevice(m->device));
  int rc = ensure(m, n);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(m->x, X, (size_t) n * IN * 4, hipMemcpyHostToDevice, stream));
  MLP_HIP(hipMemcpyAsync(m->y, Y, (size_t) n * 4, hipMemcpyHostToDevice, stream));
  const uint32_t slices = (n + MLP_SLICE - 1) / MLP_SLICE;
  G.params = m->params; G.X = m->x; G.Y = m->y; G.part = m->part; G.lpart = m->lpart; G.n = n;
  A.params = m->params; A.m = m->m; A.v = m->v; A.part = m->part; A.lpart = m->lpart;
  A.grad_out = gdev; A.loss_out = m->loss; A.slices = slices; A.n = n;
  A.lr = lr; A.beta1 = 0.9f; A.beta2 = 0.999f; A.eps = 1e-8f;
  return TM_OK;
}

void launch_step(tm_mlp *m, const mlp_grad_args &G, mlp_adam_args &A, uint32_t step)
{
  A.t = m->step + step + 1;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(A.slices), dim3(MLP_THREADS), 0, m->stream, G);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((NP + MLP_THREADS - 1) / MLP_THREADS), dim3(MLP_THREADS), 0, m->stream, A);
}

int run_train(tm_mlp *m, const float *X, const float *Y, uint32_t n, uint32_t steps, float lr, float *grad,
              float *loss)
{
  if (!m || !X || !Y || n == 0) return mlp_fail(TM_E_ARG, "mlp: bad arguments");
  float *gdev = nullptr;
  if (grad) {
    MLP_HIP(hipSetDevice(m->device));
    MLP_HIP(hipMalloc(&gdev, NP * sizeof(float)));
  }
  mlp_grad_args G;
  mlp_adam_args A;
  int rc = stage(m, X, Y, n, lr, gdev, G, A, m->stream);
  if (rc) { if (gdev) hipFree(gdev); return rc; }
  const uint32_t nsteps = grad ? 1 : steps;
  for (uint32_t step = 0; step < nsteps; step++) launch_step(m, G, A, step);
  MLP_HIP(hipGetLastError());
  if (grad) MLP_HIP(hipMemcpyAsync(grad, gdev, NP * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (loss) MLP_HIP(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  if (gdev) hipFree(gdev);
  if (!grad) m->step += steps;
  return TM_OK;
}

} // namespace

extern "C" {

int tm_mlp_create(const float *params, tm_mlp **out)
{
  if (!out || !params) return mlp_fail(TM_E_ARG, "mlp create: NULL argument");
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) return mlp_fail(TM_E_DEVICE, "no HIP device available");
  tm_mlp *m = new tm_mlp();
  hipError_t e = hipGetDevice(&m->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&m->params, NP * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->m, NP * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->v, NP * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->loss, sizeof(float));
  if (e != hipSuccess) { tm_mlp_destroy(m); return mlp_fail(TM_E_DEVICE, "mlp create: %s", hipGetErrorString(e)); }
  std::vector<float> zero(NP, 0.f);
  if (hipMemcpy(m->params, params, NP * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->m, zero.data(), NP * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->v, zero.data(), NP * 4, hipMemcpyHostToDevice) != hipSuccess) {
    tm_mlp_destroy(m);
    return mlp_fail(TM_E_DEVICE, "mlp create: upload failed");
  }
  *out = m;
  return TM_OK;
}

int tm_mlp_reset(tm_mlp *m, const float *params)
{
  if (!m || !params) return mlp_fail(TM_E_ARG, "mlp reset: NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipMemcpyAsync(m->params, params, NP * 4, hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipMemsetAsync(m->m, 0, NP * 4, m->stream));
  MLP_HIP(hipMemsetAsync(m->v, 0, NP * 4, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  m->step = 0;
  return TM_OK;
}

int tm_mlp_train(tm_mlp *m, const float *X, const float *Y, uint32_t n, uint32_t steps, float lr, float *loss)
{
  return run_train(m, X, Y, n, steps, lr, nullptr, loss);
}

int tm_mlp_train_many(tm_mlp *const *nets, uint32_t k, const float *const *X, const float *const *Y,
                      const uint32_t *n, uint32_t steps, float lr, float *loss)
{
  if (!nets || !X || !Y || !n) return mlp_fail(TM_E_ARG, "mlp train_many: NULL argument");
  for (uint32_t i = 0; i < k; i++) {
    if (!nets[i] || !X[i] || !Y[i] || n[i] == 0) return mlp_fail(TM_E_ARG, "mlp train_many: bad arguments for network %u", i);
    // two fits of one network would share its buffers and its stream
    for (uint32_t j = 0; j < i; j++)
      if (nets[j] == nets[i]) return mlp_fail(TM_E_ARG, "mlp train_many: network %u repeats network %u", i, j);
  }
  bool one_device = true;
  for (uint32_t i = 1; i < k; i++) one_device = one_device && nets[i]->device == nets[0]->device;
  if (k > 0 && k <= MLP_MANY && one_device) {
    // one launch per pass for all of them, on the first network's stream
    tm_mlp *m0 = nets[0];
    const hipStream_t st = m0->stream;
    mlp_grad_many_args GM = {};
    mlp_adam_many_args AM = {};
    uint32_t slices = 0;
    for (uint32_t i = 0; i < k; i++) {
      int rc = stage(nets[i], X[i], Y[i], n[i], lr, nullptr, GM.a[i], AM.a[i], st);
      if (rc) return rc;
      slices = max(slices, AM.a[i].slices);
    }
    MLP_HIP(hipSetDevice(m0->device));
    for (uint32_t step = 0; step < steps; step++) {
      for (uint32_t i = 0; i < k; i++) AM.a[i].t = nets[i]->step + step + 1;
      hipLaunchKernelGGL(mlp_grad_many_kernel, dim3(slices, k), dim3(MLP_THREADS), 0, st, GM);
      hipLaunchKernelGGL(mlp_adam_many_kernel, dim3((NP + MLP_THREADS - 1) / MLP_THREADS, k), dim3(MLP_THREADS), 0,
                         st, AM);
    }
    MLP_HIP(hipGetLastError());
    if (loss)
      for (uint32_t i = 0; i < k; i++)
        MLP_HIP(hipMemcpyAsync(&loss[i], nets[i]->loss, sizeof(float), hipMemcpyDeviceToHost, st));
    MLP_HIP(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < k; i++) nets[i]->step += steps;
    return TM_OK;
  }
  std::vector<mlp_grad_args> G(k);
  std::vector<mlp_adam_args> A(k);
  for (uint32_t i = 0; i < k; i++) {
    int rc = stage(nets[i], X[i], Y[i], n[i], lr, nullptr, G[i], A[i], nets[i]->stream);
    if (rc) return rc;
  }
  // more networks than one launch takes, or on several devices: step by step
  // across the networks on their own streams, so every stream has work
  // queued while the host is still launching
  int cur = -1;
  for (uint32_t step = 0; step < steps; step++)
    for (uint32_t i = 0; i < k; i++) {
      if (nets[i]->device != cur) {
        cur = nets[i]->device;
        MLP_HIP(hipSetDevice(cur));
      }
      launch_step(nets[i], G[i], A[i], step);
    }
  MLP_HIP(hipGetLastError());
  for (uint32_t i = 0; i < k; i++) {
    MLP_HIP(hipSetDevice(nets[i]->device));
    if (loss) MLP_HIP(hipMemcpyAsync(&loss[i], nets[i]->loss, sizeof(float), hipMemcpyDeviceToHost, nets[i]->stream));
  }
  for (uint32_t i = 0; i < k; i++) {
    MLP_HIP(hipStreamSynchronize(nets[i]->stream));
    nets[i]->step += steps;
  }
  return TM_OK;
}

int tm_mlp_gradient(tm_mlp *m, const float *X, const float *Y, uint32_t n, float *grad, float *loss)
{
  if (!grad) return mlp_fail(TM_E_ARG, "mlp gradient: grad is NULL");
  return run_train(m, X, Y, n, 0, 0.f, grad, loss);
}

int tm_mlp_predict(tm_mlp *m, const float *X, uint32_t n, float *Y)
{
  if (!m || !X || !Y) return mlp_fail(TM_E_ARG, "mlp predict: NULL argument");
  if (n == 0) return TM_OK;
  MLP_HIP(hipSetDevice(m->device));
  int rc = ensure(m, n);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(m->x, X, (size_t) n * IN * 4, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(mlp_predict_kernel, dim3((n + MLP_THREADS - 1) / MLP_THREADS), dim3(MLP_THREADS), 0, m->stream,
                     m->params, m->x, m->out, n);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipMemcpyAsync(Y, m->out, (size_t) n * 4, hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return TM_OK;
}

int tm_mlp_params(tm_mlp *m, float *out)
{
  if (!m || !out) return mlp_fail(TM_E_ARG, "mlp params: NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipMemcpyAsync(out, m->params, NP * 4, hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return TM_OK;
}

const char *tm_mlp_last_error(void) { return g_mlp_err; }

void tm_mlp_destroy(tm_mlp *m)
{
  if (!m) return;
  if (m->stream) hipStreamSynchronize(m->stream);
  hipFree(m->params); hipFree(m->m); hipFree(m->v); hipFree(m->loss);
  hipFree(m->x); hipFree(m->y); hipFree(m->part); hipFree(m->lpart); hipFree(m->out);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

} // extern "C"
