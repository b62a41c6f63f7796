// tm_mlp.hip -- the search surrogate: an FP32 MLP 12 -> 64 -> 64 -> 64 -> 1
// (ReLU hidden layers, linear output) with hand-written forward, backward and
// Adam kernels.  No counterpart in the reference: optimize.lua's search
// proposes trials from its samplers only (optimize.lua:187-516).
//
// Parameters are one flat float array (9217 values):
//   W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
// The network is tiny, so one workgroup trains it: weights and gradients sit in
// LDS (2 x 36.9 KB) and per-sample activations and deltas in a global scratch
// area that stays in L2.
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

struct mlp_train_args {
  float *params, *m, *v;         // weights and Adam moments (global)
  const float *X, *Y;            // n samples
  float *act;                    // scratch: A1, A2, A3, D1, D2, D3 [n][64] and dout [n]
  float *grad_out;               // mode 1: the loss gradient is written here, nothing updated
  float *loss_out;               // loss of the last step (mean squared error)
  uint32_t n, steps, step0, mode;
  float lr, beta1, beta2, eps;
};

// `steps` Adam steps of full-batch mean-squared-error training in one
// workgroup (mode 0), or one gradient evaluation (mode 1).  Every layer is a
// small matrix product: one thread per (sample, unit) output for the forward
// and delta passes, one thread per parameter for the gradient sums, with a
// workgroup barrier between passes.
__device__ __forceinline__ void dense_relu(const float *__restrict__ w, int ow, int ob, int nin, const float *in,
                                           float *out, uint32_t n, int tid)
{
  for (uint32_t idx = tid; idx < n * H; idx += MLP_THREADS) {
    const uint32_t s = idx / H, i = idx % H;
    float a = w[ob + i];
    const float *x = in + (size_t) s * nin;
    for (int k = 0; k < nin; k++) a = fmaf(w[ow + i * nin + k], x[k], a);
    out[idx] = fmaxf(a, 0.f);
  }
}

// delta of layer l from the delta of layer l+1: D[s][i] = (A[s][i] > 0) * sum_k W[k][i] Dn[s][k]
__device__ __forceinline__ void dense_back(const float *__restrict__ w, int ow, const float *dn, const float *a,
                                           float *d, uint32_t n, int tid)
{
  for (uint32_t idx = tid; idx < n * H; idx += MLP_THREADS) {
    const uint32_t s = idx / H, i = idx % H;
    float acc = 0.f;
    const float *dr = dn + (size_t) s * H;
    for (int k = 0; k < H; k++) acc = fmaf(w[ow + k * H + i], dr[k], acc);
    d[idx] = a[idx] > 0.f ? acc : 0.f;
  }
}

__global__ __launch_bounds__(MLP_THREADS) void mlp_train_kernel(mlp_train_args A)
{
  __shared__ float w[NP];
  __shared__ float g[NP];
  __shared__ float red[MLP_THREADS];
  const int tid = threadIdx.x;
  const uint32_t n = A.n;
  float *A1 = A.act, *A2 = A1 + (size_t) n * H, *A3 = A2 + (size_t) n * H;
  float *D1 = A3 + (size_t) n * H, *D2 = D1 + (size_t) n * H, *D3 = D2 + (size_t) n * H, *DO = D3 + (size_t) n * H;
  for (int i = tid; i < NP; i += MLP_THREADS) w[i] = A.params[i];
  __syncthreads();
  const uint32_t nsteps = A.mode == 1 ? 1 : A.steps;
  for (uint32_t step = 0; step < nsteps; step++) {
    dense_relu(w, OW1, OB1, IN, A.X, A1, n, tid);
    __syncthreads();
    dense_relu(w, OW2, OB2, H, A1, A2, n, tid);
    __syncthreads();
    dense_relu(w, OW3, OB3, H, A2, A3, n, tid);
    __syncthreads();
    float lsum = 0.f;
    for (uint32_t s = tid; s < n; s += MLP_THREADS) {
      float o = w[OB4];
      for (int k = 0; k < H; k++) o = fmaf(w[OW4 + k], A3[(size_t) s * H + k], o);
      const float err = o - A.Y[s];
      lsum += err * err;
      DO[s] = 2.f * err / (float) n;
    }
    red[tid] = lsum;
    __syncthreads();
    for (uint32_t idx = tid; idx < n * H; idx += MLP_THREADS) {
      const uint32_t s = idx / H, i = idx % H;
      D3[idx] = A3[idx] > 0.f ? DO[s] * w[OW4 + i] : 0.f;
    }
    __syncthreads();
    dense_back(w, OW3, D3, A2, D2, n, tid);
    __syncthreads();
    dense_back(w, OW2, D2, A1, D1, n, tid);
    __syncthreads();
    // gradients: each thread owns a strided set of parameters and sums over samples
    for (int p = tid; p < NP; p += MLP_THREADS) {
      float acc = 0.f;
      if (p < OB1) {
        const int i = (p - OW1) / IN, k = (p - OW1) % IN;
        for (uint32_t s = 0; s < n; s++) acc = fmaf(D1[(size_t) s * H + i], A.X[(size_t) s * IN + k], acc);
      } else if (p < OW2) {
        const int i = p - OB1;
        for (uint32_t s = 0; s < n; s++) acc += D1[(size_t) s * H + i];
      } else if (p < OB2) {
        const int i = (p - OW2) / H, k = (p - OW2) % H;
        for (uint32_t s = 0; s < n; s++) acc = fmaf(D2[(size_t) s * H + i], A1[(size_t) s * H + k], acc);
      } else if (p < OW3) {
        const int i = p - OB2;
        for (uint32_t s = 0; s < n; s++) acc += D2[(size_t) s * H + i];
      } else if (p < OB3) {
        const int i = (p - OW3) / H, k = (p - OW3) % H;
        for (uint32_t s = 0; s < n; s++) acc = fmaf(D3[(size_t) s * H + i], A2[(size_t) s * H + k], acc);
      } else if (p < OW4) {
        const int i = p - OB3;
        for (uint32_t s = 0; s < n; s++) acc += D3[(size_t) s * H + i];
      } else if (p < OB4) {
        const int k = p - OW4;
        for (uint32_t s = 0; s < n; s++) acc = fmaf(DO[s], A3[(size_t) s * H + k], acc);
      } else {
        for (uint32_t s = 0; s < n; s++) acc += DO[s];
      }
      g[p] = acc;
    }
    __syncthreads();
    if (tid == 0 && A.loss_out) {
      float l = 0.f;
      for (int t = 0; t < MLP_THREADS; t++) l += red[t];
      *A.loss_out = l / (float) n;
    }
    if (A.mode == 1) {
      for (int p = tid; p < NP; p += MLP_THREADS) A.grad_out[p] = g[p];
      return;
    }
    // Adam (Kingma & Ba, 2015) with bias correction
    const float t = (float)(A.step0 + step + 1);
    const float c1 = 1.f - powf(A.beta1, t), c2 = 1.f - powf(A.beta2, t);
    for (int p = tid; p < NP; p += MLP_THREADS) {
      const float gp = g[p];
      const float mp = A.beta1 * A.m[p] + (1.f - A.beta1) * gp;
      const float vp = A.beta2 * A.v[p] + (1.f - A.beta2) * gp * gp;
      A.m[p] = mp;
      A.v[p] = vp;
      w[p] -= A.lr * (mp / c1) / (sqrtf(vp / c2) + A.eps);
    }
    __syncthreads();
  }
  for (int p = tid; p < NP; p += MLP_THREADS) A.params[p] = w[p];
}

} // namespace

struct tm_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *params = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr;
  float *x = nullptr, *y = nullptr, *act = nullptr, *out = nullptr;
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
  hipFree(m->x); hipFree(m->y); hipFree(m->act); hipFree(m->out);
  m->x = m->y = m->act = m->out = nullptr;
  m->cap = 0;
  MLP_HIP(hipMalloc(&m->x, n * IN * sizeof(float)));
  MLP_HIP(hipMalloc(&m->y, n * sizeof(float)));
  MLP_HIP(hipMalloc(&m->act, (n * 6 * H + n) * sizeof(float)));
  MLP_HIP(hipMalloc(&m->out, n * sizeof(float)));
  m->cap = n;
  return TM_OK;
}

int run_train(tm_mlp *m, const float *X, const float *Y, uint32_t n, uint32_t steps, float lr, float *grad,
              float *loss)
{
  if (!m || !X || !Y || n == 0) return mlp_fail(TM_E_ARG, "mlp: bad arguments");
  MLP_HIP(hipSetDevice(m->device));
  int rc = ensure(m, n);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(m->x, X, (size_t) n * IN * 4, hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipMemcpyAsync(m->y, Y, (size_t) n * 4, hipMemcpyHostToDevice, m->stream));
  float *gdev = nullptr;
  if (grad) MLP_HIP(hipMalloc(&gdev, NP * sizeof(float)));
  mlp_train_args A;
  A.params = m->params; A.m = m->m; A.v = m->v;
  A.X = m->x; A.Y = m->y; A.act = m->act; A.grad_out = gdev; A.loss_out = m->loss;
  A.n = n; A.steps = steps; A.step0 = m->step; A.mode = grad ? 1 : 0;
  A.lr = lr; A.beta1 = 0.9f; A.beta2 = 0.999f; A.eps = 1e-8f;
  hipLaunchKernelGGL(mlp_train_kernel, dim3(1), dim3(MLP_THREADS), 0, m->stream, A);
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

int tm_mlp_train(tm_mlp *m, const float *X, const float *Y, uint32_t n, uint32_t steps, float lr, float *loss)
{
  return run_train(m, X, Y, n, steps, lr, nullptr, loss);
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
  hipFree(m->x); hipFree(m->y); hipFree(m->act); hipFree(m->out);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

} // extern "C"
