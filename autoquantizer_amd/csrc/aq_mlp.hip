// Surrogate MLP of the scheme search: hand-written training and inference
// kernels for gfx950 (fp32; one thread per activation, per delta, per
// parameter gradient).
//
// Network: in -> H (tanh) -> H (tanh) -> H (tanh) -> out, with in = 24,
// H = 64, out = 4 (AQ_MLP_* in include/aq_mpm.h).  Parameters are one flat
// fp32 array: W1 [H][in], b1 [H], W2 [H][H], b2 [H], W3 [H][H], b3 [H],
// W4 [out][H], b4 [out] (row-major, y = W x + b).  Activations and deltas
// are sample-major: act[s][3 H] (h1, h2, h3), delta[s][3 H + out].
// The loss of aq_mlp_backward is the mean squared error over samples and
// outputs, mean((y - t)^2).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "aq_mpm.h"

extern int aq_fail(int code, const char *msg);  // aq_mpm.hip: sets aq_last_error

namespace {

constexpr int IN = AQ_MLP_IN, H = AQ_MLP_HIDDEN, OUT = AQ_MLP_OUT;
constexpr int OW1 = 0, OB1 = OW1 + H * IN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * H,
              OW4 = OB3 + H, OB4 = OW4 + OUT * H, NPARAM = OB4 + OUT;
static_assert(NPARAM == AQ_MLP_PARAMS, "parameter count");

// One dense layer, one thread per (sample, output unit):
// y[s][o] = act(b[o] + sum_i W[o][i] x[s][i]); rows of x and y have strides
// xs, ys (the hidden activations share one [s][3 H] array).
__global__ void __launch_bounds__(256) k_dense_fwd(const float *__restrict__ W, const float *__restrict__ bias,
                                                   const float *__restrict__ x, int32_t xs, int32_t nin,
                                                   float *__restrict__ y, int32_t ys, int32_t nout,
                                                   int32_t nsamples, int32_t use_tanh) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nsamples * nout) return;
  const int s = (int)(t / nout), o = (int)(t % nout);
  float acc = bias[o];
  for (int i = 0; i < nin; ++i) acc = fmaf(W[o * nin + i], x[(int64_t)s * xs + i], acc);
  y[(int64_t)s * ys + o] = use_tanh ? tanhf(acc) : acc;
}

// Output deltas of mean((y - t)^2): 2 (y - t) / (nsamples OUT); the squared
// errors are summed into *loss (device double) when loss is not NULL.
__global__ void __launch_bounds__(256) k_out_delta(const float *__restrict__ Y, const float *__restrict__ T,
                                                   int32_t nsamples, float *__restrict__ delta,
                                                   double *__restrict__ loss) {
  __shared__ double part[4];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double se = 0.0;
  if (t < (int64_t)nsamples * OUT) {
    const int s = (int)(t / OUT), o = (int)(t % OUT);
    const float e = Y[t] - T[t];
    se = (double)e * (double)e;
    delta[(int64_t)s * (3 * H + OUT) + 3 * H + o] = 2.0f * e / ((float)nsamples * (float)OUT);
  }
  if (loss) {
    for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = se;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, ((part[0] + part[1]) + part[2]) + part[3]);
  }
}

// Hidden deltas, one thread per (sample, unit i of the layer below):
// d_in[s][i] = (sum_o W[o][i] d_out[s][o]) (1 - h[s][i]^2).
__global__ void __launch_bounds__(256) k_dense_bwd(const float *__restrict__ W, const float *__restrict__ dout,
                                                   int32_t nout, const float *__restrict__ h, float *__restrict__ din,
                                                   int32_t nin, int32_t nsamples) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nsamples * nin) return;
  const int s = (int)(t / nin), i = (int)(t % nin);
  float g = 0.0f;
  for (int o = 0; o < nout; ++o) g = fmaf(W[o * nin + i], dout[(int64_t)s * (3 * H + OUT) + o], g);
  const float a = h[(int64_t)s * 3 * H + i];
  din[(int64_t)s * (3 * H + OUT) + i] = g * (1.0f - a * a);
}

// Parameter gradients: one thread per parameter, summing over samples in a
// fixed order: dW_l[o][i] = sum_s delta_l[s][o] in_l[s][i] + 2 wd W_l[o][i],
// db_l[o] = sum_s delta_l[s][o].
__global__ void __launch_bounds__(256) k_mlp_grads(const float *__restrict__ prm, const float *__restrict__ X,
                                                   const float *__restrict__ act, const float *__restrict__ delta,
                                                   int32_t nsamples, float wd, float *__restrict__ grad) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NPARAM) return;
  // layer: delta column offset, input array, its row stride and column offset
  int o, i = -1, dofs;
  const float *in = act;
  int64_t stride = 3 * H;
  int col = 0;
  if (q < OB1) { o = (q - OW1) / IN; i = (q - OW1) % IN; dofs = 0; in = X; stride = IN; col = 0; }
  else if (q < OW2) { o = q - OB1; dofs = 0; }
  else if (q < OB2) { o = (q - OW2) / H; i = (q - OW2) % H; dofs = H; col = 0; }
  else if (q < OW3) { o = q - OB2; dofs = H; }
  else if (q < OB3) { o = (q - OW3) / H; i = (q - OW3) % H; dofs = 2 * H; col = H; }
  else if (q < OW4) { o = q - OB3; dofs = 2 * H; }
  else if (q < OB4) { o = (q - OW4) / H; i = (q - OW4) % H; dofs = 3 * H; col = 2 * H; }
  else { o = q - OB4; dofs = 3 * H; }
  float acc = 0.0f;
  for (int s = 0; s < nsamples; ++s) {
    const float d = delta[(int64_t)s * (3 * H + OUT) + dofs + o];
    acc = i < 0 ? acc + d : fmaf(d, in[(int64_t)s * stride + col + i], acc);
  }
  grad[q] = i < 0 ? acc : fmaf(2.0f * wd, prm[q], acc);
}

// Adam (Kingma & Ba 2015) with bias correction; t is the 1-based step.
__global__ void __launch_bounds__(256) k_adam(float *__restrict__ prm, const float *__restrict__ grad,
                                              float *__restrict__ m, float *__restrict__ v, int32_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const float g = grad[q];
  const float mq = b1 * m[q] + (1.0f - b1) * g;
  const float vq = b2 * v[q] + (1.0f - b2) * g * g;
  m[q] = mq;
  v[q] = vq;
  prm[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

}  // namespace

#define MLP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return aq_fail(AQ_ERR_HIP, hipGetErrorString(e_));           \
  } while (0)

static unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

extern "C" int aq_mlp_forward(const float *params, const float *x, int32_t nsamples, float *act, float *y,
                              void *stream) {
  if (nsamples < 0) return aq_fail(AQ_ERR_INVALID, "negative sample count");
  if (nsamples == 0) return AQ_OK;
  if (!params || !x || !y || !act) return aq_fail(AQ_ERR_INVALID, "null buffer");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nh = (int64_t)nsamples * H;
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(nh)), dim3(256), 0, st, params + OW1, params + OB1, x, IN, IN, act, 3 * H,
                     H, nsamples, 1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(nh)), dim3(256), 0, st, params + OW2, params + OB2, act, 3 * H, H,
                     act + H, 3 * H, H, nsamples, 1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(nh)), dim3(256), 0, st, params + OW3, params + OB3, act + H, 3 * H, H,
                     act + 2 * H, 3 * H, H, nsamples, 1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((int64_t)nsamples * OUT)), dim3(256), 0, st, params + OW4, params + OB4,
                     act + 2 * H, 3 * H, H, y, OUT, OUT, nsamples, 0);
  MLP_TRY(hipGetLastError());
  return AQ_OK;
}

extern "C" int aq_mlp_backward(const float *params, const float *x, const float *act, const float *y,
                               const float *target, int32_t nsamples, float weight_decay, float *delta,
                               float *grad, double *loss, void *stream) {
  if (nsamples < 1) return aq_fail(AQ_ERR_INVALID, "no samples");
  if (!(weight_decay >= 0.0f) || !std::isfinite(weight_decay))
    return aq_fail(AQ_ERR_INVALID, "weight decay must be finite and >= 0");
  if (!params || !x || !act || !y || !target || !delta || !grad) return aq_fail(AQ_ERR_INVALID, "null buffer");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nh = (int64_t)nsamples * H;
  hipLaunchKernelGGL(k_out_delta, dim3(nblk((int64_t)nsamples * OUT)), dim3(256), 0, st, y, target, nsamples, delta,
                     loss);
  hipLaunchKernelGGL(k_dense_bwd, dim3(nblk(nh)), dim3(256), 0, st, params + OW4, delta + 3 * H, OUT, act + 2 * H,
                     delta + 2 * H, H, nsamples);
  hipLaunchKernelGGL(k_dense_bwd, dim3(nblk(nh)), dim3(256), 0, st, params + OW3, delta + 2 * H, H, act + H,
                     delta + H, H, nsamples);
  hipLaunchKernelGGL(k_dense_bwd, dim3(nblk(nh)), dim3(256), 0, st, params + OW2, delta + H, H, act, delta, H,
                     nsamples);
  hipLaunchKernelGGL(k_mlp_grads, dim3(nblk(NPARAM)), dim3(256), 0, st, params, x, act, delta, nsamples,
                     weight_decay, grad);
  MLP_TRY(hipGetLastError());
  return AQ_OK;
}

// ---- allocation-free backward -----------------------------------------------
// The deltas of aq_mlp_backward_noalloc live in this file-scope device array
// (one copy per GPU, held until the process ends) instead of a caller
// workspace.  Host code never passes the array's name to a kernel: its
// device address on each GPU comes from hipGetSymbolAddress, taken once per
// device with that device current.
static __device__ __align__(256) float scratch[(int64_t)AQ_MLP_NOALLOC_MAX_SAMPLES * (3 * H + OUT)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device).
int scratch_ptr(float **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return aq_fail(AQ_ERR_HIP, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return aq_fail(AQ_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return aq_fail(AQ_ERR_HIP, (std::string("scratch address: ") + hipGetErrorString(e)).c_str());
    if (!p) return aq_fail(AQ_ERR_HIP, "scratch address is null");
    g_scratch_ptr[dev] = (float *)p;
  }
  *out = g_scratch_ptr[dev];
  return AQ_OK;
}
}  // namespace

extern "C" int aq_mlp_backward_noalloc(const float *params, const float *x, const float *act, const float *y,
                                       const float *target, int32_t nsamples, float weight_decay, float *grad,
                                       double *loss, void *stream) {
  // every check that needs no device comes before the scratch lookup
  if (nsamples < 1) return aq_fail(AQ_ERR_INVALID, "no samples");
  if (nsamples > AQ_MLP_NOALLOC_MAX_SAMPLES)
    return aq_fail(AQ_ERR_CAPACITY, "more samples than the static delta buffer holds (AQ_MLP_NOALLOC_MAX_SAMPLES)");
  if (!(weight_decay >= 0.0f) || !std::isfinite(weight_decay))
    return aq_fail(AQ_ERR_INVALID, "weight decay must be finite and >= 0");
  if (!params || !x || !act || !y || !target || !grad) return aq_fail(AQ_ERR_INVALID, "null buffer");
  float *delta = nullptr;
  const int rc = scratch_ptr(&delta);
  if (rc) return rc;
  // every delta entry the gradient kernel reads is written by this call first
  return aq_mlp_backward(params, x, act, y, target, nsamples, weight_decay, delta, grad, loss, stream);
}

extern "C" int aq_adam_step(float *params, const float *grad, float *m, float *v, int32_t n, int32_t t, float lr,
                            float beta1, float beta2, float eps, void *stream) {
  if (n < 0 || t < 1) return aq_fail(AQ_ERR_INVALID, "bad size or step");
  if (n == 0) return AQ_OK;
  if (!params || !grad || !m || !v) return aq_fail(AQ_ERR_INVALID, "null buffer");
  const float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, params, grad, m, v, n, lr,
                     beta1, beta2, eps, c1, c2);
  MLP_TRY(hipGetLastError());
  return AQ_OK;
}
