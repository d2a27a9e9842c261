// FP32 MLP (in -> h1 -> h2 -> 1, ReLU) for the learned partition cost model:
// forward, mean-squared-error backward and SGD, hand-written for gfx950.
//
// Parameters are packed W1[h1][in], b1[h1], W2[h2][h1], b2[h2], W3[h2], b3.
// Training is deterministic: each workgroup writes the gradient of its 64
// samples to its own workspace row, a second kernel sums the rows in a
// fixed order and applies the update.  The network is tiny (~6K
// parameters), so plain FMA loops over LDS-resident weights are used; it is
// not worth MFMA tiles.
#include <hip/hip_runtime.h>

#include <mutex>

#include "jq_common.hpp"

namespace jqd {
namespace {

constexpr int kMaxIn = 32, kMaxH = 64, kSamples = 64;

// Workspace of jq_mlp_train_step_noalloc: the per-workgroup gradient rows
// and losses of a batch of at most kNoAllocMaxSamples samples at the largest
// dimensions (64 rows of 6337 parameters + 64 losses, 1.6 MB).  Every GPU
// the library runs on holds its own copy until the process ends.
constexpr int kNoAllocMaxSamples = 4096;
constexpr size_t kNoAllocFloats =
    (size_t)(kNoAllocMaxSamples / kSamples) *
        (kMaxH * kMaxIn + kMaxH + kMaxH * kMaxH + kMaxH + kMaxH + 1) +
    kNoAllocMaxSamples / kSamples;
__device__ __align__(256) float scratch[kNoAllocFloats];

struct Dims {
  int in, h1, h2, np;
  __host__ __device__ int oW1() const { return 0; }
  __host__ __device__ int ob1() const { return h1 * in; }
  __host__ __device__ int oW2() const { return ob1() + h1; }
  __host__ __device__ int ob2() const { return oW2() + h2 * h1; }
  __host__ __device__ int oW3() const { return ob2() + h2; }
  __host__ __device__ int ob3() const { return oW3() + h2; }
};

inline Dims dims(int in, int h1, int h2) {
  Dims d{in, h1, h2, 0};
  d.np = d.ob3() + 1;
  return d;
}

__device__ void load_params(const float *p, float *sp, int np) {
  for (int i = threadIdx.x; i < np; i += blockDim.x) sp[i] = p[i];
}

// forward of one sample; h1o/h2o receive the activations when non-null
__device__ float fwd(const Dims &d, const float *sp, const float *x, float *h1o,
                     float *h2o) {
  float h1[kMaxH], h2[kMaxH];
  for (int k = 0; k < d.h1; ++k) {
    float a = sp[d.ob1() + k];
    for (int i = 0; i < d.in; ++i) a = fmaf(sp[d.oW1() + k * d.in + i], x[i], a);
    h1[k] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < d.h2; ++j) {
    float a = sp[d.ob2() + j];
    for (int k = 0; k < d.h1; ++k) a = fmaf(sp[d.oW2() + j * d.h1 + k], h1[k], a);
    h2[j] = a > 0.f ? a : 0.f;
  }
  float y = sp[d.ob3()];
  for (int j = 0; j < d.h2; ++j) y = fmaf(sp[d.oW3() + j], h2[j], y);
  if (h1o)
    for (int k = 0; k < d.h1; ++k) h1o[k] = h1[k];
  if (h2o)
    for (int j = 0; j < d.h2; ++j) h2o[j] = h2[j];
  return y;
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(Dims d, const float *p,
                                                          const float *x, int n,
                                                          float *y) {
  extern __shared__ float sp[];
  load_params(p, sp, d.np);
  __syncthreads();
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) y[s] = fwd(d, sp, x + (size_t)s * d.in, nullptr, nullptr);
}

// one workgroup = kSamples samples: forward/backward of every sample, then
// the workgroup's gradient sums into ws[blockIdx.x][param]
__global__ __launch_bounds__(256) void mlp_grad_kernel(Dims d, const float *p,
                                                       const float *x,
                                                       const float *target, int n,
                                                       float *ws, float *wloss) {
  extern __shared__ float lds[];
  float *sp = lds;
  float *X = sp + d.np;                 // [kSamples][in]
  float *H1 = X + kSamples * kMaxIn;    // [kSamples][h1]
  float *H2 = H1 + kSamples * kMaxH;    // [kSamples][h2]
  float *D1 = H2 + kSamples * kMaxH;    // dL/d(pre-activation 1)
  float *D2 = D1 + kSamples * kMaxH;    // dL/d(pre-activation 2)
  float *DY = D2 + kSamples * kMaxH;    // [kSamples]
  float *LS = DY + kSamples;            // [kSamples] squared errors
  load_params(p, sp, d.np);
  const int t = threadIdx.x;
  const int s0 = blockIdx.x * kSamples;
  for (int i = t; i < kSamples * d.in; i += blockDim.x) {
    int s = s0 + i / d.in;
    X[(i / d.in) * kMaxIn + (i % d.in)] = s < n ? x[(size_t)s * d.in + (i % d.in)] : 0.f;
  }
  __syncthreads();
  if (t < kSamples) {
    int s = s0 + t;
    float *h1 = H1 + t * kMaxH, *h2 = H2 + t * kMaxH;
    float y = fwd(d, sp, X + t * kMaxIn, h1, h2);
    float dy = 0.f, e2 = 0.f;
    if (s < n) {
      float e = y - target[s];
      dy = 2.f * e / (float)n;
      e2 = e * e;
    }
    DY[t] = dy;
    LS[t] = e2;
    float *d2 = D2 + t * kMaxH, *d1 = D1 + t * kMaxH;
    for (int j = 0; j < d.h2; ++j) d2[j] = h2[j] > 0.f ? dy * sp[d.oW3() + j] : 0.f;
    for (int k = 0; k < d.h1; ++k) {
      float a = 0.f;
      for (int j = 0; j < d.h2; ++j) a = fmaf(sp[d.oW2() + j * d.h1 + k], d2[j], a);
      d1[k] = h1[k] > 0.f ? a : 0.f;
    }
  }
  __syncthreads();
  float *row = ws + (size_t)blockIdx.x * d.np;
  for (int q = t; q < d.np; q += blockDim.x) {
    float g = 0.f;
    if (q < d.ob1()) {
      int k = q / d.in, i = q % d.in;
      for (int s = 0; s < kSamples; ++s) g = fmaf(D1[s * kMaxH + k], X[s * kMaxIn + i], g);
    } else if (q < d.oW2()) {
      int k = q - d.ob1();
      for (int s = 0; s < kSamples; ++s) g += D1[s * kMaxH + k];
    } else if (q < d.ob2()) {
      int j = (q - d.oW2()) / d.h1, k = (q - d.oW2()) % d.h1;
      for (int s = 0; s < kSamples; ++s) g = fmaf(D2[s * kMaxH + j], H1[s * kMaxH + k], g);
    } else if (q < d.oW3()) {
      int j = q - d.ob2();
      for (int s = 0; s < kSamples; ++s) g += D2[s * kMaxH + j];
    } else if (q < d.ob3()) {
      int j = q - d.oW3();
      for (int s = 0; s < kSamples; ++s) g = fmaf(DY[s], H2[s * kMaxH + j], g);
    } else {
      for (int s = 0; s < kSamples; ++s) g += DY[s];
    }
    row[q] = g;
  }
  if (t == 0) {
    float l = 0.f;
    for (int s = 0; s < kSamples; ++s) l += LS[s];
    wloss[blockIdx.x] = l;
  }
}

__global__ __launch_bounds__(256) void mlp_update_kernel(Dims d, float *p, float *grads,
                                                         const float *ws,
                                                         const float *wloss, int nblk,
                                                         int n, float lr,
                                                         float *loss_out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < d.np) {
    float g = 0.f;
    for (int b = 0; b < nblk; ++b) g += ws[(size_t)b * d.np + q];
    grads[q] = g;
    p[q] -= lr * g;
  }
  if (q == 0 && loss_out) {
    float l = 0.f;
    for (int b = 0; b < nblk; ++b) l += wloss[b];
    *loss_out = l / (float)n;
  }
}

size_t grad_lds_bytes(const Dims &d) {
  return sizeof(float) * ((size_t)d.np + kSamples * kMaxIn + 4 * kSamples * kMaxH + 2 * kSamples);
}

int check_dims(int in, int h1, int h2) {
  if (in < 1 || in > kMaxIn || h1 < 1 || h1 > kMaxH || h2 < 1 || h2 > kMaxH)
    return set_error("MLP dimensions out of range (in<=32, hidden<=64)");
  return 0;
}

}  // namespace
}  // namespace jqd

using namespace jqd;

extern "C" size_t jq_mlp_num_params(int in, int h1, int h2) { return (size_t)dims(in, h1, h2).np; }

extern "C" size_t jq_mlp_workspace_floats(int in, int h1, int h2, int n) {
  Dims d = dims(in, h1, h2);
  size_t nblk = ((size_t)n + kSamples - 1) / kSamples;
  return nblk * (size_t)d.np + nblk;
}

extern "C" int jq_mlp_forward(const float *params, int in, int h1, int h2,
                              const float *x, int n, float *y, void *stream) {
  if (check_dims(in, h1, h2)) return -1;
  set_error("");
  if (n <= 0) return 0;
  Dims d = dims(in, h1, h2);
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 255) / 256), dim3(256),
                     sizeof(float) * d.np, (hipStream_t)stream, d, params, x, n, y);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(hipGetErrorString(e));
}

extern "C" int jq_mlp_train_step(float *params, float *grads, float *workspace,
                                 int in, int h1, int h2, const float *x,
                                 const float *target, int n, float lr,
                                 float *loss_out, void *stream) {
  if (check_dims(in, h1, h2)) return -1;
  set_error("");
  if (n <= 0) return 0;
  if (!workspace) return set_error("MLP train step: workspace missing");
  Dims d = dims(in, h1, h2);
  int nblk = (n + kSamples - 1) / kSamples;
  size_t lds = grad_lds_bytes(d);
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&mlp_grad_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)grad_lds_bytes(dims(kMaxIn, kMaxH, kMaxH))) != hipSuccess)
      return set_error("MLP LDS attribute failed");
    attr = true;
  }
  float *ws = workspace, *wloss = workspace + (size_t)nblk * d.np;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(nblk), dim3(256), lds, (hipStream_t)stream,
                     d, params, x, target, n, ws, wloss);
  hipLaunchKernelGGL(mlp_update_kernel, dim3((d.np + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, d, params, grads, ws, wloss, nblk, n, lr,
                     loss_out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(hipGetErrorString(e));
}

namespace jqd {
namespace {

// Device address of `scratch` on each device, looked up once per device
// with that device current (the array's host-side name is not a device
// address).
constexpr int kMaxDevices = 64;
float *g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mu;

int scratch_for_current_device(float **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return set_error("MLP train step (no-alloc): no usable current device");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error("MLP train step (no-alloc): hipGetSymbolAddress(scratch) failed");
    g_scratch_ptr[dev] = static_cast<float *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace
}  // namespace jqd

extern "C" size_t jq_mlp_noalloc_max_samples(void) { return (size_t)kNoAllocMaxSamples; }

extern "C" int jq_mlp_train_step_noalloc(float *params, float *grads, int in, int h1,
                                         int h2, const float *x, const float *target,
                                         int n, float lr, float *loss_out,
                                         void *stream) {
  if (check_dims(in, h1, h2)) return -1;
  if (n > kNoAllocMaxSamples)
    return set_error("MLP train step (no-alloc): batch larger than 4096 samples");
  set_error("");
  if (n <= 0) return 0;
  float *ws = nullptr;
  if (scratch_for_current_device(&ws)) return -1;
  if (jq_mlp_workspace_floats(in, h1, h2, n) > kNoAllocFloats)
    return set_error("MLP train step (no-alloc): workspace exceeds the static buffer");
  return jq_mlp_train_step(params, grads, ws, in, h1, h2, x, target, n, lr, loss_out,
                           stream);
}
