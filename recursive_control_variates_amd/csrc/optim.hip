// K6 (image loss + its gradient) and K7 (Adam step fused with clamping).
//
// Loss: mean(|x - y|) or mean((x - y)^2) over all elements (utils.py:21-25),
// reduced deterministically: fixed grid of partial sums in a library
// workspace, summed in index order by the last block to finish. The gradient
// (sign(x - y)/N or 2(x - y)/N) is written in the same pass.
//
// Adam: Kingma & Ba (2015) in the form of their Sec. 2 remark,
//   lr_t = lr * sqrt(1 - beta2^t) / (1 - beta1^t)
//   m = beta1 m + (1 - beta1) g ;  v = beta2 v + (1 - beta2) g^2
//   p' = clamp(p - lr_t m / (sqrt(v) + eps), lo, hi)       (utils.py:172-175)
// one streaming pass; the output may go to a different buffer than the input
// so that parameter states rotate without copies.

#include <mutex>

#include "common.hpp"

namespace rcv {
namespace {

constexpr int kLossBlocks = 256;
constexpr int kLossThreads = 256;

struct LossWorkspace {
  float *partials = nullptr;  // kLossBlocks floats
  unsigned *done = nullptr;
};

LossWorkspace &loss_workspace(int device, int &rc) {
  static std::mutex mu;
  static LossWorkspace ws[64];
  std::lock_guard<std::mutex> lock(mu);
  rc = RCV_OK;
  LossWorkspace &w = ws[device & 63];
  if (!w.partials) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, kLossBlocks * sizeof(float) + 256);
    if (e != hipSuccess) {
      rc = hip_fail(e, "rcv_loss: workspace allocation");
      return w;
    }
    e = hipMemset(p, 0, kLossBlocks * sizeof(float) + 256);
    if (e != hipSuccess) {
      rc = hip_fail(e, "rcv_loss: workspace init");
      return w;
    }
    w.partials = static_cast<float *>(p);
    w.done = reinterpret_cast<unsigned *>(static_cast<char *>(p) + kLossBlocks * sizeof(float));
  }
  return w;
}

__global__ __launch_bounds__(kLossThreads) void loss_kernel(const float *__restrict__ x, const float *__restrict__ y,
                                                            int64_t n, int kind, float *__restrict__ loss_out,
                                                            float *__restrict__ grad, float *partials,
                                                            unsigned *done) {
  __shared__ float red[kLossThreads];
  const float inv_n = 1.0f / float(n);
  float acc = 0.0f;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float d = x[i] - y[i];
    if (kind == RCV_LOSS_L1) {
      acc += fabsf(d);
      if (grad) grad[i] = (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f)) * inv_n;
    } else {
      acc += d * d;
      if (grad) grad[i] = 2.0f * d * inv_n;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kLossThreads / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  __shared__ bool last;
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = red[0];
    __threadfence();
    last = atomicAdd(done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (last) {
    __threadfence();
    red[threadIdx.x] = int(threadIdx.x) < int(gridDim.x)
                           ? __hip_atomic_load(&partials[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : 0.0f;
    __syncthreads();
    for (int s = kLossThreads / 2; s > 0; s >>= 1) {
      if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      loss_out[0] = red[0] * inv_n;
      *done = 0u;
    }
  }
}

__global__ __launch_bounds__(256) void adam_kernel(float *p_out, const float *p_in,
                                                   const float *__restrict__ g, float *__restrict__ m,
                                                   float *__restrict__ v, int64_t n, float lr_t, float b1, float b2,
                                                   float eps, float lo, float hi) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  const float upd = p_in[i] - lr_t * mi / (sqrtf(vi) + eps);
  p_out[i] = fminf(fmaxf(upd, lo), hi);
}

// Scratch of rcv_loss_noalloc: the per-block partial sums and the
// last-block counter of loss_kernel in a file-scope device array instead of
// a library allocation, so the entry point allocates nothing. Every device
// the library runs on holds its own copy (a few KiB) until the process ends.
// Host code never passes the array's name to a kernel: it takes the device
// address from hipGetSymbolAddress, once per device.
constexpr int kNoAllocScratchFloats = kLossBlocks + 64;  // partials, then the counter (padded)

__global__ void zero_scratch_kernel(float *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0f;
}

}  // namespace

static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace {
// Device address of `scratch` on the current device, looked up at the first
// call on that device (with it current) and kept per device.
int noalloc_scratch(float *&out) {
  static std::mutex mu;
  static float *ptr[64] = {};
  int dev = 0;
  RCV_TRY_HIP(hipGetDevice(&dev), "rcv_loss_noalloc: hipGetDevice");
  RCV_REQUIRE(dev >= 0 && dev < 64, "rcv_loss_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void *p = nullptr;
    RCV_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "rcv_loss_noalloc: hipGetSymbolAddress");
    RCV_REQUIRE(p, "rcv_loss_noalloc: no device address for the scratch array");
    ptr[dev] = static_cast<float *>(p);
  }
  out = ptr[dev];
  return RCV_OK;
}

}  // namespace
}  // namespace rcv

using namespace rcv;

extern "C" int rcv_loss_noalloc(const float *img, const float *ref, int64_t n_elem, int32_t kind, float *loss_out,
                                float *grad_out, void *stream) {
  RCV_REQUIRE(img && ref && loss_out, "rcv_loss_noalloc: null pointer");
  RCV_REQUIRE(n_elem > 0, "rcv_loss_noalloc: n_elem must be positive");
  RCV_REQUIRE(kind == RCV_LOSS_L1 || kind == RCV_LOSS_L2, "rcv_loss_noalloc: unknown loss kind %d", kind);
  const unsigned grid = unsigned(std::min<int64_t>(kLossBlocks, (n_elem + kLossThreads - 1) / kLossThreads));
  // the scratch holds kLossBlocks partial sums; the grid never exceeds it
  RCV_REQUIRE(int(grid) <= kLossBlocks, "rcv_loss_noalloc: %u blocks exceed the scratch of %d partial sums", grid,
              kLossBlocks);
  float *ws = nullptr;
  if (int rc = noalloc_scratch(ws)) return rc;
  // the last-block counter must start at zero (the array keeps the previous
  // call's contents); a fill kernel, not a memset, so the call can be captured
  // in a graph. The partial sums are written before they are read.
  hipLaunchKernelGGL(zero_scratch_kernel, dim3(1), dim3(64), 0, as_stream(stream), ws + kLossBlocks, 64);
  hipLaunchKernelGGL(loss_kernel, dim3(grid), dim3(kLossThreads), 0, as_stream(stream), img, ref, n_elem, kind,
                     loss_out, grad_out, ws, reinterpret_cast<unsigned *>(ws + kLossBlocks));
  return check_launch("rcv_loss_noalloc");
}

extern "C" int rcv_loss(const float *img, const float *ref, int64_t n_elem, int32_t kind, float *loss_out,
                        float *grad_out, void *stream) {
  RCV_REQUIRE(img && ref && loss_out, "rcv_loss: null pointer");
  RCV_REQUIRE(n_elem > 0, "rcv_loss: n_elem must be positive");
  RCV_REQUIRE(kind == RCV_LOSS_L1 || kind == RCV_LOSS_L2, "rcv_loss: unknown loss kind %d", kind);
  int dev = 0;
  RCV_TRY_HIP(hipGetDevice(&dev), "rcv_loss: hipGetDevice");
  int rc;
  LossWorkspace &ws = loss_workspace(dev, rc);
  if (rc != RCV_OK) return rc;
  const unsigned grid = unsigned(std::min<int64_t>(kLossBlocks, (n_elem + kLossThreads - 1) / kLossThreads));
  hipLaunchKernelGGL(loss_kernel, dim3(grid), dim3(kLossThreads), 0, as_stream(stream), img, ref, n_elem, kind,
                     loss_out, grad_out, ws.partials, ws.done);
  return check_launch("rcv_loss");
}

extern "C" int rcv_adam_step(float *param_out, const float *param_in, const float *grad, float *m, float *v,
                             int64_t n_elem, int32_t t, float lr, float beta1, float beta2, float eps,
                             float clamp_lo, float clamp_hi, void *stream) {
  RCV_REQUIRE(param_out && param_in && grad && m && v, "rcv_adam_step: null pointer");
  RCV_REQUIRE(n_elem > 0 && t >= 1, "rcv_adam_step: need n_elem > 0 and t >= 1");
  RCV_REQUIRE(!(clamp_lo > clamp_hi), "rcv_adam_step: clamp_lo > clamp_hi");
  const double lr_t = double(lr) * std::sqrt(1.0 - std::pow(double(beta2), t)) / (1.0 - std::pow(double(beta1), t));
  hipLaunchKernelGGL(adam_kernel, dim3(ceil_div(n_elem, 256)), dim3(256), 0, as_stream(stream), param_out, param_in,
                     grad, m, v, n_elem, float(lr_t), beta1, beta2, eps, clamp_lo, clamp_hi);
  return check_launch("rcv_adam_step");
}
