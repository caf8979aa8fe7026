// K4: fused control-variate combine.
//
// One streaming pass over the per-pixel CV state replaces the ~4 scheduled
// kernels of the reference combine (dr.schedule at plugins/cv_integrator.py:
// 64,66,71 and plugins/welford.py:75):
//   weights   plugins/cv_integrator.py:62-68  (StatisticsEstimator.get,
//             plugins/welford.py:77-78)
//   blend     plugins/cv_integrator.py:70
//   update    plugins/welford.py:59-75 (post- or pre-update ordering,
//             plugins/metaintegrator.py:95-96,110-111)
//   it += 1   plugins/cv_integrator.py:76
//
// Numerics: every multiply/add/divide is rounded separately, in the order of
// the Python expressions (fp-contract off), and bias corrections 1 - beta^n
// are evaluated in double and rounded once to float. Results therefore match
// a float32 NumPy restatement bit for bit.
//
// Memory: per element 2 inputs (H, F) + 8 state arrays read + up to 8 written
// = 72 bytes; float4 vector accesses, grid-stride free (one float4 per
// thread), bound by HBM bandwidth.
//
// Counters live in device memory (counters[0] = it, counters[1] = n); the
// last block to finish advances them, so the launch is graph-replayable.

#include "common.hpp"

#pragma clang fp contract(off)

namespace rcv {
namespace {

struct CombineScalars {
  int it;
  int n_get;       // n used by stats.get()
  int do_update;   // it > 0
  int do_weights;  // it > warmup
  float b_get;     // 1 - beta2^n_get
  float a_upd;     // 1 - beta1^n        (before the update)
  float a_upd2;    // 1 - beta1^(n+1)    (after the update)
  int n_pos;       // n > 0 before the update
};

__device__ inline float one_minus_pow(double log_beta, int n) {
  // 1 - beta^n evaluated in double, rounded once.
  return n > 0 ? float(1.0 - exp(double(n) * log_beta)) : 0.0f;
}

struct Elem {
  float img, vn, w, mx, my, vx, vy, cv;
};

__device__ inline void stats_update(Elem &e, float h, float f, const CombineScalars &s,
                                    float b1, float b2) {
  // plugins/welford.py:61-73
  const float dx1 = s.n_pos ? h - e.mx / s.a_upd : h;
  const float dy1 = s.n_pos ? f - e.my / s.a_upd : f;
  e.my = b1 * e.my + (1.0f - b1) * f;
  e.mx = b1 * e.mx + (1.0f - b1) * h;
  const float dx2 = h - e.mx / s.a_upd2;
  const float dy2 = f - e.my / s.a_upd2;
  const float c = 1.0f - b2;
  e.vx = b2 * e.vx + (c * dx1) * dx2;
  e.vy = b2 * e.vy + (c * dy1) * dy2;
  e.cv = b2 * e.cv + (c * dx1) * dy2;
}

__device__ inline void combine_one(Elem &e, float h, float f, const CombineScalars &s,
                                   float b1, float b2, bool pre) {
  if (pre && s.do_update) stats_update(e, h, f, s, b1, b2);
  if (s.do_weights) {
    const float v0 = e.vx / s.b_get;
    const float v1 = e.vy / s.b_get;
    const float c = e.cv / s.b_get;
    e.vn = (e.w * e.w) * (e.vn + v0) + v1 - (2.0f * e.w) * c;
    float w = c / (v0 + e.vn);
    if (__builtin_isnan(w) || __builtin_isinf(w)) w = 0.0f;
    e.w = fminf(fmaxf(w, 0.0f), 1.0f);
  }
  e.img = e.w * (e.img - h) + f;
  if (!pre && s.do_update) stats_update(e, h, f, s, b1, b2);
}

__global__ __launch_bounds__(256) void cv_combine_kernel(
    rcv_cv_state st, const float *__restrict__ H, const float *__restrict__ F, int64_t n_elem,
    int warmup, float b1, float b2, double log_b1, double log_b2, int pre, int vec4) {
  __shared__ CombineScalars sh;
  if (threadIdx.x == 0) {
    const int it = __hip_atomic_load(&st.counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int n = __hip_atomic_load(&st.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    CombineScalars s;
    s.it = it;
    s.do_update = it > 0;
    s.do_weights = it > warmup;
    s.n_pos = n > 0;
    s.a_upd = one_minus_pow(log_b1, n);
    s.a_upd2 = one_minus_pow(log_b1, n + 1);
    s.n_get = (pre && s.do_update) ? n + 1 : n;
    s.b_get = one_minus_pow(log_b2, s.n_get);
    sh = s;
  }
  __syncthreads();
  const CombineScalars s = sh;
  const bool writes_w = s.do_weights;
  const bool writes_stats = s.do_update;

  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (vec4) {
    const int64_t i = tid * 4;
    if (i < n_elem) {
      const float4 h4 = *reinterpret_cast<const float4 *>(H + i);
      const float4 f4 = *reinterpret_cast<const float4 *>(F + i);
      float4 a[8];
      float *ptrs[8] = {st.img, st.v_n, st.w_s, st.mean_x, st.mean_y, st.var_x, st.var_y, st.cov};
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = *reinterpret_cast<const float4 *>(ptrs[k] + i);
      const float hh[4] = {h4.x, h4.y, h4.z, h4.w};
      const float ff[4] = {f4.x, f4.y, f4.z, f4.w};
      float out[8][4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        Elem e;
        e.img = (&a[0].x)[j];
        e.vn = (&a[1].x)[j];
        e.w = (&a[2].x)[j];
        e.mx = (&a[3].x)[j];
        e.my = (&a[4].x)[j];
        e.vx = (&a[5].x)[j];
        e.vy = (&a[6].x)[j];
        e.cv = (&a[7].x)[j];
        combine_one(e, hh[j], ff[j], s, b1, b2, pre != 0);
        out[0][j] = e.img;
        out[1][j] = e.vn;
        out[2][j] = e.w;
        out[3][j] = e.mx;
        out[4][j] = e.my;
        out[5][j] = e.vx;
        out[6][j] = e.vy;
        out[7][j] = e.cv;
      }
      *reinterpret_cast<float4 *>(st.img + i) = make_float4(out[0][0], out[0][1], out[0][2], out[0][3]);
      if (writes_w) {
        *reinterpret_cast<float4 *>(st.v_n + i) = make_float4(out[1][0], out[1][1], out[1][2], out[1][3]);
        *reinterpret_cast<float4 *>(st.w_s + i) = make_float4(out[2][0], out[2][1], out[2][2], out[2][3]);
      }
      if (writes_stats) {
#pragma unroll
        for (int k = 3; k < 8; ++k)
          *reinterpret_cast<float4 *>(ptrs[k] + i) = make_float4(out[k][0], out[k][1], out[k][2], out[k][3]);
      }
    }
  } else {
    for (int64_t i = tid * 4; i < n_elem && i < tid * 4 + 4; ++i) {
      Elem e{st.img[i], st.v_n[i], st.w_s[i], st.mean_x[i],
             st.mean_y[i], st.var_x[i], st.var_y[i], st.cov[i]};
      combine_one(e, H[i], F[i], s, b1, b2, pre != 0);
      st.img[i] = e.img;
      if (writes_w) {
        st.v_n[i] = e.vn;
        st.w_s[i] = e.w;
      }
      if (writes_stats) {
        st.mean_x[i] = e.mx;
        st.mean_y[i] = e.my;
        st.var_x[i] = e.vx;
        st.var_y[i] = e.vy;
        st.cov[i] = e.cv;
      }
    }
  }

  // Last block advances it and n (plugins/cv_integrator.py:76, welford.py:67).
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int done = atomicAdd(&st.counters[2], 1);
    if (done == int(gridDim.x) - 1) {
      st.counters[0] = s.it + 1;
      if (s.do_update) st.counters[1] = st.counters[1] + 1;
      st.counters[2] = 0;
      __threadfence();
    }
  }
}

__global__ __launch_bounds__(256) void welford_update_kernel(
    float *__restrict__ mean, float *__restrict__ var, int32_t *counter,
    const float *__restrict__ x, int64_t n_elem, float b1, float b2, double log_b1) {
  __shared__ float sh_a, sh_a2;
  __shared__ int sh_npos, sh_n;
  if (threadIdx.x == 0) {
    const int n = __hip_atomic_load(&counter[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sh_n = n;
    sh_npos = n > 0;
    sh_a = one_minus_pow(log_b1, n);
    sh_a2 = one_minus_pow(log_b1, n + 1);
  }
  __syncthreads();
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n_elem) {
    // plugins/welford.py:28-32
    const float xi = x[i];
    float m = mean[i];
    const float dx1 = sh_npos ? xi - m / sh_a : xi;
    m = b1 * m + (1.0f - b1) * xi;
    const float dx2 = xi - m / sh_a2;
    mean[i] = m;
    var[i] = b2 * var[i] + ((1.0f - b2) * dx1) * dx2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int done = atomicAdd(&counter[1], 1);
    if (done == int(gridDim.x) - 1) {
      counter[0] = sh_n + 1;
      counter[1] = 0;
      __threadfence();
    }
  }
}

}  // namespace
}  // namespace rcv

using namespace rcv;

extern "C" int rcv_cv_step(const rcv_cv_state *state, const float *H, const float *F,
                           int64_t n_elem, int32_t warmup, float beta1, float beta2,
                           int32_t pre_update, void *stream) {
  RCV_REQUIRE(state && H && F, "rcv_cv_step: null pointer");
  RCV_REQUIRE(state->img && state->v_n && state->w_s && state->mean_x && state->mean_y &&
                  state->var_x && state->var_y && state->cov && state->counters,
              "rcv_cv_step: state has a null array");
  RCV_REQUIRE(n_elem > 0, "rcv_cv_step: n_elem must be positive (got %lld)", (long long)n_elem);
  RCV_REQUIRE(beta1 > 0.f && beta1 < 1.f && beta2 > 0.f && beta2 < 1.f,
              "rcv_cv_step: betas must lie in (0, 1)");
  bool vec4 = (n_elem % 4) == 0 && aligned16(H) && aligned16(F);
  const float *arrs[8] = {state->img, state->v_n, state->w_s, state->mean_x,
                          state->mean_y, state->var_x, state->var_y, state->cov};
  for (auto p : arrs) vec4 = vec4 && aligned16(p);
  const int block = 256;
  const unsigned grid = ceil_div(ceil_div(n_elem, 4), block);
  hipLaunchKernelGGL(cv_combine_kernel, dim3(grid), dim3(block), 0, as_stream(stream), *state, H,
                     F, n_elem, int(warmup), beta1, beta2, log(double(beta1)),
                     log(double(beta2)), int(pre_update != 0), int(vec4));
  return check_launch("rcv_cv_step");
}

extern "C" int rcv_welford_update(float *mean, float *var, int32_t *counter, const float *x,
                                  int64_t n_elem, float beta1, float beta2, void *stream) {
  RCV_REQUIRE(mean && var && counter && x, "rcv_welford_update: null pointer");
  RCV_REQUIRE(n_elem > 0, "rcv_welford_update: n_elem must be positive");
  const int block = 256;
  const unsigned grid = ceil_div(n_elem, block);
  hipLaunchKernelGGL(welford_update_kernel, dim3(grid), dim3(block), 0, as_stream(stream), mean,
                     var, counter, x, n_elem, beta1, beta2, log(double(beta1)));
  return check_launch("rcv_welford_update");
}
