// Batched optimisation: many independent CV optimisations (jobs) advanced
// by one device-resident iteration.
//
// One iteration of every job is a fixed sequence of launches that reads all
// per-iteration state from device memory (iteration counters, the two-state
// parameter ring, majorants, Adam step counts), so the sequence can be
// captured once into a hipGraph and replayed:
//   1. rotate   old <- new <- next for every job whose previous step
//               produced a `next` (the TwoState parameters_changed of
//               plugins/twostatebsdf.py:46-56 / plugins/twostatemedium.py:
//               189-200 as pointer rotation); reset per-iteration scratch
//   1b. supergrid  majorant supergrids of both states of every optimised
//               grid that has one (plugins/twostatemedium.py:25-70), rebuilt
//               from the rotated ring
//   2. render   persistent work-queue kernels over all jobs' (pixel, sample-
//               chunk) units: two-state for cv_ps jobs, plain for baseline
//               jobs (plugins/metaintegrator.py:68-88), surfaces and volumes
//   3. combine  fused CV/Welford combine per job (as rcv_cv_step) or copy of
//               the plain image for baseline jobs
//   4. loss     L1/L2 gradient image per job (utils.py:21-25)
//   5. adjoint  persistent path-replay adjoint over all jobs' units
//   6. adam     Adam + clamp into the ring's `next` buffer, with the
//               majorant (max) of the result reduced on the fly
//   7. advance  iteration counters
// Work units of a job are padded to multiples of 64 so a wavefront never
// spans two jobs (its job is wave-uniform); element ranges are padded to the
// 256-thread block. The random stream of a sample depends only on the job's
// seed, its iteration, the pixel and the sample, so a job's images are the
// same whether it runs alone or in a batch.

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "path.hpp"
#include "render_internal.hpp"
#include "volpath.hpp"

namespace rcv {

constexpr int kWave = 64;
constexpr int kElemBlock = 256;
constexpr int kBatchMaxSpt = 32;         // camera samples per work unit (at most)
constexpr int64_t kBatchMinThreads = 0;  // the batch, not one job, fills the device

struct BatchJobDev {
  SceneView view;       // params[param].data[] rotated on the device
  IntegArgs integ;
  int32_t itype;        // RCV_INTEGRATOR_*
  int32_t method;       // 0 baseline (plain), 1 cv_ps (two-state + combine)
  int32_t loss;         // RCV_LOSS_*
  int32_t sensor;
  int32_t spp, spt, cl2;
  int32_t spp_g, spt_g, cl2_g;
  int32_t film_w, n_pix;
  uint32_t seed;
  int32_t param;
  int64_t n_param;
  int64_t n_img;
  int64_t unit_begin, unit_begin_g;  // in this job's render / adjoint class space
  int64_t elem_begin, pelem_begin;   // padded image / parameter element offsets
  float *ring[3];
  int32_t ring_old, ring_new, ring_next;
  uint32_t *maj_keys;   // scene majorant keys of `param`: [OLD], [NEW]
  uint32_t next_key;    // max key of the ring's next buffer (written by adam)
  int32_t iter;
  int32_t active;       // 0: the job sits out this iteration (no state changes)
  float *H, *F, *img, *grad_img;
  const float *img_ref;
  float *cv[8];
  int32_t warmup;
  float cv_b1, cv_b2;
  double cv_log_b1, cv_log_b2;
  int32_t pre_update;
  float *grad, *m, *v;
  GradTargets gt;       // gt.g[param] = grad, others null
  float lr, b1, b2, eps, clamp_lo, clamp_hi;
  double log_b1, log_b2;
  float *sg_out;        // supergrids [OLD][NEW] of the optimised grid's medium, or null
  int32_t sg_res[3];
};

namespace {

__device__ inline int find_job(const int64_t *__restrict__ ends, int n, int64_t u) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ends[mid] <= u)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ inline uint32_t float_key_dev(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline uint32_t iter_seed(const BatchJobDev &J) { return J.seed + uint32_t(J.iter); }

__device__ inline uint32_t iter_seed_grad(const BatchJobDev &J) {
  uint32_t a = iter_seed(J), b = 1u;
  tea(a, b);
  return a;
}

// ---------------------------------------------------------------- 1. rotate
__global__ void batch_rotate_kernel(BatchJobDev *jobs, int n_jobs, unsigned *queues) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < 6) queues[j] = 0u;
  if (j >= n_jobs) return;
  BatchJobDev &J = jobs[j];
  if (!J.active) return;
  if (J.iter > 0) {
    const int old_slot = J.ring_old;
    J.ring_old = J.ring_new;
    J.ring_new = J.ring_next;
    J.ring_next = old_slot;
    J.view.params[J.param].data[RCV_STATE_OLD] = J.ring[J.ring_old];
    J.view.params[J.param].data[RCV_STATE_NEW] = J.ring[J.ring_new];
    J.maj_keys[RCV_STATE_OLD] = J.maj_keys[RCV_STATE_NEW];
    J.maj_keys[RCV_STATE_NEW] = J.next_key;
  }
  J.next_key = float_key_dev(-kInf);
}

// ----------------------------------------------------------- 1b. supergrid
// One thread per (job, state, cell) over the jobs whose optimised grid has a
// majorant supergrid; ends[] are the cumulative 2 * cells counts.
__global__ void batch_supergrid_kernel(const BatchJobDev *__restrict__ jobs, const int32_t *__restrict__ ids,
                                       const int64_t *__restrict__ ends, int n, int64_t total) {
  const int64_t u = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= total) return;
  const int k = find_job(ends, n, u);
  const BatchJobDev &J = jobs[ids[k]];
  if (!J.active) return;
  const int cells = J.sg_res[0] * J.sg_res[1] * J.sg_res[2];
  const int local = int(u - (k > 0 ? ends[k - 1] : 0));
  const int state = local >= cells ? RCV_STATE_NEW : RCV_STATE_OLD;
  const int c = local - (state == RCV_STATE_NEW ? cells : 0);
  const ParamBind &P = J.view.params[J.param];
  J.sg_out[state * cells + c] =
      supergrid_cell_max(P.data[state], P.nx, P.ny, P.nz, J.sg_res[0], J.sg_res[1], J.sg_res[2], c);
}

// ---------------------------------------------------------------- 2. render
template <int NS, int INTEG, int F>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kWavesPerEuRender<NS, INTEG>))) void batch_render_kernel(const BatchJobDev *__restrict__ jobs, const int32_t *__restrict__ ids,
                                    const int64_t *__restrict__ ends, int n, int64_t total, unsigned *queue,
                                    int grab) {
  extern __shared__ int lds_stack[];
  const int lane = threadIdx.x & (kWave - 1);
  while (true) {
    unsigned b = 0;
    if (lane == 0) b = atomicAdd(queue, unsigned(grab));
    const int64_t base = int64_t(__shfl(b, 0, kWave));
    if (base >= total) break;
    const int64_t stop = min(base + int64_t(grab), total);
    for (int64_t u0 = base; u0 < stop; u0 += kWave) {
      const int k = __builtin_amdgcn_readfirstlane(find_job(ends, n, u0));
      const BatchJobDev &J = jobs[ids[k]];
      if (!J.active) continue;
      const SceneView &S = J.view;
      const LdsStack stack = lds_of(S, lds_stack);
      const int64_t local = u0 + lane - J.unit_begin;
      const int64_t pixel = local >> J.cl2;
      const int chunk = int(local & ((1 << J.cl2) - 1));
      V3 acc[NS];
#pragma unroll
      for (int q = 0; q < NS; ++q) acc[q] = v3(0.f);
      LocalCounters cnt{nullptr};
      const SensorInfo &sen = S.sensors[J.sensor];
      const uint32_t seed = iter_seed(J);
      if (pixel < J.n_pix) {
        const int px = int(pixel % J.film_w), py = int(pixel / J.film_w);
        const int s_end = sample_end(chunk, J.spp, J.cl2);
        if (INTEG == RCV_INTEGRATOR_VOLPATH)
          volpath_render_samples<NS, F>(S, J.integ, sen, seed, px, py, pixel, J.spp,
                                        sample_end(chunk - 1, J.spp, J.cl2), s_end, acc, stack, cnt);
        else for (int s = sample_end(chunk - 1, J.spp, J.cl2); s < s_end; ++s) {
          Pcg32 rng = lane_rng(seed, uint32_t(pixel * J.spp + s));
          const Ray ray = camera_ray(sen, px, py, rng);
          V3 L[NS];
          path_sample<NS, false>(S, J.integ, rng, ray, L, stack, cnt, nullptr, nullptr);
#pragma unroll
          for (int q = 0; q < NS; ++q) acc[q] += L[q];
        }
      }
      for (int off = 1; off < (1 << J.cl2); off <<= 1) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          acc[q].x += __shfl_xor(acc[q].x, off, kWave);
          acc[q].y += __shfl_xor(acc[q].y, off, kWave);
          acc[q].z += __shfl_xor(acc[q].z, off, kWave);
        }
      }
      if (pixel < J.n_pix && chunk == 0) {
        const float inv = float(J.spp);
        float *o0 = (NS == 2 ? J.H : J.F) + 3 * pixel;
        o0[0] = acc[0].x / inv;
        o0[1] = acc[0].y / inv;
        o0[2] = acc[0].z / inv;
        if (NS == 2) {
          float *o1 = J.F + 3 * pixel;
          o1[0] = acc[NS - 1].x / inv;
          o1[1] = acc[NS - 1].y / inv;
          o1[2] = acc[NS - 1].z / inv;
        }
      }
    }
  }
}

// ----------------------------------------------- 3+4. combine, loss gradient
// One block of 256 elements belongs to one job (element ranges are padded).
__global__ __launch_bounds__(kElemBlock) void batch_combine_loss_kernel(BatchJobDev *__restrict__ jobs,
                                                                        const int64_t *__restrict__ ends,
                                                                        int n_jobs) {
#pragma clang fp contract(off)
  const int64_t e0 = int64_t(blockIdx.x) * kElemBlock;
  const int j = __builtin_amdgcn_readfirstlane(find_job(ends, n_jobs, e0));
  const BatchJobDev &J = jobs[j];
  if (!J.active) return;
  const int64_t i = e0 + threadIdx.x - J.elem_begin;
  if (i >= J.n_img) return;
  const float f = J.F[i];
  float out;
  if (J.method == 1) {
    // plugins/cv_integrator.py:60-76 with it = iter, n = max(iter - 1, 0)
    const int it = J.iter;
    const int n = it > 0 ? it - 1 : 0;
    const bool pre = J.pre_update != 0;
    const bool do_update = it > 0;
    const bool do_weights = it > J.warmup;
    const float a_upd = n > 0 ? float(1.0 - exp(double(n) * J.cv_log_b1)) : 0.0f;
    const float a_upd2 = float(1.0 - exp(double(n + 1) * J.cv_log_b1));
    const int n_get = (pre && do_update) ? n + 1 : n;
    const float b_get = n_get > 0 ? float(1.0 - exp(double(n_get) * J.cv_log_b2)) : 0.0f;
    const float h = J.H[i];
    float img = J.cv[0][i], vn = J.cv[1][i], w = J.cv[2][i];
    float mx = J.cv[3][i], my = J.cv[4][i], vx = J.cv[5][i], vy = J.cv[6][i], cv = J.cv[7][i];
    const float b1 = J.cv_b1, b2 = J.cv_b2;
    auto update = [&]() {
      const float dx1 = n > 0 ? h - mx / a_upd : h;
      const float dy1 = n > 0 ? f - my / a_upd : f;
      my = b1 * my + (1.0f - b1) * f;
      mx = b1 * mx + (1.0f - b1) * h;
      const float dx2 = h - mx / a_upd2;
      const float dy2 = f - my / a_upd2;
      const float c = 1.0f - b2;
      vx = b2 * vx + (c * dx1) * dx2;
      vy = b2 * vy + (c * dy1) * dy2;
      cv = b2 * cv + (c * dx1) * dy2;
    };
    if (pre && do_update) update();
    if (do_weights) {
      const float v0 = vx / b_get, v1 = vy / b_get, c = cv / b_get;
      vn = (w * w) * (vn + v0) + v1 - (2.0f * w) * c;
      float ww = c / (v0 + vn);
      if (__builtin_isnan(ww) || __builtin_isinf(ww)) ww = 0.0f;
      w = fminf(fmaxf(ww, 0.0f), 1.0f);
    }
    img = w * (img - h) + f;
    if (!pre && do_update) update();
    J.cv[0][i] = img;
    J.cv[1][i] = vn;
    J.cv[2][i] = w;
    J.cv[3][i] = mx;
    J.cv[4][i] = my;
    J.cv[5][i] = vx;
    J.cv[6][i] = vy;
    J.cv[7][i] = cv;
    out = img;
  } else {
    out = f;
  }
  J.img[i] = out;
  const float d = out - J.img_ref[i];
  const float inv_n = 1.0f / float(J.n_img);
  J.grad_img[i] = J.loss == RCV_LOSS_L1 ? (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f)) * inv_n : 2.0f * d * inv_n;
}

// ---------------------------------------------------------- 5. adjoint
__global__ __launch_bounds__(kElemBlock) void batch_zero_grad_kernel(BatchJobDev *__restrict__ jobs,
                                                                     const int64_t *__restrict__ ends, int n_jobs) {
  const int64_t e0 = int64_t(blockIdx.x) * kElemBlock;
  const int j = __builtin_amdgcn_readfirstlane(find_job(ends, n_jobs, e0));
  const BatchJobDev &J = jobs[j];
  if (!J.active) return;
  const int64_t i = e0 + threadIdx.x - J.pelem_begin;
  if (i < J.n_param) J.grad[i] = 0.0f;
}

template <int INTEG, int F>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<INTEG>))) void batch_adjoint_kernel(const BatchJobDev *__restrict__ jobs, const int32_t *__restrict__ ids,
                                     const int64_t *__restrict__ ends, int n, int64_t total, unsigned *queue,
                                     int grab) {
  extern __shared__ int lds_stack[];
  const int lane = threadIdx.x & (kWave - 1);
  while (true) {
    unsigned b = 0;
    if (lane == 0) b = atomicAdd(queue, unsigned(grab));
    const int64_t base = int64_t(__shfl(b, 0, kWave));
    if (base >= total) break;
    const int64_t stop = min(base + int64_t(grab), total);
    for (int64_t u0 = base; u0 < stop; u0 += kWave) {
      const int k = __builtin_amdgcn_readfirstlane(find_job(ends, n, u0));
      const BatchJobDev &J = jobs[ids[k]];
      if (!J.active) continue;
      const SceneView &S = J.view;
      const LdsStack stack = lds_of(S, lds_stack);
      const int64_t local = u0 + lane - J.unit_begin_g;
      const int64_t pixel = local >> J.cl2_g;
      const int chunk = int(local & ((1 << J.cl2_g) - 1));
      LocalCounters cnt{nullptr};
      if (pixel < J.n_pix) {
        const SensorInfo &sen = S.sensors[J.sensor];
        const int px = int(pixel % J.film_w), py = int(pixel / J.film_w);
        const float inv = float(J.spp_g);
        const V3 dL = v3(J.grad_img[3 * pixel] / inv, J.grad_img[3 * pixel + 1] / inv,
                         J.grad_img[3 * pixel + 2] / inv);
        const uint32_t seed = iter_seed_grad(J);
        if (any_nonzero(dL) && INTEG == RCV_INTEGRATOR_VOLPATH) {
          volpath_adjoint_samples<F>(S, J.integ, sen, seed, px, py, pixel, J.spp_g,
                                     sample_end(chunk - 1, J.spp_g, J.cl2_g), sample_end(chunk, J.spp_g, J.cl2_g), dL,
                                     &J.gt, stack, cnt, nullptr);
        } else if (any_nonzero(dL)) {
          const int s_end = sample_end(chunk, J.spp_g, J.cl2_g);
          for (int s = sample_end(chunk - 1, J.spp_g, J.cl2_g); s < s_end; ++s) {
            Pcg32 rng = lane_rng(seed, uint32_t(pixel * J.spp_g + s));
            const Ray ray = camera_ray(sen, px, py, rng);
            const Pcg32 rng0 = rng;
            V3 L[1];
            path_sample<1, false>(S, J.integ, rng, ray, L, stack, cnt, nullptr, nullptr);
            if (!any_nonzero(L[0])) continue;
            AdjState adj{dL, L[0]};
            rng = rng0;
            V3 L2[1];
            path_sample<1, true>(S, J.integ, rng, ray, L2, stack, cnt, &adj, &J.gt);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------- 6. adam
__global__ __launch_bounds__(kElemBlock) void batch_adam_kernel(BatchJobDev *__restrict__ jobs,
                                                                const int64_t *__restrict__ ends, int n_jobs) {
  __shared__ float red[kElemBlock];
  const int64_t e0 = int64_t(blockIdx.x) * kElemBlock;
  const int j = __builtin_amdgcn_readfirstlane(find_job(ends, n_jobs, e0));
  BatchJobDev &J = jobs[j];
  if (!J.active) return;
  const int64_t i = e0 + threadIdx.x - J.pelem_begin;
  float out = -kInf;
  if (i < J.n_param) {
    const int t = J.iter + 1;
    const float lr_t = float(double(J.lr) * sqrt(1.0 - exp(double(t) * J.log_b2)) / (1.0 - exp(double(t) * J.log_b1)));
    const float g = J.grad[i];
    const float m = J.b1 * J.m[i] + (1.0f - J.b1) * g;
    const float v = J.b2 * J.v[i] + (1.0f - J.b2) * (g * g);
    J.m[i] = m;
    J.v[i] = v;
    const float p = J.ring[J.ring_new][i] - lr_t * m / (sqrtf(v) + J.eps);
    out = fminf(fmaxf(p, J.clamp_lo), J.clamp_hi);
    J.ring[J.ring_next][i] = out;
  }
  red[threadIdx.x] = out;
  __syncthreads();
  for (int s = kElemBlock / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(&J.next_key, float_key_dev(red[0]));
}

// ---------------------------------------------------------- 7. advance
__global__ void batch_advance_kernel(BatchJobDev *jobs, int n_jobs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n_jobs && jobs[j].active) jobs[j].iter += 1;
}

}  // namespace

// ------------------------------------------------------------------ host
struct BatchClass {
  int itype, ns;
  std::vector<int32_t> ids;
  std::vector<int64_t> ends;  // cumulative padded unit counts
  int64_t total = 0;
};

struct Batch {
  std::vector<BatchJobDev> host_jobs;
  std::vector<const Scene *> scenes;
  bool committed = false;
  void *dev = nullptr;
  BatchJobDev *d_jobs = nullptr;
  int64_t *d_elem_ends = nullptr, *d_pelem_ends = nullptr;
  int64_t elem_total = 0, pelem_total = 0;
  unsigned *d_queues = nullptr;
  BatchClass render_cls[4];  // (itype, ns): 0 path/1, 1 path/2, 2 vol/1, 3 vol/2
  BatchClass adj_cls[2];     // itype
  int32_t *d_ids[6] = {nullptr};
  int64_t *d_ends[6] = {nullptr};
  std::vector<int32_t> sg_ids;   // jobs whose optimised grid has a supergrid
  std::vector<int64_t> sg_ends;  // cumulative 2 * cells
  int32_t *d_sg_ids = nullptr;
  int64_t *d_sg_ends = nullptr;
  int block = 256, grab = 256, grid_render = 0, grid_adj = 0;
  int feat = 0;  // kFeat* union over the volume jobs' scenes
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  ~Batch() {
    for (const Scene *s : scenes) --const_cast<Scene *>(s)->batch_refs;
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (dev) (void)hipFree(dev);
  }
};

}  // namespace rcv

using namespace rcv;

struct rcv_batch : public rcv::Batch {};

extern "C" int rcv_batch_create(rcv_batch **out) {
  RCV_REQUIRE(out, "rcv_batch_create: null out");
  *out = new rcv_batch();
  return RCV_OK;
}

extern "C" int rcv_batch_destroy(rcv_batch *b) {
  delete b;
  return RCV_OK;
}

extern "C" int rcv_batch_add_job(rcv_batch *b, const rcv_batch_job *d, int32_t *out_job) {
  RCV_REQUIRE(b && d, "rcv_batch_add_job: null argument");
  if (b->committed) return fail(RCV_E_STATE, "rcv_batch_add_job: batch already committed");
  const Scene *s = static_cast<const Scene *>(d->scene);
  RCV_REQUIRE(s && s->committed, "rcv_batch_add_job: scene not committed");
  for (auto *o : b->scenes)
    RCV_REQUIRE(o != s, "rcv_batch_add_job: a scene can belong to one job only");
  RCV_REQUIRE(d->integrator.type == RCV_INTEGRATOR_PATH || d->integrator.type == RCV_INTEGRATOR_VOLPATH,
              "rcv_batch_add_job: bad integrator type");
  RCV_REQUIRE(d->method == 0 || d->method == 1, "rcv_batch_add_job: method must be 0 (baseline) or 1 (cv_ps)");
  RCV_REQUIRE(d->loss == RCV_LOSS_L1 || d->loss == RCV_LOSS_L2, "rcv_batch_add_job: bad loss");
  RCV_REQUIRE(d->sensor >= 0 && d->sensor < int(s->sensors.size()), "rcv_batch_add_job: bad sensor");
  RCV_REQUIRE(d->param >= 0 && d->param < int(s->params.size()), "rcv_batch_add_job: bad param");
  RCV_REQUIRE(d->spp > 0 && d->spp_grad > 0, "rcv_batch_add_job: spp must be positive");
  for (int k = 0; k < 3; ++k) RCV_REQUIRE(d->param_ring[k], "rcv_batch_add_job: null parameter ring buffer");
  RCV_REQUIRE(d->adam_m && d->adam_v && d->grad && d->img_ref && d->H && d->F && d->img && d->grad_img,
              "rcv_batch_add_job: null buffer");
  if (d->method == 1)
    for (int k = 0; k < 8; ++k) RCV_REQUIRE(d->cv_state[k], "rcv_batch_add_job: null CV state array");
  const SensorInfo &sen = s->sensors[d->sensor];
  const Param &P = s->params[d->param];
  BatchJobDev J{};
  J.view = s->view;
  J.view.counters = nullptr;
  J.integ = IntegArgs{d->integrator.max_depth, d->integrator.rr_depth, d->integrator.hide_emitters};
  J.itype = d->integrator.type;
  J.method = d->method;
  J.loss = d->loss;
  J.sensor = d->sensor;
  J.film_w = sen.film_w;
  J.n_pix = sen.film_w * sen.film_h;
  RCV_REQUIRE(int64_t(J.n_pix) * std::max(d->spp, d->spp_grad) < (int64_t(1) << 32),
              "rcv_batch_add_job: pixels * spp must stay below 2^32");
  J.spp = d->spp;
  J.spp_g = d->spp_grad;
  // finer work units than a single launch: a batch has parallelism to
  // spare, and short units keep the end of a persistent launch balanced
  if (int rc = plan_launch(&d->integrator, J.n_pix, J.spp, J.spt, J.cl2, kBatchMaxSpt, kBatchMinThreads)) return rc;
  if (int rc = plan_launch(&d->integrator, J.n_pix, J.spp_g, J.spt_g, J.cl2_g, kBatchMaxSpt, kBatchMinThreads))
    return rc;
  J.seed = d->seed;
  J.param = d->param;
  J.n_param = int64_t(P.nx) * P.ny * P.nz * P.ch;
  J.n_img = int64_t(J.n_pix) * 3;
  for (int k = 0; k < 3; ++k) J.ring[k] = d->param_ring[k];
  J.ring_old = d->resume ? d->ring_old : 0;
  J.ring_new = d->resume ? d->ring_new : 1;
  J.ring_next = d->resume ? d->ring_next : 2;
  RCV_REQUIRE(J.ring_old != J.ring_new && J.ring_new != J.ring_next && J.ring_old != J.ring_next &&
                  J.ring_old >= 0 && J.ring_old < 3 && J.ring_new >= 0 && J.ring_new < 3 && J.ring_next >= 0 &&
                  J.ring_next < 3,
              "rcv_batch_add_job: ring slots must be a permutation of 0, 1, 2");
  J.view.params[d->param].data[RCV_STATE_OLD] = J.ring[J.ring_old];
  J.view.params[d->param].data[RCV_STATE_NEW] = J.ring[J.ring_new];
  J.maj_keys = s->majorant_keys + 2 * d->param;
  J.iter = d->resume ? d->iteration : 0;
  J.active = 1;
  J.next_key = d->resume ? d->next_key : 0u;
  J.H = d->H;
  J.F = d->F;
  J.img = d->img;
  J.img_ref = d->img_ref;
  J.grad_img = d->grad_img;
  for (int k = 0; k < 8; ++k) J.cv[k] = d->cv_state[k];
  J.warmup = d->warmup;
  J.cv_b1 = d->cv_beta1;
  J.cv_b2 = d->cv_beta2;
  J.cv_log_b1 = log(double(d->cv_beta1));
  J.cv_log_b2 = log(double(d->cv_beta2));
  J.pre_update = d->pre_update;
  J.grad = d->grad;
  J.gt.clear();
  J.gt.g[d->param] = d->grad;
  J.m = d->adam_m;
  J.v = d->adam_v;
  J.lr = d->lr;
  J.b1 = d->beta1;
  J.b2 = d->beta2;
  J.eps = d->eps;
  J.clamp_lo = d->clamp_lo;
  J.clamp_hi = d->clamp_hi;
  J.log_b1 = log(double(d->beta1));
  J.log_b2 = log(double(d->beta2));
  J.sg_out = nullptr;
  int n_sg = 0;
  for (const MediumInfo &m : s->media)
    if (m.sigma_param == d->param && m.supergrid_offset >= 0) {
      ++n_sg;
      J.sg_out = s->supergrids + m.supergrid_offset;
      for (int k = 0; k < 3; ++k) J.sg_res[k] = m.sg_res[k];
    }
  RCV_REQUIRE(n_sg <= 1, "rcv_batch_add_job: the optimised grid drives %d media with supergrids (at most 1)", n_sg);
  RCV_REQUIRE(!(s->features() & kFeatHomog), "rcv_batch_add_job: batched jobs take heterogeneous media only");
  if (J.itype == RCV_INTEGRATOR_VOLPATH) b->feat |= s->features();
  b->host_jobs.push_back(J);
  b->scenes.push_back(s);
  ++const_cast<Scene *>(s)->batch_refs;
  if (out_job) *out_job = int(b->host_jobs.size()) - 1;
  return RCV_OK;
}

static int64_t pad(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

extern "C" int rcv_batch_commit(rcv_batch *b, void *stream) {
  RCV_REQUIRE(b, "rcv_batch_commit: null batch");
  if (b->committed) return fail(RCV_E_STATE, "rcv_batch_commit: already committed");
  const int n = int(b->host_jobs.size());
  RCV_REQUIRE(n > 0 && n <= RCV_BATCH_MAX_JOBS, "rcv_batch_commit: 1..%d jobs", RCV_BATCH_MAX_JOBS);
  std::vector<int64_t> elem_ends(n), pelem_ends(n);
  int64_t e = 0, pe = 0;
  for (int j = 0; j < n; ++j) {
    BatchJobDev &J = b->host_jobs[j];
    J.elem_begin = e;
    e += pad(J.n_img, kElemBlock);
    elem_ends[j] = e;
    J.pelem_begin = pe;
    pe += pad(J.n_param, kElemBlock);
    pelem_ends[j] = pe;
    const int rc_i = (J.itype == RCV_INTEGRATOR_VOLPATH ? 2 : 0) + (J.method == 1 ? 1 : 0);
    BatchClass &rc = b->render_cls[rc_i];
    J.unit_begin = rc.total;
    rc.total += pad(int64_t(J.n_pix) << J.cl2, kWave);
    rc.ids.push_back(j);
    rc.ends.push_back(rc.total);
    BatchClass &ac = b->adj_cls[J.itype == RCV_INTEGRATOR_VOLPATH ? 1 : 0];
    J.unit_begin_g = ac.total;
    ac.total += pad(int64_t(J.n_pix) << J.cl2_g, kWave);
    ac.ids.push_back(j);
    ac.ends.push_back(ac.total);
  }
  for (auto &c : b->render_cls)
    RCV_REQUIRE(c.total < (int64_t(1) << 31), "rcv_batch_commit: too many work units");
  for (auto &c : b->adj_cls)
    RCV_REQUIRE(c.total < (int64_t(1) << 31), "rcv_batch_commit: too many work units");
  b->elem_total = e;
  b->pelem_total = pe;
  b->sg_ids.clear();
  b->sg_ends.clear();
  int64_t sg_total = 0;
  for (int j = 0; j < n; ++j) {
    const BatchJobDev &J = b->host_jobs[j];
    if (!J.sg_out) continue;
    sg_total += 2 * int64_t(J.sg_res[0]) * J.sg_res[1] * J.sg_res[2];
    b->sg_ids.push_back(j);
    b->sg_ends.push_back(sg_total);
  }
  // device block: jobs | elem ends | pelem ends | queues | class ids/ends
  size_t bytes = 0;
  auto reserve = [&](size_t sz) {
    size_t off = (bytes + 255) & ~size_t(255);
    bytes = off + std::max<size_t>(sz, 16);
    return off;
  };
  const size_t o_jobs = reserve(sizeof(BatchJobDev) * n);
  const size_t o_ee = reserve(8 * n);
  const size_t o_pe = reserve(8 * n);
  const size_t o_q = reserve(64);
  const size_t o_sgi = reserve(4 * b->sg_ids.size());
  const size_t o_sge = reserve(8 * b->sg_ends.size());
  size_t o_ids[6], o_ends[6];
  for (int c = 0; c < 6; ++c) {
    const BatchClass &bc = c < 4 ? b->render_cls[c] : b->adj_cls[c - 4];
    o_ids[c] = reserve(4 * bc.ids.size());
    o_ends[c] = reserve(8 * bc.ends.size());
  }
  RCV_TRY_HIP(hipMalloc(&b->dev, bytes), "rcv_batch_commit: hipMalloc");
  char *base = static_cast<char *>(b->dev);
  std::vector<char> blob(bytes, 0);
  std::memcpy(blob.data() + o_jobs, b->host_jobs.data(), sizeof(BatchJobDev) * n);
  std::memcpy(blob.data() + o_ee, elem_ends.data(), 8 * n);
  std::memcpy(blob.data() + o_pe, pelem_ends.data(), 8 * n);
  if (!b->sg_ids.empty()) {
    std::memcpy(blob.data() + o_sgi, b->sg_ids.data(), 4 * b->sg_ids.size());
    std::memcpy(blob.data() + o_sge, b->sg_ends.data(), 8 * b->sg_ends.size());
  }
  b->d_sg_ids = reinterpret_cast<int32_t *>(base + o_sgi);
  b->d_sg_ends = reinterpret_cast<int64_t *>(base + o_sge);
  for (int c = 0; c < 6; ++c) {
    const BatchClass &bc = c < 4 ? b->render_cls[c] : b->adj_cls[c - 4];
    if (!bc.ids.empty()) {
      std::memcpy(blob.data() + o_ids[c], bc.ids.data(), 4 * bc.ids.size());
      std::memcpy(blob.data() + o_ends[c], bc.ends.data(), 8 * bc.ends.size());
    }
    b->d_ids[c] = reinterpret_cast<int32_t *>(base + o_ids[c]);
    b->d_ends[c] = reinterpret_cast<int64_t *>(base + o_ends[c]);
  }
  hipStream_t st = as_stream(stream);
  RCV_TRY_HIP(hipMemcpyAsync(b->dev, blob.data(), bytes, hipMemcpyHostToDevice, st), "rcv_batch_commit: upload");
  RCV_TRY_HIP(hipStreamSynchronize(st), "rcv_batch_commit: synchronize");
  b->d_jobs = reinterpret_cast<BatchJobDev *>(base + o_jobs);
  b->d_elem_ends = reinterpret_cast<int64_t *>(base + o_ee);
  b->d_pelem_ends = reinterpret_cast<int64_t *>(base + o_pe);
  b->d_queues = reinterpret_cast<unsigned *>(base + o_q);
  b->committed = true;
  return rcv_batch_set_launch(b, 0, 0);
}

template <typename K>
static int persistent_grid(K kernel, int block, size_t lds, int &grid) {
  int per_cu = 0;
  RCV_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds), "occupancy query");
  int dev = 0, cus = 0;
  RCV_TRY_HIP(hipGetDevice(&dev), "hipGetDevice");
  RCV_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
  grid = std::max(1, per_cu) * std::max(1, cus);
  return RCV_OK;
}

extern "C" int rcv_batch_set_launch(rcv_batch *b, int32_t block_size, int32_t queue_grab) {
  RCV_REQUIRE(b && b->committed, "rcv_batch_set_launch: batch not committed");
  if (block_size == 0) block_size = 256;
  if (queue_grab == 0) queue_grab = 64;
  RCV_REQUIRE(block_size >= 64 && block_size <= 512 && block_size % 64 == 0,
              "rcv_batch_set_launch: block size must be a multiple of 64 in [64, 512]");
  RCV_REQUIRE(queue_grab >= 64 && queue_grab <= 8192 && queue_grab % 64 == 0,
              "rcv_batch_set_launch: queue grab must be a multiple of 64 in [64, 8192]");
  const size_t lds = size_t(kLdsWordsPerThread) * block_size * sizeof(int);
  int dev = 0, max_lds = 0;
  RCV_TRY_HIP(hipGetDevice(&dev), "hipGetDevice");
  RCV_TRY_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "LDS per block");
  RCV_REQUIRE(lds <= size_t(max_lds), "rcv_batch_set_launch: block size %d needs %zu B of LDS per block, above the %d B "
              "this device allows", block_size, lds, max_lds);
  b->block = block_size;
  b->grab = queue_grab;
  int g1 = 0, g2 = 0;
  int rc = RCV_OK;
  with_features(b->feat, [&](auto f) {
    rc = persistent_grid(batch_render_kernel<2, RCV_INTEGRATOR_VOLPATH, decltype(f)::value>, block_size, lds, g1);
    if (rc == RCV_OK)
      rc = persistent_grid(batch_adjoint_kernel<RCV_INTEGRATOR_VOLPATH, decltype(f)::value>, block_size, lds, g2);
  });
  if (rc != RCV_OK) return rc;
  b->grid_render = g1;
  b->grid_adj = g2;
  if (b->graph_exec) {
    (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return RCV_OK;
}

static int enqueue_step(rcv_batch *b, hipStream_t st) {
  const int n = int(b->host_jobs.size());
  const size_t lds = size_t(kLdsWordsPerThread) * b->block * sizeof(int);
  hipLaunchKernelGGL(batch_rotate_kernel, dim3(ceil_div(std::max(n, 6), 64)), dim3(64), 0, st, b->d_jobs, n,
                     b->d_queues);
  if (!b->sg_ids.empty()) {
    const int64_t total = b->sg_ends.back();
    hipLaunchKernelGGL(batch_supergrid_kernel, dim3(ceil_div(total, 128)), dim3(128), 0, st, b->d_jobs,
                       b->d_sg_ids, b->d_sg_ends, int(b->sg_ids.size()), total);
  }
  // render classes: (path, 1), (path, 2), (vol, 1), (vol, 2); queues 0..3 reused per class in sequence
  for (int c = 0; c < 4; ++c) {
    const BatchClass &bc = b->render_cls[c];
    if (bc.total == 0) continue;
    unsigned *q = b->d_queues + c;
    const int nc = int(bc.ids.size());
    const dim3 grid(b->grid_render), blk(b->block);
    if (c == 0)
      hipLaunchKernelGGL((batch_render_kernel<1, RCV_INTEGRATOR_PATH, 0>), grid, blk, lds, st, b->d_jobs,
                         b->d_ids[c], b->d_ends[c], nc, bc.total, q, b->grab);
    else if (c == 1)
      hipLaunchKernelGGL((batch_render_kernel<2, RCV_INTEGRATOR_PATH, 0>), grid, blk, lds, st, b->d_jobs,
                         b->d_ids[c], b->d_ends[c], nc, bc.total, q, b->grab);
    else if (c == 2)
      with_features(b->feat, [&](auto f) {
        hipLaunchKernelGGL((batch_render_kernel<1, RCV_INTEGRATOR_VOLPATH, decltype(f)::value>), grid, blk, lds, st,
                           b->d_jobs, b->d_ids[c], b->d_ends[c], nc, bc.total, q, b->grab);
      });
    else
      with_features(b->feat, [&](auto f) {
        hipLaunchKernelGGL((batch_render_kernel<2, RCV_INTEGRATOR_VOLPATH, decltype(f)::value>), grid, blk, lds, st,
                           b->d_jobs, b->d_ids[c], b->d_ends[c], nc, bc.total, q, b->grab);
      });
    if (int rc = check_launch("rcv_batch_step: render")) return rc;
  }
  hipLaunchKernelGGL(batch_combine_loss_kernel, dim3(unsigned(b->elem_total / kElemBlock)), dim3(kElemBlock), 0,
                     st, b->d_jobs, b->d_elem_ends, n);
  hipLaunchKernelGGL(batch_zero_grad_kernel, dim3(unsigned(b->pelem_total / kElemBlock)), dim3(kElemBlock), 0, st,
                     b->d_jobs, b->d_pelem_ends, n);
  // adjoint classes use queues 4 and 5 (also reset by the rotate kernel)
  for (int c = 0; c < 2; ++c) {
    const BatchClass &bc = b->adj_cls[c];
    if (bc.total == 0) continue;
    unsigned *q = b->d_queues + 4 + c;
    const int nc = int(bc.ids.size());
    if (c == 0)
      hipLaunchKernelGGL((batch_adjoint_kernel<RCV_INTEGRATOR_PATH, 0>), dim3(b->grid_adj), dim3(b->block), lds, st,
                         b->d_jobs, b->d_ids[4 + c], b->d_ends[4 + c], nc, bc.total, q, b->grab);
    else
      with_features(b->feat, [&](auto f) {
        hipLaunchKernelGGL((batch_adjoint_kernel<RCV_INTEGRATOR_VOLPATH, decltype(f)::value>), dim3(b->grid_adj),
                           dim3(b->block), lds, st, b->d_jobs, b->d_ids[4 + c], b->d_ends[4 + c], nc, bc.total, q,
                           b->grab);
      });
    if (int rc = check_launch("rcv_batch_step: adjoint")) return rc;
  }
  hipLaunchKernelGGL(batch_adam_kernel, dim3(unsigned(b->pelem_total / kElemBlock)), dim3(kElemBlock), 0, st,
                     b->d_jobs, b->d_pelem_ends, n);
  hipLaunchKernelGGL(batch_advance_kernel, dim3(ceil_div(n, 64)), dim3(64), 0, st, b->d_jobs, n);
  return check_launch("rcv_batch_step");
}

extern "C" int rcv_batch_step(rcv_batch *b, int32_t use_graph, void *stream) {
  RCV_REQUIRE(b && b->committed, "rcv_batch_step: batch not committed");
  hipStream_t st = as_stream(stream);
  if (!use_graph) return enqueue_step(b, st);
  if (!b->graph_exec) {
    RCV_TRY_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "rcv_batch_step: begin capture");
    int rc = enqueue_step(b, st);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != RCV_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    RCV_TRY_HIP(e, "rcv_batch_step: end capture");
    b->graph = g;
    RCV_TRY_HIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0), "rcv_batch_step: instantiate");
  }
  RCV_TRY_HIP(hipGraphLaunch(b->graph_exec, st), "rcv_batch_step: graph launch");
  return RCV_OK;
}

namespace rcv {
namespace {
struct ActiveBits {
  uint32_t w[RCV_BATCH_MAX_JOBS / 32];
};
__global__ void batch_set_active_kernel(BatchJobDev *jobs, int n_jobs, ActiveBits bits) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n_jobs) jobs[j].active = (bits.w[j >> 5] >> (j & 31)) & 1u;
}
}  // namespace
}  // namespace rcv

extern "C" int rcv_batch_set_active(rcv_batch *b, const uint32_t *active_bits, int32_t n_jobs, void *stream) {
  RCV_REQUIRE(b && b->committed, "rcv_batch_set_active: batch not committed");
  RCV_REQUIRE(active_bits, "rcv_batch_set_active: null flags");
  RCV_REQUIRE(n_jobs == int(b->host_jobs.size()), "rcv_batch_set_active: expected %d jobs, got %d",
              int(b->host_jobs.size()), n_jobs);
  ActiveBits bits{};
  for (int w = 0; w < (n_jobs + 31) / 32; ++w) bits.w[w] = active_bits[w];
  hipLaunchKernelGGL(batch_set_active_kernel, dim3(ceil_div(n_jobs, 64)), dim3(64), 0, as_stream(stream),
                     b->d_jobs, n_jobs, bits);
  return check_launch("rcv_batch_set_active");
}

extern "C" int rcv_batch_read_job(const rcv_batch *b, int32_t job, int32_t *iteration, int32_t *ring_old,
                                  int32_t *ring_new, int32_t *ring_next, uint32_t *next_key, void *stream) {
  RCV_REQUIRE(b && b->committed, "rcv_batch_read_job: batch not committed");
  RCV_REQUIRE(job >= 0 && job < int(b->host_jobs.size()), "rcv_batch_read_job: bad job");
  BatchJobDev J;
  hipStream_t st = as_stream(stream);
  RCV_TRY_HIP(hipMemcpyAsync(&J, b->d_jobs + job, sizeof(J), hipMemcpyDeviceToHost, st), "rcv_batch_read_job");
  RCV_TRY_HIP(hipStreamSynchronize(st), "rcv_batch_read_job: synchronize");
  if (iteration) *iteration = J.iter;
  if (ring_old) *ring_old = J.ring_old;
  if (ring_new) *ring_new = J.ring_new;
  if (ring_next) *ring_next = J.ring_next;
  if (next_key) *next_key = J.next_key;
  return RCV_OK;
}

// Current parameter state (NEW) pointer and iteration of a job (host read,
// synchronises the stream).
extern "C" int rcv_batch_job_state(const rcv_batch *b, int32_t job, int32_t *iteration, int32_t *ring_new,
                                   void *stream) {
  RCV_REQUIRE(b && b->committed, "rcv_batch_job_state: batch not committed");
  RCV_REQUIRE(job >= 0 && job < int(b->host_jobs.size()), "rcv_batch_job_state: bad job");
  BatchJobDev J;
  hipStream_t st = as_stream(stream);
  RCV_TRY_HIP(hipMemcpyAsync(&J, b->d_jobs + job, sizeof(J), hipMemcpyDeviceToHost, st), "rcv_batch_job_state");
  RCV_TRY_HIP(hipStreamSynchronize(st), "rcv_batch_job_state: synchronize");
  if (iteration) *iteration = J.iter;
  // after step k the optimizer's latest result is ring_next (rotated in at the next step)
  if (ring_new) *ring_new = J.iter > 0 ? J.ring_next : J.ring_new;
  return RCV_OK;
}
