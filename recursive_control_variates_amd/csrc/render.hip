// Render entry points of librcv: plain, two-state and adjoint launches of the
// surface (path.hpp) and volume (volpath.hpp) integrators.
//
// Launch shape: the spp camera samples of a pixel are split over 2^chunk_log2
// threads (chunk c traces samples [c spp >> chunk_log2, (c + 1) spp >>
// chunk_log2) in sequence); the threads of a pixel sit in the same wavefront
// and their sums are combined with cross-lane shuffles, so the film needs no
// atomics and the pixel average (box filter develop, plugins/
// twostatepath.py:168-177) is written once, coalesced. The random stream of a
// camera sample depends only on (seed, pixel * spp + sample), so the launch
// shape changes results only through the float summation order of a pixel.
// The BVH traversal stack of each thread lives in LDS (as deep as the
// scene's BVH, strided by the block size), followed by per-lane slots (the
// adjoint's NEE weights): kLdsWordsPerThread words per thread in all.

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"
#include "path.hpp"
#include "render_internal.hpp"
#include "volpath.hpp"
#include "volsimple.hpp"

namespace rcv {
namespace {

constexpr int kMaxBlock = 256;      // launch bound of the render kernels
constexpr int kDefaultBlock = 256;  // threads per block when the descriptor says 0
constexpr int kAdjUnitSamples = 16;  // camera samples per work unit of the persistent per-sample adjoints
// Camera samples per work unit of the volume path tracer's persistent
// launches (the adjoint; the renders, within kRenderMaxUnitsPerPixel). With
// small units the lanes of a wavefront trace samples of the same few pixels,
// whose paths start alike and share L1 lines of the grid, and the last units
// of a launch are short. The 64^3 adjoint at the mid / final state of the
// optimisation: 46.1 / 67.5 ms with 16-sample units, 39.7 / 51.4 with 8,
// 37.0 / 47.8 with 4, 36.2 / 46.4 with 2.
constexpr int kVolUnitSamples = 2;

template <int NS, int INTEG, int F>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(kWavesPerEuRender<NS, INTEG>))) void render_kernel(
    SceneView S, IntegArgs I, int sensor, uint32_t seed,
                                                        int spp, int spt, int chunk_log2, float *img_old,
                                                        float *img_new) {
  extern __shared__ int lds_stack[];  // kLdsWordsPerThread words per thread
  const LdsStack stack = lds_of(S, lds_stack);
  const SensorInfo &sen = S.sensors[sensor];
  const int film_w = sen.film_w;
  const int64_t n_pix = int64_t(film_w) * sen.film_h;
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t pixel = tid >> chunk_log2;
  const int chunk = int(tid & ((1 << chunk_log2) - 1));
  V3 acc[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) acc[k] = v3(0.f);
  LocalCounters cnt{S.counters};
  if (pixel < n_pix) {
    const int px = int(pixel % film_w), py = int(pixel / film_w);
    const int s_end = sample_end(chunk, spp, chunk_log2);
    if (INTEG == RCV_INTEGRATOR_VOLPATH)
      volpath_render_samples<NS, F>(S, I, sen, seed, px, py, pixel, spp, sample_end(chunk - 1, spp, chunk_log2),
                                    s_end, acc, stack, cnt);
    else for (int s = sample_end(chunk - 1, spp, chunk_log2); s < s_end; ++s) {
      Pcg32 rng = lane_rng(seed, uint32_t(pixel * spp + s));
      const Ray ray = camera_ray(sen, px, py, rng);
      V3 L[NS];
      if (INTEG == RCV_INTEGRATOR_VOLPATH_SIMPLE)
        L[0] = simple_sample<false, F>(S, I, 0, rng, ray, simple_camera_start(), stack, cnt, nullptr, nullptr,
                                       nullptr);
      else
        path_sample<NS, false>(S, I, rng, ray, L, stack, cnt, nullptr, nullptr);
#pragma unroll
      for (int k = 0; k < NS; ++k) acc[k] += L[k];
    }
  }
  for (int off = 1; off < (1 << chunk_log2); off <<= 1) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      acc[k].x += __shfl_xor(acc[k].x, off, 64);
      acc[k].y += __shfl_xor(acc[k].y, off, 64);
      acc[k].z += __shfl_xor(acc[k].z, off, 64);
    }
  }
  if (pixel < n_pix && chunk == 0) {
    const float inv = float(spp);
    float *o0 = (NS == 2 ? img_old : img_new) + 3 * pixel;
    o0[0] = acc[0].x / inv;
    o0[1] = acc[0].y / inv;
    o0[2] = acc[0].z / inv;
    if (NS == 2) {
      float *o1 = img_new + 3 * pixel;
      o1[0] = acc[NS - 1].x / inv;
      o1[1] = acc[NS - 1].y / inv;
      o1[2] = acc[NS - 1].z / inv;
    }
  }
}

// The adjoint of camera sample s of a pixel for the per-sample integrators
// (surface path tracer, volpathsimple): the primal pass, then the adjoint pass
// with the same random numbers -- a replay that scatters d(log weight)/d(theta)
// times the radiance still to come (path tracer, skipped when L = 0), or
// volpathsimple's adjoint with its gradient-only stream as well (DRT
// contributes even where L = 0).
template <int INTEG, int F>
__device__ inline void adjoint_one_sample(const SceneView &S, const IntegArgs &I, const SensorInfo &sen,
                                          uint32_t seed, int px, int py, int64_t pixel, int spp, int s, V3 dL,
                                          const GradTargets &G, const LdsStack &stack, LocalCounters &cnt) {
  Pcg32 rng = lane_rng(seed, uint32_t(pixel * spp + s));
  const Ray ray = camera_ray(sen, px, py, rng);
  const Pcg32 rng0 = rng;
  V3 L[1];
  if constexpr (INTEG == RCV_INTEGRATOR_VOLPATH_SIMPLE) {
    L[0] = simple_sample<false, F>(S, I, 0, rng, ray, simple_camera_start(), stack, cnt, nullptr, nullptr, nullptr);
    AdjState adj{dL, L[0]};
    rng = rng0;
    Pcg32 alt = lane_rng(simple_alt_seed(seed), uint32_t(pixel * spp + s));
    (void)simple_sample<true, F>(S, I, 0, rng, ray, simple_camera_start(), stack, cnt, &adj, &alt, &G);
  } else {
    path_sample<1, false>(S, I, rng, ray, L, stack, cnt, nullptr, nullptr);
    if (!any_nonzero(L[0])) return;
    AdjState adj{dL, L[0]};
    rng = rng0;
    V3 L2[1];
    path_sample<1, true>(S, I, rng, ray, L2, stack, cnt, &adj, &G);
  }
}

// Path-replay adjoint (NEW state): primal pass for L, then a replay with the
// same random numbers that scatters d(log weight)/d(theta) times the radiance
// still to come into the parameter gradients. dimage/dL = 1/spp (box filter).
template <int INTEG, int F>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<INTEG>))) void adjoint_kernel(
    SceneView S, IntegArgs I, int sensor, uint32_t seed,
                                                         int spp, int spt, int chunk_log2,
                                                         const float *__restrict__ grad_img, GradTargets G) {
  extern __shared__ int lds_stack[];  // kLdsWordsPerThread words per thread
  const LdsStack stack = lds_of(S, lds_stack);
  const SensorInfo &sen = S.sensors[sensor];
  const int film_w = sen.film_w;
  const int64_t n_pix = int64_t(film_w) * sen.film_h;
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t pixel = tid >> chunk_log2;
  const int chunk = int(tid & ((1 << chunk_log2) - 1));
  LocalCounters cnt{S.counters};
  HomSum hom{0.0f, v3(0.f)};
  if (pixel < n_pix) {
    const int px = int(pixel % film_w), py = int(pixel / film_w);
    const float inv = float(spp);
    const V3 dL = v3(grad_img[3 * pixel] / inv, grad_img[3 * pixel + 1] / inv, grad_img[3 * pixel + 2] / inv);
    if (any_nonzero(dL) && INTEG == RCV_INTEGRATOR_VOLPATH) {
      volpath_adjoint_samples<F>(S, I, sen, seed, px, py, pixel, spp, sample_end(chunk - 1, spp, chunk_log2),
                                 sample_end(chunk, spp, chunk_log2), dL, &G, stack, cnt, &hom);
    } else if (any_nonzero(dL)) {
      const int s_end = sample_end(chunk, spp, chunk_log2);
      for (int s = sample_end(chunk - 1, spp, chunk_log2); s < s_end; ++s)
        adjoint_one_sample<INTEG, F>(S, I, sen, seed, px, py, pixel, spp, s, dL, G, stack, cnt);
    }
  }
  if constexpr (INTEG == RCV_INTEGRATOR_VOLPATH && (F & kFeatHomog) != 0) flush_homogeneous(S, G, hom);
}

// Forward mode (NEW state): per pixel, d(image)/d(theta) . tangent, from the
// primal pass of each sample and its replay with the same random numbers,
// which gathers the tangents (in T) where the adjoint scatters (the
// transpose of adjoint_kernel on the same samples). Launch shape and pixel
// reduction as render_kernel.
template <int INTEG, int F>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<INTEG>))) void forward_kernel(
    SceneView S, IntegArgs I, int sensor, uint32_t seed, int spp, int spt, int chunk_log2, GradTargets T,
    float *__restrict__ dimg) {
  extern __shared__ int lds_stack[];  // kLdsWordsPerThread words per thread
  const LdsStack stack = lds_of(S, lds_stack);
  const SensorInfo &sen = S.sensors[sensor];
  const int film_w = sen.film_w;
  const int64_t n_pix = int64_t(film_w) * sen.film_h;
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t pixel = tid >> chunk_log2;
  const int chunk = int(tid & ((1 << chunk_log2) - 1));
  LocalCounters cnt{S.counters};
  V3 acc = v3(0.f);
  if (pixel < n_pix) {
    const int px = int(pixel % film_w), py = int(pixel / film_w);
    const int s0 = sample_end(chunk - 1, spp, chunk_log2), s1 = sample_end(chunk, spp, chunk_log2);
    if constexpr (INTEG == RCV_INTEGRATOR_VOLPATH) {
      volpath_forward_samples<F>(S, I, sen, seed, px, py, pixel, spp, s0, s1, &T, stack, cnt, acc);
    } else {
      for (int s = s0; s < s1; ++s) {
        Pcg32 rng = lane_rng(seed, uint32_t(pixel * spp + s));
        const Ray ray = camera_ray(sen, px, py, rng);
        const Pcg32 rng0 = rng;
        V3 L[1];
        path_sample<1, false>(S, I, rng, ray, L, stack, cnt, nullptr, nullptr);
        if (!any_nonzero(L[0])) continue;
        AdjState adj{v3(0.f), L[0], v3(0.f)};
        rng = rng0;
        V3 L2[1];
        path_sample<1, true, true>(S, I, rng, ray, L2, stack, cnt, &adj, &T);
        acc += adj.dI;
      }
    }
  }
  for (int off = 1; off < (1 << chunk_log2); off <<= 1) {
    acc.x += __shfl_xor(acc.x, off, 64);
    acc.y += __shfl_xor(acc.y, off, 64);
    acc.z += __shfl_xor(acc.z, off, 64);
  }
  if (pixel < n_pix && chunk == 0) {
    const float inv = float(spp);
    float *o = dimg + 3 * pixel;
    o[0] = acc.x / inv;
    o[1] = acc.y / inv;
    o[2] = acc.z / inv;
  }
}

// Path-replay adjoint of the volume path tracer as a persistent launch: one
// grid of resident waves whose lanes pull work units (a pixel's chunk of
// kVolUnitSamples camera samples) from a device counter, a wavefront's
// requests combined into one atomic, and keep tracing until the units run
// out, so no wave waits at the end of the launch for a few long pixels. The
// result per sample is that of adjoint_kernel; only the order of the gradient
// scatter-adds differs.
template <int F, int W>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(W))) void
adjoint_persistent_kernel(SceneView S, IntegArgs I, int sensor, uint32_t seed, int spp, int chunk_log2,
                          const float *__restrict__ grad_img, GradTargets G, unsigned *queue, unsigned total_units) {
  extern __shared__ int lds_stack[];  // kLdsWordsPerThread words per thread
  const LdsStack stack = lds_of(S, lds_stack);
  const SensorInfo &sen = S.sensors[sensor];
  const int film_w = sen.film_w;
  LocalCounters cnt{S.counters};
  VolPath<1, true, F> t;
  t.L[0] = v3(0.f);
  t.init_hom();
  log_begin(G, stack);
  bool live = false, want_replay = false;
  int s = 0, s_end = 0, px = 0, py = 0;
  int64_t pixel = 0;
  V3 dL = v3(0.f);
  const float inv = float(spp);
  while (true) {
    if (!live) {
      if (!want_replay && s >= s_end) {
        // the lanes that need a unit in this trip share one atomic
        const uint64_t m = __ballot(1);
        const int leader = __ffsll((unsigned long long)m) - 1;
        unsigned base = 0;
        if (int(__lane_id()) == leader) base = atomicAdd(queue, unsigned(__popcll(m)));
        base = __shfl(base, leader, 64);
        const unsigned u = base + unsigned(__popcll(m & __lanemask_lt()));
        if (u >= total_units) break;
        pixel = int64_t(u >> chunk_log2);
        const int chunk = int(u & ((1u << chunk_log2) - 1u));
        px = int(pixel % film_w);
        py = int(pixel / film_w);
        s = sample_end(chunk - 1, spp, chunk_log2);
        s_end = sample_end(chunk, spp, chunk_log2);
        dL = v3(grad_img[3 * pixel] / inv, grad_img[3 * pixel + 1] / inv, grad_img[3 * pixel + 2] / inv);
        if (!any_nonzero(dL)) {
          s = s_end;
          continue;
        }
      }
      Pcg32 rng = lane_rng(seed, uint32_t(pixel * spp + s));
      const Ray r = camera_ray(sen, px, py, rng);
      t.begin_adjoint(rng, r, cnt, want_replay, dL, t.L[0]);
    }
    live = t.bounce(S, I, stack, cnt, &G);
    if (!live) {
      want_replay = !t.end_adjoint(S, cnt, &G);
      if (!want_replay) ++s;
    }
  }
  if constexpr ((F & kFeatHomog) != 0) flush_homogeneous(S, G, t.hom);
  // the whole wavefront is past its loop
  log_end(G, stack);
}

// Scatter log reduction. Block (x, y) takes tile y of the gradient
// (voxels [y tile, (y + 1) tile)) and its share of the chunks of the three
// bins whose records reach into it (tile y, the seams below and above it);
// a record is read once, or twice when it lies across a seam. A block of
// kReduceBlock threads takes kStepChunks chunks per step: thread i reads
// slot i % kLogChunk of kReduceUnroll of them, all loads in flight together
// (unused slots are read, every index stays inside the pool, and skipped).
// The records of one chunk come from one wavefront of the adjoint, whose
// lanes trace neighbouring pixels; the 8-lane groups of a wavefront take
// their chunks in rotated order so that one instruction adds records of
// several chunks.
constexpr int kReduceBlock = 1024;
constexpr int kReduceUnroll = 8;
constexpr unsigned kChunkGroups = unsigned(kReduceBlock) / kLogChunk;
constexpr unsigned kStepChunks = kChunkGroups * kReduceUnroll;
static_assert(kReduceBlock % int(kLogChunk) == 0, "a reduction block covers whole chunks");

// Calls fn(record) for each record of the block's share.
template <class Fn>
__device__ inline void log_block_records(const int4 *__restrict__ rec, const unsigned *__restrict__ ctl,
                                         const unsigned *__restrict__ list,
                                         const unsigned *__restrict__ fill, unsigned nchunks, Fn &&fn) {
  const int y = int(blockIdx.y);
  const unsigned slot = threadIdx.x % kLogChunk;
  const unsigned sub = threadIdx.x / kLogChunk;
  const unsigned grp = (threadIdx.x >> 3) % kReduceUnroll;
  for (int k = 0; k < 3; ++k) {
    if (k == 1 && y == 0) continue;
    const int b = k == 0 ? y : kLogMaxTiles + y - (k == 1 ? 1 : 0);
    const unsigned nb = min(ctl[1 + b], nchunks);
    const unsigned *bl = list + size_t(b) * nchunks;
    for (unsigned p0 = blockIdx.x * kStepChunks; p0 < nb; p0 += gridDim.x * kStepChunks) {
      unsigned c[kReduceUnroll], n[kReduceUnroll];
      int4 e[kReduceUnroll];
#pragma unroll
      for (int u = 0; u < kReduceUnroll; ++u) {
        const unsigned j = p0 + sub + kChunkGroups * ((grp + unsigned(u)) % kReduceUnroll);
        c[u] = j < nb ? bl[j] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kReduceUnroll; ++u) {
        const unsigned j = p0 + sub + kChunkGroups * ((grp + unsigned(u)) % kReduceUnroll);
        n[u] = j < nb ? min(fill[c[u]], kLogChunk) : 0u;
      }
#pragma unroll
      for (int u = 0; u < kReduceUnroll; ++u) e[u] = rec[size_t(c[u]) * kLogChunk + slot];
#pragma unroll
      for (int u = 0; u < kReduceUnroll; ++u)
        if (slot < n[u]) fn(e[u]);
    }
  }
}

// Calls add(voxel - lo, value x weight) for the corners of a record
// (log_record) that fall in [lo, hi): the products of scatter_grid, with
// the fractions as stored.
template <class Add>
__device__ inline void record_corners(const int4 &e, int nx, int nxy, int64_t lo, int64_t hi, Add &&add) {
  const unsigned cell = unsigned(e.x);
  const unsigned long long q = (unsigned long long)unsigned(e.y) | ((unsigned long long)unsigned(e.z) << 32);
  constexpr unsigned kMask = (1u << 21) - 1u;
  GridFoot f;
  f.base = int(cell & kRecBaseMask);
  f.dx = int((cell >> 26) & 1u);
  f.dy = ((cell >> 27) & 1u) ? nx : 0;
  f.dz = ((cell >> 28) & 1u) ? nxy : 0;
  f.tx = float(unsigned(q) & kMask) / kLogFracScale;
  f.ty = float(unsigned(q >> 21) & kMask) / kLogFracScale;
  f.tz = float(unsigned(q >> 42) & kMask) / kLogFracScale;
  const float x = __int_as_float(e.w);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int64_t idx = f.idx(q);
    if (idx >= lo && idx < hi) add(idx - lo, x * f.w(q));
  }
}

// Adds the scatter log of a persistent adjoint launch to the gradient of
// an nx x ny x nz grid, tile by tile (see log_block_records). The tile is
// summed in LDS in 64-bit fixed point: integer LDS atomics run ~5x faster
// than float ones here. A first pass over the share's values finds its
// record count N and largest magnitude M; terms are scaled by 2^s with M N
// 2^s <= 2^62, so no partial sum overflows and each term is kept to ~2^-44
// of M (float atomics keep 2^-24 of the running sum). A share holding a
// non-finite value is summed with float atomics as it stands.
__global__ __launch_bounds__(kReduceBlock) void log_reduce_kernel(const int4 *__restrict__ rec,
                                                              const unsigned *__restrict__ ctl,
                                                              const unsigned *__restrict__ list,
                                                              const unsigned *__restrict__ fill, unsigned nchunks,
                                                              int nx, int nxy, int64_t nvox, int tile,
                                                              float *__restrict__ g) {
  extern __shared__ long long acc[];  // tile words (floats in the first half for a non-finite share)
  __shared__ float wave_max[kReduceBlock / 64];
  __shared__ unsigned wave_count[kReduceBlock / 64];
  const int64_t lo = int64_t(blockIdx.y) * tile;
  const int64_t hi = min(lo + int64_t(tile), nvox);
  float m = 0.0f;
  unsigned total = 0;
  log_block_records(rec, ctl, list, fill, nchunks, [&](const int4 &e) {
    const float a = fabsf(__int_as_float(e.w));
    m = (a <= m) ? m : (a == a ? a : INFINITY);  // NaN and inf both give inf
    ++total;
  });
  for (int off = 32; off > 0; off >>= 1) {
    m = fmaxf(m, __shfl_xor(m, off, 64));
    total += __shfl_xor(total, off, 64);
  }
  if ((threadIdx.x & 63u) == 0) {
    wave_max[threadIdx.x >> 6] = m;
    wave_count[threadIdx.x >> 6] = total;
  }
  __syncthreads();
  m = 0.0f;
  total = 0;
  for (int w = 0; w < kReduceBlock / 64; ++w) {
    m = fmaxf(m, wave_max[w]);
    total += wave_count[w];
  }
  if (total == 0 || m == 0.0f) return;
  if (isfinite(m)) {
    int em = 0, en = 0;
    (void)frexpf(m, &em);               // m < 2^em
    (void)frexpf(float(total), &en);    // total <= 2^en
    const int sc = 62 - em - en;
    for (int v = threadIdx.x; v < tile; v += blockDim.x) acc[v] = 0;
    __syncthreads();
    log_block_records(rec, ctl, list, fill, nchunks, [&](const int4 &e) {
      record_corners(e, nx, nxy, lo, hi, [&](int64_t a, float t) {
        atomicAdd(reinterpret_cast<unsigned long long *>(&acc[a]), (unsigned long long)__float2ll_rn(ldexpf(t, sc)));
      });
    });
    __syncthreads();
    for (int64_t v = threadIdx.x; v < hi - lo; v += blockDim.x)
      if (acc[v] != 0) atomicAdd(&g[lo + v], ldexpf(__ll2float_rn(acc[v]), -sc));
  } else {
    float *facc = reinterpret_cast<float *>(acc);
    for (int v = threadIdx.x; v < tile; v += blockDim.x) facc[v] = 0.0f;
    __syncthreads();
    log_block_records(rec, ctl, list, fill, nchunks, [&](const int4 &e) {
      record_corners(e, nx, nxy, lo, hi, [&](int64_t a, float t) { atomicAdd(&facc[a], t); });
    });
    __syncthreads();
    for (int64_t v = threadIdx.x; v < hi - lo; v += blockDim.x)
      if (facc[v] != 0.0f) atomicAdd(&g[lo + v], facc[v]);
  }
}

// The per-sample integrators' adjoint (volpathsimple, surface path tracer) as
// a persistent launch (see adjoint_persistent_kernel): lanes pull work units
// (a pixel's chunk of camera samples) from a device counter and run
// adjoint_one_sample for each, so long pixels do not hold their wavefronts at
// the end of the launch. The result per sample is that of adjoint_kernel;
// only the order of the gradient scatter-adds differs.
template <int INTEG, int F>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<INTEG>))) void
adjoint_sample_persistent_kernel(
    SceneView S, IntegArgs I, int sensor, uint32_t seed, int spp, int chunk_log2, const float *__restrict__ grad_img,
    GradTargets G, unsigned *queue, unsigned total_units) {
  extern __shared__ int lds_stack[];  // kLdsWordsPerThread words per thread
  const LdsStack stack = lds_of(S, lds_stack);
  const SensorInfo &sen = S.sensors[sensor];
  const int film_w = sen.film_w;
  LocalCounters cnt{S.counters};
  int s = 0, s_end = 0, px = 0, py = 0;
  int64_t pixel = 0;
  V3 dL = v3(0.f);
  const float inv = float(spp);
  while (true) {
    if (s >= s_end) {
      // the lanes that need a unit in this trip share one atomic
      const uint64_t m = __ballot(1);
      const int leader = __ffsll((unsigned long long)m) - 1;
      unsigned base = 0;
      if (int(__lane_id()) == leader) base = atomicAdd(queue, unsigned(__popcll(m)));
      base = __shfl(base, leader, 64);
      const unsigned u = base + unsigned(__popcll(m & __lanemask_lt()));
      if (u >= total_units) break;
      pixel = int64_t(u >> chunk_log2);
      const int chunk = int(u & ((1u << chunk_log2) - 1u));
      px = int(pixel % film_w);
      py = int(pixel / film_w);
      s = sample_end(chunk - 1, spp, chunk_log2);
      s_end = sample_end(chunk, spp, chunk_log2);
      dL = v3(grad_img[3 * pixel] / inv, grad_img[3 * pixel + 1] / inv, grad_img[3 * pixel + 2] / inv);
      if (!any_nonzero(dL)) {
        s = s_end;
        continue;
      }
    }
    adjoint_one_sample<INTEG, F>(S, I, sen, seed, px, py, pixel, spp, s, dL, G, stack, cnt);
    ++s;
  }
}

// The volume path tracer's render as a persistent launch (see
// adjoint_persistent_kernel): a lane traces the samples of a work unit (a
// pixel's chunk) in order and stores their sum in the unit's slot of
// `partial`; reduce_units_kernel then forms each pixel's mean from its units
// in order, so the image does not depend on which lane took which unit.
template <int NS, int F>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(kWavesPerEuRender<NS, RCV_INTEGRATOR_VOLPATH>))) void
render_persistent_kernel(SceneView S, IntegArgs I, int sensor, uint32_t seed, int spp, int chunk_log2,
                         float *__restrict__ partial, unsigned *queue, unsigned total_units) {
  extern __shared__ int lds_stack[];  // kLdsWordsPerThread words per thread
  const LdsStack stack = lds_of(S, lds_stack);
  const SensorInfo &sen = S.sensors[sensor];
  const int film_w = sen.film_w;
  LocalCounters cnt{S.counters};
  VolPath<NS, false, F> t;
  bool live = false;
  int s = 0, s_end = 0, px = 0, py = 0;
  unsigned u = 0;
  int64_t pixel = 0;
  V3 acc[NS];
  while (true) {
    if (!live) {
      if (s >= s_end) {
        const uint64_t m = __ballot(1);
        const int leader = __ffsll((unsigned long long)m) - 1;
        unsigned base = 0;
        if (int(__lane_id()) == leader) base = atomicAdd(queue, unsigned(__popcll(m)));
        base = __shfl(base, leader, 64);
        u = base + unsigned(__popcll(m & __lanemask_lt()));
        if (u >= total_units) break;
        pixel = int64_t(u >> chunk_log2);
        const int chunk = int(u & ((1u << chunk_log2) - 1u));
        px = int(pixel % film_w);
        py = int(pixel / film_w);
        s = sample_end(chunk - 1, spp, chunk_log2);
        s_end = sample_end(chunk, spp, chunk_log2);
#pragma unroll
        for (int k = 0; k < NS; ++k) acc[k] = v3(0.f);
      }
      Pcg32 rng = lane_rng(seed, uint32_t(pixel * spp + s));
      const Ray r = camera_ray(sen, px, py, rng);
      t.begin(rng, r, cnt, false);
    }
    live = t.bounce(S, I, stack, cnt, nullptr);
    if (!live) {
#pragma unroll
      for (int k = 0; k < NS; ++k) acc[k] += t.L[k];
      if (++s >= s_end) {
        float *o = partial + int64_t(u) * (3 * NS);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          o[3 * k] = acc[k].x;
          o[3 * k + 1] = acc[k].y;
          o[3 * k + 2] = acc[k].z;
        }
      }
    }
  }
}

// Pixel mean from the unit sums, units in order (a pixel of two units adds
// them as the cross-lane reduction of render_kernel does).
template <int NS>
__global__ void reduce_units_kernel(const float *__restrict__ partial, int64_t n_pix, int chunk_log2, int spp,
                                    float *__restrict__ img_old, float *__restrict__ img_new) {
  const int64_t pixel = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pixel >= n_pix) return;
  const int units = 1 << chunk_log2;
  const float inv = float(spp);
  for (int k = 0; k < NS; ++k) {
    float x = 0.f, y = 0.f, z = 0.f;
    for (int c = 0; c < units; ++c) {
      const float *src = partial + ((pixel << chunk_log2) + c) * (3 * NS) + 3 * k;
      if (c == 0) {
        x = src[0];
        y = src[1];
        z = src[2];
      } else {
        x += src[0];
        y += src[1];
        z += src[2];
      }
    }
    float *o = (NS == 2 ? (k == 0 ? img_old : img_new) : img_new) + 3 * pixel;
    o[0] = x / inv;
    o[1] = y / inv;
    o[2] = z / inv;
  }
}

__global__ void zero_u32_kernel(unsigned *p) { p[threadIdx.x] = 0u; }

__global__ void zero_counters_kernel(unsigned long long *c) {
  if (threadIdx.x < kNumCounters) c[threadIdx.x] = 0ull;
}


}  // namespace

static size_t stack_bytes(int block) { return size_t(kLdsWordsPerThread) * block * sizeof(int); }

// Blocks of `kern` the current device holds at once (compute units x
// resident blocks per unit), cached per (device, kernel, block size).
static unsigned resident_grid(const void *kern, int block) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void *, int>, unsigned> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_tuple(dev, kern, block);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cus = 0, per_cu = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, stack_bytes(block));
  const unsigned g = unsigned(std::max(1, per_cu) * std::max(1, cus));
  cache[key] = g;
  return g;
}

// Scatter log of the persistent volume adjoint: a pool of kLogPool chunks
// of kLogChunk 16-byte records (3.2 GB, and 100 MB of chunk lists), one per
// (device, stream), kept for the process (launches on one stream run in
// order; other streams get their own). The reduction tiles hold up to
// 16384 voxels (128 KB of LDS, less where the device allows less); grids of
// more than kLogMaxTiles tiles, or whose footprints span more than two
// tiles, keep the atomics. kLogSplits blocks share each tile's reduction.
constexpr unsigned kLogPool = 786432;
constexpr int kLogSplits = 64;

struct LogBuf {
  int4 *rec = nullptr;
  unsigned *ctl = nullptr;   // 1 + kLogBins counters
  unsigned *list = nullptr;  // kLogBins x kLogPool chunk indices
  unsigned *fill = nullptr;  // kLogPool record counts
};

static int log_buffer(hipStream_t stream, LogBuf &out) {
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, LogBuf> bufs;
  int dev = 0;
  RCV_TRY_HIP(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(dev, stream);
  auto it = bufs.find(key);
  if (it == bufs.end()) {
    LogBuf nb;
    const size_t recs = size_t(kLogPool) * kLogChunk;
    hipError_t e = hipMalloc(&nb.rec, recs * sizeof(int4));
    if (e == hipSuccess) e = hipMalloc(&nb.ctl, (1 + kLogBins) * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&nb.list, size_t(kLogBins) * kLogPool * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&nb.fill, size_t(kLogPool) * sizeof(unsigned));
    if (e != hipSuccess) {
      (void)hipFree(nb.rec);
      (void)hipFree(nb.ctl);
      (void)hipFree(nb.list);
      (void)hipFree(nb.fill);
      return hip_fail(e, "scatter log: allocation");
    }
    it = bufs.emplace(key, nb).first;
  }
  out = it->second;
  return RCV_OK;
}

// Chunks of the pool a launch may draw: kLogPool, or fewer when the
// environment sets RCV_SCATTER_LOG_CHUNKS (a test of the adjoint past the
// pool's end, where the scatter-adds fall back to atomics).
static unsigned log_pool_chunks() {
  const char *e = std::getenv("RCV_SCATTER_LOG_CHUNKS");
  if (!e || !*e) return kLogPool;
  char *end = nullptr;
  const unsigned long v = std::strtoul(e, &end, 10);
  if (end == e || *end != '\0' || v == 0) return kLogPool;
  return unsigned(std::min<unsigned long>(v, kLogPool));
}

// Voxels per reduction tile on the current device: 16384 (a 64-bit word
// each, 128 KB of LDS), or the largest power of two its LDS per block holds.
static int log_tile() {
  static std::mutex mu;
  static std::map<int, int> tiles;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  auto it = tiles.find(dev);
  if (it != tiles.end()) return it->second;
  int max_lds = 0;
  (void)hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  constexpr size_t kStatic = 512;  // the kernel's static LDS, rounded up
  int tile = 16384;
  while (tile > 1024 && size_t(tile) * sizeof(long long) + kStatic > size_t(max_lds)) tile >>= 1;
  if (tile * sizeof(long long) > 65536 &&
      hipFuncSetAttribute(reinterpret_cast<const void *>(log_reduce_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, int(tile * sizeof(long long))) != hipSuccess) {
    (void)hipGetLastError();
    tile = 65536 / int(sizeof(long long)) / 2;
  }
  tiles[dev] = tile;
  return tile;
}

// Work-unit split of a persistent volume launch: the chunk count of the
// static launch (enough units to fill the device), raised while a unit keeps
// at least min_samples samples.
static int unit_chunk_log2(int spp, int static_cl2, int min_samples) {
  int ucl2 = 0;
  while ((1 << ucl2) < 64 && (spp >> (ucl2 + 1)) >= min_samples) ++ucl2;
  return std::max(ucl2, static_cl2);
}

int plan_launch(const rcv_integrator_desc *d, int64_t n_pix, int spp, int &spt, int &chunk_log2, int max_spt,
                int64_t min_threads) {
  if (d && d->block_size != 0 && (d->block_size < 64 || d->block_size > kMaxBlock || d->block_size % 64 != 0))
    return fail(RCV_E_INVALID, "block_size %d must be 0 or a multiple of 64 in [64, %d]", d->block_size, kMaxBlock);
  int chunks;
  int cap = 1;
  while (cap < 64 && 2 * cap <= spp) cap <<= 1;  // largest power of two <= min(spp, 64)
  if (d && d->samples_per_thread > 0) {
    if (spp % d->samples_per_thread != 0)
      return fail(RCV_E_INVALID, "samples_per_thread %d must divide spp %d", d->samples_per_thread, spp);
    chunks = spp / d->samples_per_thread;
    if (chunks > 64 || (chunks & (chunks - 1)) != 0)
      return fail(RCV_E_INVALID, "spp / samples_per_thread = %d must be a power of two <= 64", chunks);
  } else {
    chunks = 1;
    while (chunks < cap && (n_pix * chunks < min_threads || (max_spt > 0 && spp > max_spt * chunks)))
      chunks <<= 1;
  }
  spt = (spp + chunks - 1) / chunks;  // samples of the busiest thread
  chunk_log2 = 0;
  while ((1 << chunk_log2) < chunks) ++chunk_log2;
  return RCV_OK;
}

static int check_states(const Scene *s, int ns, const char *fn) {
  auto need = [&](int param, const char *what) -> int {
    if (param < 0) return RCV_OK;
    const Param &p = s->params[param];
    if (!p.data[RCV_STATE_NEW]) return fail(RCV_E_STATE, "%s: %s parameter %d has no NEW state bound", fn, what, param);
    if (ns == 2 && !p.data[RCV_STATE_OLD])
      return fail(RCV_E_STATE, "%s: %s parameter %d has no OLD state bound", fn, what, param);
    return RCV_OK;
  };
  for (auto &b : s->bsdfs)
    if (int rc = need(b.texture_param, "texture")) return rc;
  for (auto &m : s->media) {
    if (int rc = need(m.sigma_param, "sigma_t")) return rc;
    if (int rc = need(m.albedo_param, "albedo")) return rc;
  }
  return RCV_OK;
}

struct RenderLaunch {
  dim3 grid;
  int block;
  hipStream_t stream;
  const SceneView *view;
  IntegArgs I;
  int sensor;
  uint32_t seed;
  int spp, spt, cl2;
  float *img_old, *img_new;
};

template <int NS, int INTEG, int F>
static void launch_render_kernel(const RenderLaunch &L) {
  hipLaunchKernelGGL((render_kernel<NS, INTEG, F>), L.grid, dim3(L.block), stack_bytes(L.block), L.stream, *L.view,
                     L.I, L.sensor,
                     L.seed, L.spp, L.spt, L.cl2, L.img_old, L.img_new);
}

// get_single_medium and the class docstring of plugins/volpathsimple.py
// (:9-30): one shape with an interior medium and a null BSDF, the
// environment as the only emitter.
static int check_simple(const Scene *s, const char *fn) {
  if (s->shapes.size() != 1)
    return fail(RCV_E_SCENE, "%s: volpathsimple: Not supported: more than 1 shape in the scene (found %d).", fn,
                int(s->shapes.size()));
  if (s->media.size() != 1 || s->shapes[0].interior != 0)
    return fail(RCV_E_SCENE, "%s: volpathsimple: Expected a single shape with an interior medium.", fn);
  if (s->features() & kFeatSurfaces)
    return fail(RCV_E_SCENE, "%s: volpathsimple: the medium boundary must use a null BSDF", fn);
  if (s->features() & kFeatHomog)
    return fail(RCV_E_SCENE, "%s: volpathsimple: heterogeneous media only", fn);
  if (s->emitters.size() != 1 || s->emitters[0].type != RCV_EMITTER_CONSTANT)
    return fail(RCV_E_SCENE, "%s: volpathsimple: the only emitter must be an infinite light", fn);
  return RCV_OK;
}

// with_features over the masks without surfaces (volpathsimple scenes).
template <typename Fn>
static void with_volume_features(int f, Fn &&fn) {
  switch (f & (kFeatAlbedoGrid | kFeatSupergrid)) {
    case 0: fn(std::integral_constant<int, 0>{}); break;
    case kFeatAlbedoGrid: fn(std::integral_constant<int, kFeatAlbedoGrid>{}); break;
    case kFeatSupergrid: fn(std::integral_constant<int, kFeatSupergrid>{}); break;
    default: fn(std::integral_constant<int, kFeatAlbedoGrid | kFeatSupergrid>{}); break;
  }
}

int launch_render(const Scene *s, const rcv_integrator_desc *d, int sensor, uint32_t seed, int spp,
                  float *img_old, float *img_new, int ns, hipStream_t stream, const char *fn) {
  RCV_REQUIRE(s && s->committed, "%s: scene not committed", fn);
  RCV_REQUIRE(d, "%s: null integrator", fn);
  RCV_REQUIRE(d->type == RCV_INTEGRATOR_PATH || d->type == RCV_INTEGRATOR_VOLPATH ||
                  d->type == RCV_INTEGRATOR_VOLPATH_SIMPLE,
              "%s: bad integrator type", fn);
  if (d->type == RCV_INTEGRATOR_VOLPATH_SIMPLE) {
    RCV_REQUIRE(ns == 1, "%s: volpathsimple renders one state", fn);
    if (int rc = check_simple(s, fn)) return rc;
  }
  RCV_REQUIRE(sensor >= 0 && sensor < int(s->sensors.size()), "%s: bad sensor %d", fn, sensor);
  RCV_REQUIRE(spp > 0, "%s: spp must be positive", fn);
  RCV_REQUIRE(img_new && (ns == 1 || img_old), "%s: null image", fn);
  if (int rc = check_states(s, ns, fn)) return rc;
  const SensorInfo &sen = s->sensors[sensor];
  const int64_t n_pix = int64_t(sen.film_w) * sen.film_h;
  RCV_REQUIRE(n_pix * spp < (int64_t(1) << 32), "%s: pixels * spp must stay below 2^32", fn);
  int spt, cl2;
  if (int rc = plan_launch(d, n_pix, spp, spt, cl2)) return rc;
  const IntegArgs I{d->max_depth, d->rr_depth, d->hide_emitters, d->flags};
  const int64_t threads = n_pix << cl2;
  const int block = d->block_size ? d->block_size : kDefaultBlock;
  const unsigned grid = ceil_div(threads, block);
  if (s->view.counters) hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, stream, s->view.counters);
  const RenderLaunch L{dim3(grid), block, stream, &s->view, I, sensor, seed, spp, spt, cl2, img_old, img_new};
  if (d->type == RCV_INTEGRATOR_VOLPATH_SIMPLE) {
    with_volume_features(s->features(), [&](auto f) {
      launch_render_kernel<1, RCV_INTEGRATOR_VOLPATH_SIMPLE, decltype(f)::value>(L);
    });
  } else if (d->type == RCV_INTEGRATOR_VOLPATH && !d->samples_per_thread && (1 << cl2) <= kRenderMaxUnitsPerPixel) {
    // persistent, with at most kRenderMaxUnitsPerPixel units per pixel (the
    // unit-sum buffer); smaller images that need more units per pixel to
    // fill the device take the static launch below
    int ucl2 = cl2;
    while ((2 << ucl2) <= kRenderMaxUnitsPerPixel && (spp >> (ucl2 + 1)) >= kVolUnitSamples) ++ucl2;
    const unsigned units = unsigned(n_pix << ucl2);
    RCV_REQUIRE(int64_t(units) * 3 * ns <= s->partial_cap, "%s: unit buffer too small", fn);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, stream, s->work_queue);
    auto go = [&](auto kern) {
      const unsigned pgrid = std::min(resident_grid(reinterpret_cast<const void *>(kern), block), grid);
      hipLaunchKernelGGL(kern, dim3(pgrid), dim3(block), stack_bytes(block), stream, s->view, I, sensor, seed, spp,
                         ucl2, s->partial, s->work_queue, units);
    };
    if (ns == 2) {
      with_render_features(s->features(), [&](auto f) { go(render_persistent_kernel<2, decltype(f)::value>); });
      hipLaunchKernelGGL(reduce_units_kernel<2>, dim3(ceil_div(n_pix, 256)), dim3(256), 0, stream, s->partial, n_pix,
                         ucl2, spp, img_old, img_new);
    } else {
      with_render_features(s->features(), [&](auto f) { go(render_persistent_kernel<1, decltype(f)::value>); });
      hipLaunchKernelGGL(reduce_units_kernel<1>, dim3(ceil_div(n_pix, 256)), dim3(256), 0, stream, s->partial, n_pix,
                         ucl2, spp, img_old, img_new);
    }
  } else if (d->type == RCV_INTEGRATOR_VOLPATH) {
    if (ns == 2)
      with_render_features(s->features(), [&](auto f) { launch_render_kernel<2, RCV_INTEGRATOR_VOLPATH, decltype(f)::value>(L); });
    else
      with_render_features(s->features(), [&](auto f) { launch_render_kernel<1, RCV_INTEGRATOR_VOLPATH, decltype(f)::value>(L); });
  } else {
    if (ns == 2)
      launch_render_kernel<2, RCV_INTEGRATOR_PATH, 0>(L);
    else
      launch_render_kernel<1, RCV_INTEGRATOR_PATH, 0>(L);
  }
  return check_launch(fn);
}

int launch_adjoint(const Scene *s, const rcv_integrator_desc *d, int sensor, uint32_t seed, int spp,
                   const float *grad_img, int n_params, const int32_t *param_ids, float *const *grad_ptrs,
                   hipStream_t stream) {
  const char *fn = "rcv_render_backward";
  RCV_REQUIRE(s && s->committed, "%s: scene not committed", fn);
  RCV_REQUIRE(d, "%s: null integrator", fn);
  RCV_REQUIRE(sensor >= 0 && sensor < int(s->sensors.size()), "%s: bad sensor %d", fn, sensor);
  RCV_REQUIRE(spp > 0 && grad_img, "%s: bad spp or null grad_img", fn);
  RCV_REQUIRE(n_params >= 0 && (n_params == 0 || (param_ids && grad_ptrs)), "%s: bad gradient targets", fn);
  RCV_REQUIRE(d->type == RCV_INTEGRATOR_PATH || d->type == RCV_INTEGRATOR_VOLPATH ||
                  d->type == RCV_INTEGRATOR_VOLPATH_SIMPLE,
              "%s: bad integrator type", fn);
  if (d->type == RCV_INTEGRATOR_VOLPATH_SIMPLE)
    if (int rc = check_simple(s, fn)) return rc;
  if (int rc = check_states(s, 1, fn)) return rc;
  GradTargets G;
  G.clear();
  for (int i = 0; i < n_params; ++i) {
    RCV_REQUIRE(param_ids[i] >= 0 && param_ids[i] < int(s->params.size()), "%s: bad param id %d", fn, param_ids[i]);
    RCV_REQUIRE(grad_ptrs[i], "%s: null gradient buffer for param %d", fn, param_ids[i]);
    G.g[param_ids[i]] = grad_ptrs[i];
  }
  const SensorInfo &sen = s->sensors[sensor];
  const int64_t n_pix = int64_t(sen.film_w) * sen.film_h;
  RCV_REQUIRE(n_pix * spp < (int64_t(1) << 32), "%s: pixels * spp must stay below 2^32", fn);
  int spt, cl2;
  if (int rc = plan_launch(d, n_pix, spp, spt, cl2)) return rc;
  const IntegArgs I{d->max_depth, d->rr_depth, d->hide_emitters, d->flags};
  const int block = d->block_size ? d->block_size : kDefaultBlock;
  const unsigned grid = ceil_div(n_pix << cl2, block);
  if (s->view.counters) hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, stream, s->view.counters);
  // the per-sample integrators' persistent adjoint, as the volume path
  // tracer's below
  auto launch_sample_persistent = [&](auto kern) {
    const int ucl2 = unit_chunk_log2(spp, cl2, kAdjUnitSamples);
    const unsigned units = unsigned(n_pix << ucl2);
    const unsigned pgrid =
        std::min(resident_grid(reinterpret_cast<const void *>(kern), block), ceil_div(int64_t(units), block));
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, stream, s->work_queue);
    hipLaunchKernelGGL(kern, dim3(pgrid), dim3(block), stack_bytes(block), stream, s->view, I, sensor, seed, spp, ucl2,
                       grad_img, G, s->work_queue, units);
  };
  if (d->type == RCV_INTEGRATOR_VOLPATH_SIMPLE && !d->samples_per_thread)
    with_volume_features(s->features(), [&](auto f) {
      launch_sample_persistent(adjoint_sample_persistent_kernel<RCV_INTEGRATOR_VOLPATH_SIMPLE, decltype(f)::value>);
    });
  else if (d->type == RCV_INTEGRATOR_VOLPATH_SIMPLE)
    with_volume_features(s->features(), [&](auto f) {
      hipLaunchKernelGGL((adjoint_kernel<RCV_INTEGRATOR_VOLPATH_SIMPLE, decltype(f)::value>), dim3(grid),
                         dim3(block), stack_bytes(block), stream, s->view, I, sensor, seed, spp, spt, cl2, grad_img,
                         G);
    });
  else if (d->type == RCV_INTEGRATOR_VOLPATH && !d->samples_per_thread)
    with_render_features(s->features(), [&](auto f) {
      // the waves that fit, pulling units of >= kVolUnitSamples samples
      // (fewer where the image is too small to fill the device otherwise)
      constexpr int F = decltype(f)::value;
      // (taped kernels keep kAdjUnitSamples: their per-lane tape in scratch
      // gains nothing from coherence; the 64^3 k = 8 adjoint took 59.4 ms
      // with 2-sample units against 51.2 ms)
      const int ucl2 = unit_chunk_log2(spp, cl2, kTapeAdjoint<F> ? kAdjUnitSamples : kVolUnitSamples);
      const unsigned units = unsigned(n_pix << ucl2);
      // the density gradient of the first heterogeneous medium that has one
      // goes through the scatter log when its grid fits the reduction tiles
      // and a footprint (up to nx ny + nx + 1 voxels past its base) spans
      // at most two of them
      GradTargets GL = G;
      int lp = -1;
      int64_t nvox = 0;
      const int tile = log_tile();
      if constexpr (!kTapeAdjoint<F>) {
        for (const MediumInfo &M : s->media) {
          if (M.homogeneous || M.sigma_param < 0 || !G.g[M.sigma_param]) continue;
          const Param &P = s->params[M.sigma_param];
          const int64_t n = int64_t(P.nx) * P.ny * P.nz;
          if (n <= int64_t(tile) * kLogMaxTiles && n < (int64_t(1) << kRecBaseBits) &&
              int64_t(P.nx) * P.ny + P.nx + 1 <= tile) {
            lp = M.sigma_param;
            nvox = n;
          }
          break;
        }
      }
      LogBuf lb;
      // without memory for the log the scatter-adds stay atomics
      if (lp >= 0 && log_buffer(stream, lb) != RCV_OK) {
        (void)hipGetLastError();
        lp = -1;
      }
      if (lp >= 0) {
        GL.log_rec = lb.rec;
        GL.log_ctl = lb.ctl;
        GL.log_list = lb.list;
        GL.log_fill = lb.fill;
        GL.log_chunks = log_pool_chunks();
        GL.log_shift = __builtin_ctz(unsigned(tile));
        GL.log_param = lp;
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1 + kLogBins), 0, stream, lb.ctl);
      }
      // three waves per SIMD with the log, four with atomics (kWavesPerEuLoggedAdjoint)
      auto go = [&](auto kern) {
        const unsigned pgrid =
            std::min(resident_grid(reinterpret_cast<const void *>(kern), block), ceil_div(int64_t(units), block));
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, stream, s->work_queue);
        hipLaunchKernelGGL(kern, dim3(pgrid), dim3(block), stack_bytes(block), stream, s->view, I, sensor, seed, spp,
                           ucl2, grad_img, GL, s->work_queue, units);
      };
      if constexpr (kTapeAdjoint<F>)
        go(adjoint_persistent_kernel<F, kWavesPerEu<RCV_INTEGRATOR_VOLPATH>>);
      else if (lp >= 0)
        go(adjoint_persistent_kernel<F, kWavesPerEuLoggedAdjoint>);
      else
        go(adjoint_persistent_kernel<F, kWavesPerEu<RCV_INTEGRATOR_VOLPATH>>);
      if (lp >= 0) {
        const Param &P = s->params[lp];
        const int tiles = int((nvox + tile - 1) / tile);
        hipLaunchKernelGGL(log_reduce_kernel, dim3(kLogSplits, tiles), dim3(kReduceBlock), size_t(tile) * sizeof(long long),
                           stream, lb.rec, lb.ctl, lb.list, lb.fill, GL.log_chunks, int(P.nx),
                           int(P.nx) * int(P.ny), nvox, tile, G.g[lp]);
      }
    });
  else if (d->type == RCV_INTEGRATOR_VOLPATH)
    with_render_features(s->features(), [&](auto f) {
      hipLaunchKernelGGL((adjoint_kernel<RCV_INTEGRATOR_VOLPATH, decltype(f)::value>), dim3(grid), dim3(block),
                         stack_bytes(block), stream, s->view, I, sensor, seed, spp, spt, cl2, grad_img, G);
    });
  else if (!d->samples_per_thread)
    launch_sample_persistent(adjoint_sample_persistent_kernel<RCV_INTEGRATOR_PATH, 0>);
  else
    hipLaunchKernelGGL((adjoint_kernel<RCV_INTEGRATOR_PATH, 0>), dim3(grid), dim3(block), stack_bytes(block), stream,
                       s->view, I, sensor, seed, spp, spt, cl2, grad_img, G);
  return check_launch(fn);
}

int launch_forward(const Scene *s, const rcv_integrator_desc *d, int sensor, uint32_t seed, int spp,
                   int n_params, const int32_t *param_ids, const float *const *tangents, float *dimg,
                   hipStream_t stream) {
  const char *fn = "rcv_render_forward";
  RCV_REQUIRE(s && s->committed, "%s: scene not committed", fn);
  RCV_REQUIRE(d, "%s: null integrator", fn);
  RCV_REQUIRE(sensor >= 0 && sensor < int(s->sensors.size()), "%s: bad sensor %d", fn, sensor);
  RCV_REQUIRE(spp > 0 && dimg, "%s: bad spp or null output image", fn);
  RCV_REQUIRE(n_params >= 0 && (n_params == 0 || (param_ids && tangents)), "%s: bad tangent arguments", fn);
  // volpathsimple's adjoint is written with backward-mode calls only
  // (dr.backward_from, plugins/volpathsimple.py:183,502,593,620): no forward mode
  RCV_REQUIRE(d->type == RCV_INTEGRATOR_PATH || d->type == RCV_INTEGRATOR_VOLPATH,
              "%s: forward mode exists for the path-replay integrators (prb, prbvolpath) only", fn);
  if (int rc = check_states(s, 1, fn)) return rc;
  GradTargets T;
  T.clear();
  for (int i = 0; i < n_params; ++i) {
    RCV_REQUIRE(param_ids[i] >= 0 && param_ids[i] < int(s->params.size()), "%s: bad param id %d", fn, param_ids[i]);
    RCV_REQUIRE(tangents[i], "%s: null tangent for param %d", fn, param_ids[i]);
    T.g[param_ids[i]] = const_cast<float *>(tangents[i]);
  }
  const SensorInfo &sen = s->sensors[sensor];
  const int64_t n_pix = int64_t(sen.film_w) * sen.film_h;
  RCV_REQUIRE(n_pix * spp < (int64_t(1) << 32), "%s: pixels * spp must stay below 2^32", fn);
  int spt, cl2;
  if (int rc = plan_launch(d, n_pix, spp, spt, cl2)) return rc;
  const IntegArgs I{d->max_depth, d->rr_depth, d->hide_emitters, d->flags};
  const int block = d->block_size ? d->block_size : kDefaultBlock;
  const unsigned grid = ceil_div(n_pix << cl2, block);
  if (s->view.counters) hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, stream, s->view.counters);
  if (d->type == RCV_INTEGRATOR_VOLPATH)
    with_render_features(s->features(), [&](auto f) {
      hipLaunchKernelGGL((forward_kernel<RCV_INTEGRATOR_VOLPATH, decltype(f)::value>), dim3(grid), dim3(block),
                         stack_bytes(block), stream, s->view, I, sensor, seed, spp, spt, cl2, T, dimg);
    });
  else
    hipLaunchKernelGGL((forward_kernel<RCV_INTEGRATOR_PATH, 0>), dim3(grid), dim3(block), stack_bytes(block), stream,
                       s->view, I, sensor, seed, spp, spt, cl2, T, dimg);
  return check_launch(fn);
}

namespace {
__global__ void supergrid_kernel(const float *__restrict__ grid, int nx, int ny, int nz, int sx, int sy, int sz,
                                 float *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= sx * sy * sz) return;
  out[c] = supergrid_cell_max(grid, nx, ny, nz, sx, sy, sz, c);
}
}  // namespace

int Scene::build_supergrid(int medium, int state, hipStream_t stream) {
  const MediumInfo &m = media[medium];
  const Param &p = params[m.sigma_param];
  const int cells = m.sg_res[0] * m.sg_res[1] * m.sg_res[2];
  float *out = supergrids + m.supergrid_offset + state * cells;
  hipLaunchKernelGGL(supergrid_kernel, dim3(ceil_div(cells, 128)), dim3(128), 0, stream, p.data[state], p.nx, p.ny,
                     p.nz, m.sg_res[0], m.sg_res[1], m.sg_res[2], out);
  return check_launch("majorant supergrid");
}

}  // namespace rcv

using namespace rcv;

extern "C" int rcv_render(const rcv_scene *scene, const rcv_integrator_desc *integrator, int32_t sensor,
                          uint32_t seed, int32_t spp, float *img, void *stream) {
  return launch_render(static_cast<const Scene *>(scene), integrator, sensor, seed, spp, nullptr, img, 1,
                       as_stream(stream), "rcv_render");
}

extern "C" int rcv_render_twostates(const rcv_scene *scene, const rcv_integrator_desc *integrator,
                                    int32_t sensor, uint32_t seed, int32_t spp, float *img_old, float *img_new,
                                    void *stream) {
  return launch_render(static_cast<const Scene *>(scene), integrator, sensor, seed, spp, img_old, img_new, 2,
                       as_stream(stream), "rcv_render_twostates");
}

extern "C" int rcv_render_forward(const rcv_scene *scene, const rcv_integrator_desc *integrator, int32_t sensor,
                                  uint32_t seed, int32_t spp, int32_t n_params, const int32_t *param_ids,
                                  const float *const *tangents, float *dimg, void *stream) {
  return launch_forward(static_cast<const Scene *>(scene), integrator, sensor, seed, spp, n_params, param_ids,
                        tangents, dimg, as_stream(stream));
}

extern "C" int rcv_render_backward(const rcv_scene *scene, const rcv_integrator_desc *integrator, int32_t sensor,
                                   uint32_t seed, int32_t spp, const float *grad_img, int32_t n_params,
                                   const int32_t *param_ids, float *const *grad_ptrs, void *stream) {
  return launch_adjoint(static_cast<const Scene *>(scene), integrator, sensor, seed, spp, grad_img, n_params,
                        param_ids, grad_ptrs, as_stream(stream));
}
