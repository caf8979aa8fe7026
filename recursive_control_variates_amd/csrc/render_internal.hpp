// Launch helpers shared by the render entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/rcv.h"
#include "scene_internal.hpp"

namespace rcv {

// Calls fn(std::integral_constant<int, F>{}) for the run-time feature mask
// f, so a launch site can pick the kernel compiled for exactly the scene's
// features (kFeat* bits).
template <typename Fn>
inline void with_features(int f, Fn &&fn) {
  switch (f & kFeatAll) {
    case 0: fn(std::integral_constant<int, 0>{}); break;
    case 1: fn(std::integral_constant<int, 1>{}); break;
    case 2: fn(std::integral_constant<int, 2>{}); break;
    case 3: fn(std::integral_constant<int, 3>{}); break;
    case 4: fn(std::integral_constant<int, 4>{}); break;
    case 5: fn(std::integral_constant<int, 5>{}); break;
    case 6: fn(std::integral_constant<int, 6>{}); break;
    default: fn(std::integral_constant<int, 7>{}); break;
  }
}

// with_features for the single-scene volume kernels, which also take scenes
// with a homogeneous medium: those run one of two kernels, with surface code
// (and heterogeneous media without albedo grids or supergrids), or with every
// feature.
template <typename Fn>
inline void with_render_features(int f, Fn &&fn) {
  if (f & kFeatHomog) {
    if (f & (kFeatAlbedoGrid | kFeatSupergrid))
      fn(std::integral_constant<int, kFeatAll | kFeatHomog>{});
    else
      fn(std::integral_constant<int, kFeatSurfaces | kFeatHomog>{});
    return;
  }
  with_features(f, fn);
}

// Waves per SIMD of the render kernels: the trackers are bound by memory
// latency, so occupancy pays while spills stay small. Measured: the volume
// path tracer (<= 128 VGPRs, few spills) is fastest at four; volpathsimple and
// the surface path tracer, whose state is larger, at three (<= 168 VGPRs).
template <int INTEG>
constexpr int kWavesPerEu = INTEG == RCV_INTEGRATOR_VOLPATH ? 4 : 3;

// The volume path tracer's two-state render and its adjoint carry more
// state (two throughputs and images; the replay's gradient terms): at four
// waves they spilled 60 / 72 B per lane, and at three (up to 168 VGPRs, no
// spills) they run 12-14 % / 17-18 % faster over the optimisation (two
// waves: no faster than three). That holds for the adjoint whose density
// gradient goes through the scatter log; adjoints that add with device-scope
// atomics (batched, static, grids past the log) are faster at four, whose
// extra wave covers the atomics' waits (the 64-job batch: 348 ms per
// iteration against 376 ms with its adjoint at three). The single-state
// render stays at four.
template <int NS, int INTEG>
constexpr int kWavesPerEuRender = INTEG == RCV_INTEGRATOR_VOLPATH && NS == 2 ? 3 : kWavesPerEu<INTEG>;
constexpr int kWavesPerEuLoggedAdjoint = 3;

// Splits the spp camera samples of a pixel over 2^chunk_log2 threads (at
// most spt samples each): a user-given samples_per_thread must divide spp
// into a power of two <= 64 chunks; otherwise chunks double, up to the
// largest power of two <= min(spp, 64), while the launch has fewer than
// min_threads threads or (max_spt > 0) a thread would trace more than
// max_spt samples.
int plan_launch(const rcv_integrator_desc *d, int64_t n_pix, int spp, int &spt, int &chunk_log2,
                int max_spt = 0, int64_t min_threads = int64_t(1) << 19);

// One past the last sample of chunk c (chunk -1 ends at 0): chunk c traces
// samples [sample_end(c - 1), sample_end(c)), an even split when the chunk
// count divides spp.
__host__ __device__ inline int sample_end(int c, int spp, int chunk_log2) {
  return int((int64_t(c + 1) * spp) >> chunk_log2);
}

int launch_render(const Scene *s, const rcv_integrator_desc *d, int sensor, uint32_t seed, int spp,
                  float *img_old, float *img_new, int ns, hipStream_t stream, const char *fn);

int launch_adjoint(const Scene *s, const rcv_integrator_desc *d, int sensor, uint32_t seed, int spp,
                   const float *grad_img, int n_params, const int32_t *param_ids, float *const *grad_ptrs,
                   hipStream_t stream);

}  // namespace rcv
