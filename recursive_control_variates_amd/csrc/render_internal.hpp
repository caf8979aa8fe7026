// Launch helpers shared by the render entry points.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rcv.h"
#include "scene_internal.hpp"

namespace rcv {

// Splits spp camera samples of a pixel over 2^chunk_log2 threads of spt
// samples each (see render.hip).
int plan_launch(const rcv_integrator_desc *d, int64_t n_pix, int spp, int &spt, int &chunk_log2);

int launch_render(const Scene *s, const rcv_integrator_desc *d, int sensor, uint32_t seed, int spp,
                  float *img_old, float *img_new, int ns, hipStream_t stream, const char *fn);

int launch_adjoint(const Scene *s, const rcv_integrator_desc *d, int sensor, uint32_t seed, int spp,
                   const float *grad_img, int n_params, const int32_t *param_ids, float *const *grad_ptrs,
                   hipStream_t stream);

}  // namespace rcv
