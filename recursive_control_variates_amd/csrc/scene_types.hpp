// Plain-old-data layout of a committed scene in device memory, shared by the
// host builder (scene.hip) and the render kernels.
//
// Geometry is a flat primitive list (triangles, then spheres) ordered by a
// binned-SAH BVH whose nodes are 32 bytes; triangles are stored as one
// float4 triple (v0 | e1 | e2, with the shape id in v0.w) so a leaf test is
// three 16-byte loads. Small per-object tables (shapes, BSDFs, emitters,
// media, sensors) live in one device block. Parameter data pointers (the
// optimised textures/grids, two states each) are NOT in device memory: they
// travel by value in the kernel arguments, so rebinding a state is free.
#pragma once

#include <stdint.h>

namespace rcv {

constexpr int kMaxParams = 8;
constexpr int kMaxViews = 8;
constexpr int kMaxBvhDepth = 30;
constexpr int kStackSize = 32;
// Per-thread LDS of the trace kernels, in 4-byte words: the BVH traversal
// stack (as deep as the scene's BVH, SceneView::lds_stack) and then per-lane
// slots (the adjoint's NEE weights, SceneView::lds_slots, at most
// kMaxLdsSlots). 40 words = 160 KB of LDS per CU over the 16 waves of 64
// lanes that the trace kernels' registers allow (four per SIMD).
constexpr int kLdsWordsPerThread = 40;
constexpr int kMaxLdsSlots = 32;
static_assert(kMaxBvhDepth < kLdsWordsPerThread, "a BVH of the deepest kind leaves no LDS for its stack");
// LDS slots a render's deferred NEE keeps its shadow ray in.
constexpr int kNeeFlightWords = 16;
constexpr int kNumCounters = 8;

enum Counter : int {
  kCntPaths = 0,        // camera samples traced
  kCntCollisions = 1,   // tentative collisions sampled in the main loop
  kCntNeeSteps = 2,     // ratio-tracking steps in NEE
  kCntTraversals = 3,   // BVH traversals (closest + any hit)
  kCntRealScatter = 4,  // real medium scattering events
  kCntSurfaceHits = 5,  // non-null surface vertices
  kCntGradAtomics = 6,  // gradient scatter-adds issued (8 per lookup)
  kCntNullEvents = 7,   // null collisions
};

struct BVHNode {
  float lo[3];
  int32_t first;  // leaf: first primitive slot; inner: index of left child (right = first+1)
  float hi[3];
  int32_t count;  // > 0: leaf with `count` primitives
};

// Oriented box primitive: the image of [-1, 1]^3 under an affine map (the
// 'cube' shape, traced as one primitive instead of 12 triangles). inv is the
// world-to-object map (3x4, row-major); n the outward unit world normals of
// the faces +x, -x, +y, -y, +z, -z.
struct BoxPrim {
  float inv[12];
  float n[18];
  int32_t shape;
  int32_t pad;
};

struct ShapeInfo {
  int32_t bsdf;      // -1: none
  int32_t emitter;   // -1: none
  int32_t interior;  // medium id or -1
  int32_t exterior;
};

struct BsdfInfo {
  int32_t type;           // RCV_BSDF_*
  int32_t texture_param;  // -1: constant rgb
  float rgb[3];
  int32_t pad;
};

struct EmitterInfo {
  int32_t type;  // RCV_EMITTER_*
  int32_t shape;
  int32_t tri_begin;  // area: range in the emitter triangle table
  int32_t tri_count;
  float rgb[3];
  float inv_area;
};

struct MediumInfo {
  int32_t sigma_param;   // grid (heterogeneous) or 1-value parameter (homogeneous)
  int32_t albedo_param;  // -1: constant albedo; a 3-channel grid or 3-value parameter
  float albedo[3];
  float scale;
  float lo[3];
  float hi[3];
  int32_t sample_emitters;
  int32_t majorant_factor;
  int32_t supergrid_offset;  // float offset of this medium's supergrids (2 states) or -1
  int32_t sg_res[3];         // supergrid resolution
  int32_t homogeneous;       // 1: constant sigma_t and albedo, unbounded (lo/hi unused)
};

struct ViewInfo {
  float origin[3];
  float left[3];
  float up[3];
  float fwd[3];
  float tan_half;  // tan(fov_x / 2)
  float pad;
};

struct SensorInfo {
  int32_t film_w, film_h, n_views, pad;
  ViewInfo views[kMaxViews];
};

// Scene features a volume kernel is compiled for (its template parameter F):
// code for an absent feature is left out, which keeps the register count of
// the common configurations low enough for four waves per SIMD.
constexpr int kFeatSurfaces = 1;    // a shape whose BSDF is not null (diffuse shading, surface NEE)
constexpr int kFeatAlbedoGrid = 2;  // a medium whose albedo is a grid
constexpr int kFeatSupergrid = 4;   // a medium with a majorant supergrid
constexpr int kFeatAll = 7;
constexpr int kFeatHomog = 8;       // a homogeneous medium

struct ParamBind {
  const float *data[2];  // RCV_STATE_OLD, RCV_STATE_NEW
  int32_t nx, ny, nz, ch;
};

// Kernel-argument view of a committed scene.
struct SceneView {
  const BVHNode *nodes;
  const float4 *tris;        // 3 float4 per triangle slot
  const float2 *tri_uv;      // 3 per triangle slot
  const int32_t *prim_ids;   // slot -> primitive id: triangles, then spheres, then boxes
  const float4 *spheres;     // center.xyz, radius
  const int32_t *sphere_shape;
  const BoxPrim *boxes;
  const ShapeInfo *shapes;
  const BsdfInfo *bsdfs;
  const EmitterInfo *emitters;
  const float *em_cdf;       // per emitter triangle: cumulative area fraction
  const int32_t *em_tris;    // per emitter triangle: primitive id
  const MediumInfo *media;
  const SensorInfo *sensors;
  const float *majorants;    // [param][state] max of the grid data
  const float *supergrids;   // per-medium supergrid majorants (two states)
  unsigned long long *counters;  // kNumCounters or null
  int32_t n_tris;
  int32_t n_sph;
  int32_t n_emitters;
  int32_t env_emitter;  // -1 if none
  int32_t n_media;
  // Integrator flags of upstream prbvolpath's prepare_scene, reached from
  // plugins/twostatevolpath.py:99: null scattering (delta tracking) when any
  // medium is heterogeneous, otherwise free-flight sampling weighted by
  // transmittance / pdf; analytic transmittance of homogeneous media in NEE.
  int32_t handle_null;
  int32_t nee_homog;
  int32_t hom_medium;   // the scene's homogeneous medium (at most one) or -1
  int32_t lds_stack;    // LDS words per thread for the BVH traversal stack (the BVH depth)
  int32_t lds_slots;    // LDS words per thread after it: kLdsWordsPerThread - lds_stack, at most kMaxLdsSlots
  float bs_center[3];
  float bs_radius;
  ParamBind params[kMaxParams];
};

}  // namespace rcv
