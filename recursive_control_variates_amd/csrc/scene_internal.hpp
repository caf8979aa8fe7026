// Host-side scene object behind the opaque rcv_scene handle.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "scene_types.hpp"

namespace rcv {

// Persistent volume renders split a pixel into at most this many work units;
// their unit-sum buffer is sized for it at scene commit. Four units (8 of 32
// samples each) against two: the 64^3 two-state render at the mid / final
// state of the optimisation 15.9 / 21.5 -> 12.6 / 15.8 ms, 3.1 -> 3.7 ms at
// the uniform initial state (eight: 12.3 / 14.8, 4.4 ms).
constexpr int kRenderMaxUnitsPerPixel = 4;

struct Param {
  int kind = 0;
  int nx = 0, ny = 0, nz = 0, ch = 0;
  const float *data[2] = {nullptr, nullptr};
};

struct Scene {
  // collected before commit
  std::vector<Param> params;
  std::vector<BsdfInfo> bsdfs;
  std::vector<MediumInfo> media;
  std::vector<EmitterInfo> emitters;
  std::vector<ShapeInfo> shapes;
  std::vector<SensorInfo> sensors;
  std::vector<float> tri_v;     // 9 floats per triangle
  std::vector<float> tri_uv;    // 6 floats per triangle
  std::vector<int32_t> tri_shape;
  std::vector<float> sph;       // center xyz, radius
  std::vector<int32_t> sph_shape;
  std::vector<float> box_xf;    // per box: 3x4 object-to-world map, row-major
  std::vector<int32_t> box_shape;

  // after commit
  bool committed = false;
  void *dev_block = nullptr;
  uint32_t *majorant_keys = nullptr;  // [param][state], order-preserving float keys
  float *supergrids = nullptr;      // sg_block: [medium][state][cell]
  void *sg_block = nullptr;
  int64_t sg_capacity = 0;          // floats
  MediumInfo *media_dev = nullptr;  // device copy of `media`
  int batch_refs = 0;               // batches holding this scene (no relayout then)
  unsigned long long *counters = nullptr;
  unsigned *work_queue = nullptr;  // unit counter of the persistent launches
  float *partial = nullptr;        // per work unit radiance sums of a persistent render
  int64_t partial_cap = 0;         // floats
  int n_nodes = 0;
  float bs_center[3] = {0, 0, 0};
  float bs_radius = 0;
  SceneView view{};

  Scene() = default;
  Scene(const Scene &) = delete;
  Scene &operator=(const Scene &) = delete;
  ~Scene();

  int upload(hipStream_t stream);
  int reduce_param(int param, int state, hipStream_t stream);
  int build_supergrid(int medium, int state, hipStream_t stream);
  int64_t plan_supergrids();
  int alloc_supergrids();
  int relayout_supergrids(hipStream_t stream);
  int features() const;  // kFeat* mask of the committed scene
};

// Decodes an order-preserving key written by the majorant reduction.
__host__ __device__ inline float key_to_float(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __builtin_bit_cast(float, u);
}

}  // namespace rcv

// The opaque C handle is the scene object itself.
struct rcv_scene : public rcv::Scene {};
