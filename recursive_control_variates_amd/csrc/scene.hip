// Scene construction for librcv: the C-ABI object that replaces the scene
// graph mi.load_dict builds for the hot path (shapes, BSDFs incl. the
// 'twostate' adapter, media incl. 'twostatemedium', emitters, batch sensor).
//
// Host side: collects primitives, builds a binned-SAH BVH (16 bins, leaves of
// at most 4 primitives, depth <= kMaxBvhDepth so traversal stacks of
// kStackSize entries never overflow), packs everything into one device block.
// Parameter states are bound by pointer; binding a grid reduces its maximum
// on the device (the majorant of plugins/twostatemedium.py:73-75) with an
// order-independent atomic max, so no host synchronisation is needed.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "scene_internal.hpp"

namespace rcv {

std::string &last_error_slot() {
  static thread_local std::string s;
  return s;
}

namespace {

struct AABB {
  float lo[3] = {kInf, kInf, kInf};
  float hi[3] = {-kInf, -kInf, -kInf};
  void grow(const float *p) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[k]);
      hi[k] = std::max(hi[k], p[k]);
    }
  }
  void grow(const AABB &b) {
    grow(b.lo);
    grow(b.hi);
  }
  float area() const {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
  }
};

struct BuildPrim {
  AABB box;
  float c[3];
  int id;
};

struct Builder {
  std::vector<BuildPrim> &prims;
  std::vector<BVHNode> nodes;

  int max_depth = 0;  // deepest leaf (root: 0), the most entries a traversal stack holds
  explicit Builder(std::vector<BuildPrim> &p) : prims(p) {}

  void make_leaf(int node, int begin, int end, const AABB &box) {
    BVHNode &n = nodes[node];
    for (int k = 0; k < 3; ++k) {
      n.lo[k] = box.lo[k];
      n.hi[k] = box.hi[k];
    }
    n.first = begin;
    n.count = end - begin;
  }

  void build(int node, int begin, int end, int depth) {
    AABB box, cbox;
    for (int i = begin; i < end; ++i) {
      box.grow(prims[i].box);
      cbox.grow(prims[i].c);
    }
    const int count = end - begin;
    if (count <= 4 || depth >= kMaxBvhDepth - 1) {
      if (count > 4 && depth >= kMaxBvhDepth - 1) {
        // Depth limit reached: the leaf keeps all remaining primitives.
      }
      make_leaf(node, begin, end, box);
      max_depth = std::max(max_depth, depth);
      return;
    }
    // Binned SAH over the centroid bounds.
    constexpr int kBins = 16;
    float best_cost = kInf;
    int best_axis = -1, best_split = -1;
    for (int axis = 0; axis < 3; ++axis) {
      const float lo = cbox.lo[axis], hi = cbox.hi[axis];
      if (!(hi > lo)) continue;
      AABB bins[kBins];
      int cnt[kBins] = {0};
      const float scale = kBins / (hi - lo);
      for (int i = begin; i < end; ++i) {
        int b = std::min(kBins - 1, int((prims[i].c[axis] - lo) * scale));
        bins[b].grow(prims[i].box);
        cnt[b]++;
      }
      AABB lacc;
      int lcnt = 0;
      float larea[kBins];
      int lc[kBins];
      for (int b = 0; b < kBins - 1; ++b) {
        lacc.grow(bins[b]);
        lcnt += cnt[b];
        larea[b] = lacc.area();
        lc[b] = lcnt;
      }
      AABB racc;
      int rcnt = 0;
      for (int b = kBins - 1; b > 0; --b) {
        racc.grow(bins[b]);
        rcnt += cnt[b];
        const int split = b - 1;  // left = [0, split]
        if (lc[split] == 0 || rcnt == 0) continue;
        const float cost = larea[split] * lc[split] + racc.area() * rcnt;
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_split = split;
        }
      }
    }
    int mid;
    if (best_axis < 0) {
      // All centroids coincide: split by count.
      mid = begin + count / 2;
    } else {
      const float lo = cbox.lo[best_axis], hi = cbox.hi[best_axis];
      const float scale = kBins / (hi - lo);
      auto it = std::partition(prims.begin() + begin, prims.begin() + end, [&](const BuildPrim &p) {
        int b = std::min(kBins - 1, int((p.c[best_axis] - lo) * scale));
        return b <= best_split;
      });
      mid = int(it - prims.begin());
      if (mid == begin || mid == end) mid = begin + count / 2;
    }
    const int left = int(nodes.size());
    nodes.push_back(BVHNode{});
    nodes.push_back(BVHNode{});
    BVHNode &n = nodes[node];
    for (int k = 0; k < 3; ++k) {
      n.lo[k] = box.lo[k];
      n.hi[k] = box.hi[k];
    }
    n.first = left;
    n.count = 0;
    build(left, begin, mid, depth + 1);
    build(left + 1, mid, end, depth + 1);
  }
};

inline uint32_t float_key(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v) { *p = v; }

__global__ __launch_bounds__(256) void max_reduce_kernel(const float *__restrict__ x, int64_t n,
                                                         uint32_t *__restrict__ out) {
  __shared__ float red[256];
  float m = -kInf;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    m = fmaxf(m, x[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint32_t u = __builtin_bit_cast(uint32_t, red[0]);
    const uint32_t key = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    atomicMax(out, key);
  }
}

}  // namespace

int Scene::upload(hipStream_t stream) {
  // ---- primitives
  const int n_tris = int(tri_v.size() / 9);
  const int n_sph = int(sph.size() / 4);
  const int n_box = int(box_shape.size());
  std::vector<BuildPrim> bp;
  bp.reserve(n_tris + n_sph + n_box);
  for (int t = 0; t < n_tris; ++t) {
    BuildPrim p;
    for (int k = 0; k < 3; ++k) p.box.grow(&tri_v[9 * t + 3 * k]);
    for (int a = 0; a < 3; ++a) p.c[a] = 0.5f * (p.box.lo[a] + p.box.hi[a]);
    p.id = t;
    bp.push_back(p);
  }
  for (int s = 0; s < n_sph; ++s) {
    BuildPrim p;
    const float *c = &sph[4 * s];
    const float r = c[3];
    for (int a = 0; a < 3; ++a) {
      p.box.lo[a] = c[a] - r;
      p.box.hi[a] = c[a] + r;
      p.c[a] = c[a];
    }
    p.id = n_tris + s;
    bp.push_back(p);
  }
  std::vector<BoxPrim> boxes(std::max(1, n_box));
  for (int b = 0; b < n_box; ++b) {
    const float *m = &box_xf[12 * b];
    BuildPrim p;
    for (int c = 0; c < 8; ++c) {
      const float q[3] = {c & 1 ? 1.f : -1.f, c & 2 ? 1.f : -1.f, c & 4 ? 1.f : -1.f};
      float w[3];
      for (int r = 0; r < 3; ++r) w[r] = m[4 * r] * q[0] + m[4 * r + 1] * q[1] + m[4 * r + 2] * q[2] + m[4 * r + 3];
      p.box.grow(w);
    }
    for (int a = 0; a < 3; ++a) p.c[a] = m[4 * a + 3];
    p.id = n_tris + n_sph + b;
    bp.push_back(p);
    // inverse of the linear part (cofactors, in double) and of the offset
    double A[3][3], I[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) A[r][c] = m[4 * r + c];
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
        I[r][c] = (A[r1][c1] * A[r2][c2] - A[r1][c2] * A[r2][c1]) / det;
      }
    BoxPrim &B = boxes[b];
    for (int r = 0; r < 3; ++r) {
      double off = 0;
      for (int c = 0; c < 3; ++c) {
        B.inv[4 * r + c] = float(I[r][c]);
        off -= I[r][c] * m[4 * c + 3];
      }
      B.inv[4 * r + 3] = float(off);
    }
    // outward world normal of face (axis k, sign sg): inverse-transpose image
    // of the object normal, i.e. sg times row k of the inverse
    for (int f = 0; f < 6; ++f) {
      const int k = f / 2;
      const double sg = (f & 1) ? -1.0 : 1.0;
      const double n[3] = {sg * I[k][0], sg * I[k][1], sg * I[k][2]};
      const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int a = 0; a < 3; ++a) B.n[3 * f + a] = float(n[a] / len);
    }
    B.shape = box_shape[b];
    B.pad = 0;
  }
  RCV_REQUIRE(!bp.empty(), "rcv_scene_commit: the scene has no shapes");
  Builder b(bp);
  b.nodes.push_back(BVHNode{});
  b.build(0, 0, int(bp.size()), 0);
  const int bvh_depth = b.max_depth;
  const int n_slots = int(bp.size());

  std::vector<float4> tris(3 * n_slots, make_float4(0, 0, 0, 0));
  std::vector<float2> uvs(3 * n_slots, make_float2(0, 0));
  std::vector<int32_t> prim_ids(n_slots);
  std::vector<int32_t> slot_of(n_slots);
  for (int s = 0; s < n_slots; ++s) {
    const int id = bp[s].id;
    prim_ids[s] = id;
    slot_of[id] = s;
    if (id < n_tris) {
      const float *v = &tri_v[9 * id];
      const int shape = tri_shape[id];
      float sh;
      std::memcpy(&sh, &shape, 4);
      tris[3 * s + 0] = make_float4(v[0], v[1], v[2], sh);
      tris[3 * s + 1] = make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], 0.f);
      tris[3 * s + 2] = make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], 0.f);
      for (int k = 0; k < 3; ++k) uvs[3 * s + k] = make_float2(tri_uv[6 * id + 2 * k], tri_uv[6 * id + 2 * k + 1]);
    }
  }
  std::vector<float4> spheres(std::max(1, n_sph));
  std::vector<int32_t> sphere_shape_v(std::max(1, n_sph), -1);
  for (int s = 0; s < n_sph; ++s) {
    spheres[s] = make_float4(sph[4 * s], sph[4 * s + 1], sph[4 * s + 2], sph[4 * s + 3]);
    sphere_shape_v[s] = sph_shape[s];
  }

  // ---- area emitters: triangle CDF per emitter
  std::vector<float> em_cdf;
  std::vector<int32_t> em_tris;
  for (auto &e : emitters) {
    if (e.type != RCV_EMITTER_AREA) continue;
    e.tri_begin = int(em_tris.size());
    double total = 0;
    std::vector<double> areas;
    for (int t = 0; t < n_tris; ++t) {
      if (tri_shape[t] != e.shape) continue;
      const float *v = &tri_v[9 * t];
      const double e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
      const double e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
      const double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
      const double a = 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      areas.push_back(a);
      em_tris.push_back(slot_of[t]);
      total += a;
    }
    RCV_REQUIRE(total > 0, "area emitter attached to shape %d without area", e.shape);
    double acc = 0;
    for (double a : areas) {
      acc += a;
      em_cdf.push_back(float(acc / total));
    }
    em_cdf.back() = 1.0f;
    e.tri_count = int(areas.size());
    e.inv_area = float(1.0 / total);
  }
  if (em_cdf.empty()) {
    em_cdf.push_back(1.f);
    em_tris.push_back(0);
  }

  // ---- bounding sphere of the geometry (constant-emitter sampling distance)
  AABB all;
  for (auto &p : bp) all.grow(p.box);
  for (int k = 0; k < 3; ++k) bs_center[k] = 0.5f * (all.lo[k] + all.hi[k]);
  {
    const float d[3] = {all.hi[0] - all.lo[0], all.hi[1] - all.lo[1], all.hi[2] - all.lo[2]};
    bs_radius = 0.5f * std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  }

  plan_supergrids();

  // ---- pack one device block
  std::vector<char> blob;
  auto put = [&](const void *src, size_t bytes) {
    size_t off = (blob.size() + 255) & ~size_t(255);
    blob.resize(off + std::max<size_t>(bytes, 16));
    if (bytes) std::memcpy(blob.data() + off, src, bytes);
    return off;
  };
  const size_t o_nodes = put(b.nodes.data(), b.nodes.size() * sizeof(BVHNode));
  const size_t o_tris = put(tris.data(), tris.size() * sizeof(float4));
  const size_t o_uv = put(uvs.data(), uvs.size() * sizeof(float2));
  const size_t o_pid = put(prim_ids.data(), prim_ids.size() * 4);
  const size_t o_sph = put(spheres.data(), spheres.size() * sizeof(float4));
  const size_t o_sphs = put(sphere_shape_v.data(), sphere_shape_v.size() * 4);
  const size_t o_box = put(boxes.data(), boxes.size() * sizeof(BoxPrim));
  const size_t o_shapes = put(shapes.data(), shapes.size() * sizeof(ShapeInfo));
  const size_t o_bsdfs = put(bsdfs.data(), bsdfs.size() * sizeof(BsdfInfo));
  const size_t o_em = put(emitters.data(), emitters.size() * sizeof(EmitterInfo));
  const size_t o_cdf = put(em_cdf.data(), em_cdf.size() * 4);
  const size_t o_emt = put(em_tris.data(), em_tris.size() * 4);
  const size_t o_med = put(media.data(), media.size() * sizeof(MediumInfo));
  const size_t o_sens = put(sensors.data(), sensors.size() * sizeof(SensorInfo));
  std::vector<uint32_t> maj_init(2 * kMaxParams, float_key(-kInf));
  const size_t o_maj = put(maj_init.data(), maj_init.size() * 4);
  std::vector<unsigned long long> cnt(kNumCounters, 0ull);
  const size_t o_cnt = put(cnt.data(), cnt.size() * 8);
  const unsigned q_init[4] = {0u, 0u, 0u, 0u};
  const size_t o_q = put(q_init, sizeof(q_init));

  RCV_TRY_HIP(hipMalloc(&dev_block, blob.size()), "rcv_scene_commit: hipMalloc");
  // unit sums of persistent renders: up to kRenderMaxUnitsPerPixel units of
  // (old, new) RGB per pixel of the largest sensor
  {
    int64_t max_pix = 0;
    for (auto &sen : sensors) max_pix = std::max(max_pix, int64_t(sen.film_w) * sen.film_h);
    partial_cap = max_pix * kRenderMaxUnitsPerPixel * 6;
    if (partial_cap > 0)
      RCV_TRY_HIP(hipMalloc(&partial, size_t(partial_cap) * sizeof(float)), "rcv_scene_commit: hipMalloc");
  }
  RCV_TRY_HIP(hipMemcpyAsync(dev_block, blob.data(), blob.size(), hipMemcpyHostToDevice, stream),
              "rcv_scene_commit: upload");
  RCV_TRY_HIP(hipStreamSynchronize(stream), "rcv_scene_commit: synchronize");

  char *base = static_cast<char *>(dev_block);
  SceneView &v = view;
  v.nodes = reinterpret_cast<const BVHNode *>(base + o_nodes);
  v.tris = reinterpret_cast<const float4 *>(base + o_tris);
  v.tri_uv = reinterpret_cast<const float2 *>(base + o_uv);
  v.prim_ids = reinterpret_cast<const int32_t *>(base + o_pid);
  v.spheres = reinterpret_cast<const float4 *>(base + o_sph);
  v.sphere_shape = reinterpret_cast<const int32_t *>(base + o_sphs);
  v.boxes = reinterpret_cast<const BoxPrim *>(base + o_box);
  v.shapes = reinterpret_cast<const ShapeInfo *>(base + o_shapes);
  v.bsdfs = reinterpret_cast<const BsdfInfo *>(base + o_bsdfs);
  v.emitters = reinterpret_cast<const EmitterInfo *>(base + o_em);
  v.em_cdf = reinterpret_cast<const float *>(base + o_cdf);
  v.em_tris = reinterpret_cast<const int32_t *>(base + o_emt);
  media_dev = reinterpret_cast<MediumInfo *>(base + o_med);
  v.media = media_dev;
  v.sensors = reinterpret_cast<const SensorInfo *>(base + o_sens);
  majorant_keys = reinterpret_cast<uint32_t *>(base + o_maj);
  v.majorants = reinterpret_cast<const float *>(majorant_keys);
  counters = reinterpret_cast<unsigned long long *>(base + o_cnt);
  work_queue = reinterpret_cast<unsigned *>(base + o_q);
  v.counters = nullptr;
  v.n_tris = n_tris;
  v.n_sph = n_sph;
  v.n_emitters = int(emitters.size());
  v.env_emitter = -1;
  for (int i = 0; i < int(emitters.size()); ++i)
    if (emitters[i].type == RCV_EMITTER_CONSTANT) v.env_emitter = i;
  v.n_media = int(media.size());
  v.handle_null = 0;
  v.nee_homog = 0;
  v.hom_medium = -1;
  // a traversal pushes at most one node per level above the current one
  v.lds_stack = std::max(1, bvh_depth);
  v.lds_slots = std::min(kMaxLdsSlots, kLdsWordsPerThread - v.lds_stack);
  for (int i = 0; i < int(media.size()); ++i) {
    if (media[i].homogeneous) {
      v.nee_homog = 1;
      v.hom_medium = i;
    } else {
      v.handle_null = 1;
    }
  }
  for (int k = 0; k < 3; ++k) v.bs_center[k] = bs_center[k];
  v.bs_radius = bs_radius;
  for (int i = 0; i < kMaxParams; ++i) {
    v.params[i] = ParamBind{{nullptr, nullptr}, 0, 0, 0, 0};
    if (i < int(params.size())) {
      v.params[i].nx = params[i].nx;
      v.params[i].ny = params[i].ny;
      v.params[i].nz = params[i].nz;
      v.params[i].ch = params[i].ch;
      v.params[i].data[0] = params[i].data[0];
      v.params[i].data[1] = params[i].data[1];
    }
  }
  n_nodes = int(b.nodes.size());
  if (int rc = alloc_supergrids()) return rc;
  committed = true;
  // Majorants of states bound before commit.
  for (int i = 0; i < int(params.size()); ++i)
    for (int s = 0; s < 2; ++s)
      if (params[i].data[s]) {
        int rc = reduce_param(i, s, stream);
        if (rc != RCV_OK) return rc;
      }
  return RCV_OK;
}

int Scene::reduce_param(int param, int state, hipStream_t stream) {
  const Param &p = params[param];
  bool used_by_medium = false;
  for (auto &m : media) used_by_medium |= (m.sigma_param == param);
  if (!used_by_medium) return RCV_OK;
  const int64_t n = int64_t(p.nx) * p.ny * p.nz * p.ch;
  uint32_t *slot = majorant_keys + 2 * param + state;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(1), 0, stream, slot, float_key(-kInf));
  const unsigned grid = unsigned(std::min<int64_t>(1024, (n + 255) / 256));
  hipLaunchKernelGGL(max_reduce_kernel, dim3(grid), dim3(256), 0, stream, p.data[state], n, slot);
  int rc = check_launch("rcv_scene_bind_param: majorant reduction");
  if (rc != RCV_OK) return rc;
  for (int mi = 0; mi < int(media.size()); ++mi)
    if (media[mi].sigma_param == param && media[mi].majorant_factor > 0) {
      rc = build_supergrid(mi, state, stream);
      if (rc != RCV_OK) return rc;
    }
  return RCV_OK;
}

// Supergrid resolution (ceil(n / factor) cells per axis) and storage offset
// of every medium with a majorant factor; returns the floats needed.
int64_t Scene::plan_supergrids() {
  int64_t total = 0;
  for (auto &m : media) {
    m.supergrid_offset = -1;
    m.sg_res[0] = m.sg_res[1] = m.sg_res[2] = 0;
    if (m.majorant_factor > 0) {
      const Param &p = params[m.sigma_param];
      for (int k = 0; k < 3; ++k) {
        const int n = k == 0 ? p.nx : (k == 1 ? p.ny : p.nz);
        m.sg_res[k] = std::max(1, (n + m.majorant_factor - 1) / m.majorant_factor);
      }
      m.supergrid_offset = int32_t(total);
      total += 2 * int64_t(m.sg_res[0]) * m.sg_res[1] * m.sg_res[2];
    }
  }
  return total;
}

// (Re)allocates the supergrid storage for the current plan (grown only).
int Scene::alloc_supergrids() {
  int64_t total = 0;
  for (auto &m : media)
    if (m.supergrid_offset >= 0) total += 2 * int64_t(m.sg_res[0]) * m.sg_res[1] * m.sg_res[2];
  RCV_REQUIRE(total < (int64_t(1) << 31), "majorant supergrids: %lld cells exceed 2^31", (long long)total);
  if (total > sg_capacity) {
    if (sg_block) RCV_TRY_HIP(hipFree(sg_block), "majorant supergrids: hipFree");
    sg_block = nullptr;
    sg_capacity = 0;
    RCV_TRY_HIP(hipMalloc(&sg_block, size_t(total) * sizeof(float)), "majorant supergrids: hipMalloc");
    sg_capacity = total;
  }
  supergrids = static_cast<float *>(sg_block);
  view.supergrids = supergrids;
  return RCV_OK;
}

// After a change of a grid's resolution or of a majorant factor: new
// supergrid layout, media table re-uploaded, supergrids of bound states
// rebuilt.
int Scene::relayout_supergrids(hipStream_t stream) {
  plan_supergrids();
  if (int rc = alloc_supergrids()) return rc;
  RCV_TRY_HIP(hipMemcpyAsync(media_dev, media.data(), media.size() * sizeof(MediumInfo), hipMemcpyHostToDevice,
                             stream),
              "majorant supergrids: media upload");
  for (int mi = 0; mi < int(media.size()); ++mi) {
    if (media[mi].supergrid_offset < 0) continue;
    for (int s = 0; s < 2; ++s)
      if (params[media[mi].sigma_param].data[s])
        if (int rc = build_supergrid(mi, s, stream)) return rc;
  }
  // the host copy of the media table must outlive the upload
  RCV_TRY_HIP(hipStreamSynchronize(stream), "majorant supergrids: synchronize");
  return RCV_OK;
}

int Scene::features() const {
  int f = 0;
  for (const ShapeInfo &sh : shapes)
    if (sh.bsdf < 0 || bsdfs[sh.bsdf].type != RCV_BSDF_NULL) f |= kFeatSurfaces;
  // the kernels with surface code track NEE shadow rays at once, needing no
  // LDS slots for them: the choice when a deep BVH's stack leaves too few
  if (view.lds_slots < kNeeFlightWords) f |= kFeatSurfaces;
  for (const MediumInfo &m : media) {
    if (m.homogeneous) {
      f |= kFeatHomog;
      continue;
    }
    if (m.albedo_param >= 0) f |= kFeatAlbedoGrid;
    if (m.supergrid_offset >= 0) f |= kFeatSupergrid;
  }
  return f;
}

Scene::~Scene() {
  if (dev_block) (void)hipFree(dev_block);
  if (sg_block) (void)hipFree(sg_block);
  if (partial) (void)hipFree(partial);
}

}  // namespace rcv

using namespace rcv;

extern "C" int rcv_version(void) { return 1; }

extern "C" const char *rcv_last_error(void) { return last_error_slot().c_str(); }

extern "C" int rcv_scene_create(rcv_scene **out) {
  RCV_REQUIRE(out, "rcv_scene_create: null out");
  *out = new rcv_scene();
  return RCV_OK;
}

extern "C" int rcv_scene_destroy(rcv_scene *scene) {
  if (scene && scene->batch_refs > 0)
    return fail(RCV_E_STATE, "rcv_scene_destroy: the scene belongs to %d batch(es); destroy them first",
                scene->batch_refs);
  delete scene;
  return RCV_OK;
}

#define REQUIRE_OPEN(s, fn)                                                       \
  do {                                                                            \
    RCV_REQUIRE((s) != nullptr, fn ": null scene");                               \
    if ((s)->committed) return fail(RCV_E_STATE, fn ": scene already committed"); \
  } while (0)

extern "C" int rcv_scene_add_param(rcv_scene *s, int32_t kind, int32_t nx, int32_t ny, int32_t nz,
                                   int32_t channels, int32_t *out_id) {
  REQUIRE_OPEN(s, "rcv_scene_add_param");
  RCV_REQUIRE(kind == RCV_PARAM_TEXTURE || kind == RCV_PARAM_GRID || kind == RCV_PARAM_VALUE,
              "rcv_scene_add_param: bad kind %d", kind);
  RCV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && channels > 0, "rcv_scene_add_param: bad shape");
  RCV_REQUIRE(kind == RCV_PARAM_GRID || nz == 1, "rcv_scene_add_param: textures have nz == 1");
  RCV_REQUIRE(kind != RCV_PARAM_VALUE || (nx == 1 && ny == 1), "rcv_scene_add_param: a value has nx = ny = nz = 1");
  RCV_REQUIRE(channels == 1 || channels == 3, "rcv_scene_add_param: channels must be 1 or 3");
  RCV_REQUIRE(int(s->params.size()) < kMaxParams, "rcv_scene_add_param: at most %d parameters", kMaxParams);
  Param p;
  p.kind = kind;
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.ch = channels;
  s->params.push_back(p);
  if (out_id) *out_id = int(s->params.size()) - 1;
  return RCV_OK;
}

extern "C" int rcv_scene_bind_param(rcv_scene *s, int32_t param, int32_t state, const float *data,
                                    void *stream) {
  RCV_REQUIRE(s, "rcv_scene_bind_param: null scene");
  RCV_REQUIRE(param >= 0 && param < int(s->params.size()), "rcv_scene_bind_param: bad param id %d", param);
  RCV_REQUIRE(state == RCV_STATE_OLD || state == RCV_STATE_NEW, "rcv_scene_bind_param: bad state %d", state);
  RCV_REQUIRE(data, "rcv_scene_bind_param: null data");
  s->params[param].data[state] = data;
  if (!s->committed) return RCV_OK;
  s->view.params[param].data[state] = data;
  return s->reduce_param(param, state, as_stream(stream));
}

extern "C" int rcv_scene_resize_param(rcv_scene *s, int32_t param, int32_t nx, int32_t ny, int32_t nz,
                                      void *stream) {
  RCV_REQUIRE(s && s->committed, "rcv_scene_resize_param: scene not committed");
  RCV_REQUIRE(param >= 0 && param < int(s->params.size()), "rcv_scene_resize_param: bad param id %d", param);
  if (s->batch_refs > 0) return fail(RCV_E_STATE, "rcv_scene_resize_param: the scene belongs to a batch");
  Param &p = s->params[param];
  RCV_REQUIRE(nx > 0 && ny > 0 && nz > 0, "rcv_scene_resize_param: sizes must be positive");
  RCV_REQUIRE(p.kind != RCV_PARAM_TEXTURE || nz == 1, "rcv_scene_resize_param: a texture has nz = 1");
  RCV_REQUIRE(p.kind != RCV_PARAM_VALUE, "rcv_scene_resize_param: a value parameter has no resolution");
  RCV_REQUIRE(int64_t(nx) * ny * nz * p.ch < (int64_t(1) << 31), "rcv_scene_resize_param: too many elements");
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.data[0] = p.data[1] = nullptr;
  ParamBind &v = s->view.params[param];
  v.nx = nx;
  v.ny = ny;
  v.nz = nz;
  v.data[0] = v.data[1] = nullptr;
  return s->relayout_supergrids(as_stream(stream));
}

extern "C" int rcv_scene_set_majorant_factor(rcv_scene *s, int32_t medium, int32_t factor, void *stream) {
  RCV_REQUIRE(s && s->committed, "rcv_scene_set_majorant_factor: scene not committed");
  RCV_REQUIRE(medium >= 0 && medium < int(s->media.size()), "rcv_scene_set_majorant_factor: bad medium %d",
              medium);
  RCV_REQUIRE(factor >= 0, "rcv_scene_set_majorant_factor: factor must be >= 0");
  if (s->batch_refs > 0) return fail(RCV_E_STATE, "rcv_scene_set_majorant_factor: the scene belongs to a batch");
  RCV_REQUIRE(!s->media[medium].homogeneous || factor == 0,
              "rcv_scene_set_majorant_factor: a homogeneous medium has no majorant supergrid");
  s->media[medium].majorant_factor = factor;
  return s->relayout_supergrids(as_stream(stream));
}

extern "C" int rcv_scene_add_bsdf(rcv_scene *s, int32_t type, const float rgb[3], int32_t texture_param,
                                  int32_t *out_id) {
  REQUIRE_OPEN(s, "rcv_scene_add_bsdf");
  RCV_REQUIRE(type == RCV_BSDF_DIFFUSE || type == RCV_BSDF_NULL, "rcv_scene_add_bsdf: bad type %d", type);
  BsdfInfo b{};
  b.type = type;
  b.texture_param = texture_param;
  if (texture_param >= 0) {
    RCV_REQUIRE(texture_param < int(s->params.size()) && s->params[texture_param].kind == RCV_PARAM_TEXTURE &&
                    s->params[texture_param].ch == 3,
                "rcv_scene_add_bsdf: texture param %d must be an RGB texture", texture_param);
  }
  for (int k = 0; k < 3; ++k) b.rgb[k] = rgb ? rgb[k] : 0.f;
  s->bsdfs.push_back(b);
  if (out_id) *out_id = int(s->bsdfs.size()) - 1;
  return RCV_OK;
}

extern "C" int rcv_scene_add_medium(rcv_scene *s, const rcv_medium_desc *d, int32_t *out_id) {
  REQUIRE_OPEN(s, "rcv_scene_add_medium");
  RCV_REQUIRE(d, "rcv_scene_add_medium: null desc");
  const bool hom = d->homogeneous != 0;
  const int pk = hom ? RCV_PARAM_VALUE : RCV_PARAM_GRID;
  RCV_REQUIRE(d->sigma_t_param >= 0 && d->sigma_t_param < int(s->params.size()) &&
                  s->params[d->sigma_t_param].kind == pk && s->params[d->sigma_t_param].ch == 1,
              "rcv_scene_add_medium: sigma_t must be a 1-channel %s parameter", hom ? "value" : "grid");
  if (d->albedo_param >= 0)
    RCV_REQUIRE(d->albedo_param < int(s->params.size()) && s->params[d->albedo_param].kind == pk &&
                    s->params[d->albedo_param].ch == 3,
                "rcv_scene_add_medium: albedo must be a 3-channel %s parameter", hom ? "value" : "grid");
  RCV_REQUIRE(d->majorant_factor >= 0, "rcv_scene_add_medium: majorant_factor must be >= 0");
  if (hom) {
    RCV_REQUIRE(d->majorant_factor == 0, "rcv_scene_add_medium: a homogeneous medium has no majorant supergrid");
    for (const MediumInfo &o : s->media)
      RCV_REQUIRE(!o.homogeneous, "rcv_scene_add_medium: at most one homogeneous medium per scene");
  } else {
    for (int k = 0; k < 3; ++k)
      RCV_REQUIRE(d->bbox_max[k] > d->bbox_min[k], "rcv_scene_add_medium: empty bbox");
  }
  MediumInfo m{};
  m.homogeneous = hom ? 1 : 0;
  m.sigma_param = d->sigma_t_param;
  m.albedo_param = d->albedo_param;
  for (int k = 0; k < 3; ++k) {
    m.albedo[k] = d->albedo_rgb[k];
    m.lo[k] = d->bbox_min[k];
    m.hi[k] = d->bbox_max[k];
  }
  m.scale = d->scale;
  m.sample_emitters = d->sample_emitters;
  m.majorant_factor = d->majorant_factor;
  m.supergrid_offset = -1;
  s->media.push_back(m);
  if (out_id) *out_id = int(s->media.size()) - 1;
  return RCV_OK;
}

extern "C" int rcv_scene_add_emitter(rcv_scene *s, int32_t type, const float rgb[3], int32_t *out_id) {
  REQUIRE_OPEN(s, "rcv_scene_add_emitter");
  RCV_REQUIRE(type == RCV_EMITTER_AREA || type == RCV_EMITTER_CONSTANT, "rcv_scene_add_emitter: bad type %d",
              type);
  RCV_REQUIRE(rgb, "rcv_scene_add_emitter: null radiance");
  if (type == RCV_EMITTER_CONSTANT)
    for (auto &e : s->emitters)
      RCV_REQUIRE(e.type != RCV_EMITTER_CONSTANT, "rcv_scene_add_emitter: only one constant emitter");
  EmitterInfo e{};
  e.type = type;
  e.shape = -1;
  for (int k = 0; k < 3; ++k) e.rgb[k] = rgb[k];
  s->emitters.push_back(e);
  if (out_id) *out_id = int(s->emitters.size()) - 1;
  return RCV_OK;
}

static int check_refs(rcv_scene *s, int32_t bsdf, int32_t emitter, int32_t interior, int32_t exterior,
                      const char *fn) {
  RCV_REQUIRE(bsdf >= -1 && bsdf < int(s->bsdfs.size()), "%s: bad bsdf id %d", fn, bsdf);
  RCV_REQUIRE(emitter >= -1 && emitter < int(s->emitters.size()), "%s: bad emitter id %d", fn, emitter);
  RCV_REQUIRE(interior >= -1 && interior < int(s->media.size()), "%s: bad interior medium %d", fn, interior);
  RCV_REQUIRE(exterior >= -1 && exterior < int(s->media.size()), "%s: bad exterior medium %d", fn, exterior);
  if (emitter >= 0) {
    RCV_REQUIRE(s->emitters[emitter].type == RCV_EMITTER_AREA, "%s: emitter %d is not an area emitter", fn, emitter);
    RCV_REQUIRE(s->emitters[emitter].shape < 0, "%s: emitter %d already attached", fn, emitter);
  }
  return RCV_OK;
}

extern "C" int rcv_scene_add_mesh(rcv_scene *s, const float *vertices, const float *uv, int32_t n_vertices,
                                  const int32_t *triangles, int32_t n_triangles, int32_t bsdf, int32_t emitter,
                                  int32_t interior, int32_t exterior, int32_t *out_id) {
  REQUIRE_OPEN(s, "rcv_scene_add_mesh");
  RCV_REQUIRE(vertices && triangles && n_vertices > 0 && n_triangles > 0, "rcv_scene_add_mesh: empty mesh");
  int rc = check_refs(s, bsdf, emitter, interior, exterior, "rcv_scene_add_mesh");
  if (rc) return rc;
  const int shape = int(s->shapes.size());
  for (int t = 0; t < n_triangles; ++t) {
    for (int k = 0; k < 3; ++k) {
      const int vi = triangles[3 * t + k];
      RCV_REQUIRE(vi >= 0 && vi < n_vertices, "rcv_scene_add_mesh: triangle %d index %d out of range", t, vi);
      for (int a = 0; a < 3; ++a) s->tri_v.push_back(vertices[3 * vi + a]);
      s->tri_uv.push_back(uv ? uv[2 * vi] : 0.f);
      s->tri_uv.push_back(uv ? uv[2 * vi + 1] : 0.f);
    }
    s->tri_shape.push_back(shape);
  }
  s->shapes.push_back(ShapeInfo{bsdf, emitter, interior, exterior});
  if (emitter >= 0) s->emitters[emitter].shape = shape;
  if (out_id) *out_id = shape;
  return RCV_OK;
}

extern "C" int rcv_scene_add_sphere(rcv_scene *s, const float center[3], float radius, int32_t bsdf,
                                    int32_t interior, int32_t exterior, int32_t *out_id) {
  REQUIRE_OPEN(s, "rcv_scene_add_sphere");
  RCV_REQUIRE(center && radius > 0.f, "rcv_scene_add_sphere: bad sphere");
  int rc = check_refs(s, bsdf, -1, interior, exterior, "rcv_scene_add_sphere");
  if (rc) return rc;
  const int shape = int(s->shapes.size());
  for (int k = 0; k < 3; ++k) s->sph.push_back(center[k]);
  s->sph.push_back(radius);
  s->sph_shape.push_back(shape);
  s->shapes.push_back(ShapeInfo{bsdf, -1, interior, exterior});
  if (out_id) *out_id = shape;
  return RCV_OK;
}

extern "C" int rcv_scene_add_box(rcv_scene *s, const float to_world[12], int32_t bsdf, int32_t interior,
                                 int32_t exterior, int32_t *out_id) {
  REQUIRE_OPEN(s, "rcv_scene_add_box");
  RCV_REQUIRE(to_world, "rcv_scene_add_box: null transform");
  for (int i = 0; i < 12; ++i) RCV_REQUIRE(std::isfinite(to_world[i]), "rcv_scene_add_box: non-finite transform");
  const float *m = to_world;
  const double det = double(m[0]) * (double(m[5]) * m[10] - double(m[6]) * m[9]) -
                     double(m[1]) * (double(m[4]) * m[10] - double(m[6]) * m[8]) +
                     double(m[2]) * (double(m[4]) * m[9] - double(m[5]) * m[8]);
  RCV_REQUIRE(det != 0.0, "rcv_scene_add_box: singular transform");
  int rc = check_refs(s, bsdf, -1, interior, exterior, "rcv_scene_add_box");
  if (rc) return rc;
  const int shape = int(s->shapes.size());
  s->box_xf.insert(s->box_xf.end(), to_world, to_world + 12);
  s->box_shape.push_back(shape);
  s->shapes.push_back(ShapeInfo{bsdf, -1, interior, exterior});
  if (out_id) *out_id = shape;
  return RCV_OK;
}

extern "C" int rcv_scene_add_sensor(rcv_scene *s, int32_t film_width, int32_t film_height, int32_t n_views,
                                    const float *to_world, const float *fov_x, int32_t *out_id) {
  REQUIRE_OPEN(s, "rcv_scene_add_sensor");
  RCV_REQUIRE(to_world && fov_x, "rcv_scene_add_sensor: null transform");
  RCV_REQUIRE(n_views >= 1 && n_views <= kMaxViews, "rcv_scene_add_sensor: 1..%d views", kMaxViews);
  RCV_REQUIRE(film_width > 0 && film_height > 0 && film_width % n_views == 0,
              "rcv_scene_add_sensor: film width %d must be a positive multiple of n_views %d", film_width, n_views);
  SensorInfo si{};
  si.film_w = film_width;
  si.film_h = film_height;
  si.n_views = n_views;
  for (int v = 0; v < n_views; ++v) {
    const float *m = to_world + 16 * v;  // row-major: column k = (m[k], m[4+k], m[8+k])
    ViewInfo &vi = si.views[v];
    for (int r = 0; r < 3; ++r) {
      vi.left[r] = m[4 * r + 0];
      vi.up[r] = m[4 * r + 1];
      vi.fwd[r] = m[4 * r + 2];
      vi.origin[r] = m[4 * r + 3];
    }
    RCV_REQUIRE(fov_x[v] > 0.f && fov_x[v] < 180.f, "rcv_scene_add_sensor: fov out of range");
    vi.tan_half = std::tan(double(fov_x[v]) * 0.5 * M_PI / 180.0);
  }
  s->sensors.push_back(si);
  if (out_id) *out_id = int(s->sensors.size()) - 1;
  return RCV_OK;
}

extern "C" int rcv_scene_commit(rcv_scene *s, void *stream) {
  REQUIRE_OPEN(s, "rcv_scene_commit");
  RCV_REQUIRE(!s->sensors.empty(), "rcv_scene_commit: the scene has no sensor");
  for (auto &e : s->emitters)
    if (e.type == RCV_EMITTER_AREA) RCV_REQUIRE(e.shape >= 0, "rcv_scene_commit: area emitter without a shape");
  return s->upload(as_stream(stream));
}

extern "C" int rcv_scene_check_twostate(const rcv_scene *s, int32_t integrator_type, const int32_t *bsdf_twostate,
                                        int32_t n_bsdf, const int32_t *medium_twostate, int32_t n_media) {
  RCV_REQUIRE(s, "rcv_scene_check_twostate: null scene");
  bool any_bsdf = false, any_medium = false;
  for (int i = 0; i < n_bsdf; ++i) any_bsdf |= bsdf_twostate && bsdf_twostate[i];
  for (int i = 0; i < n_media; ++i) any_medium |= medium_twostate && medium_twostate[i];
  if (integrator_type == RCV_INTEGRATOR_PATH) {
    // plugins/twostatepath.py:158-166
    if (any_medium) return fail(RCV_E_SCENE, "TwoStateMedium is not supported in two state prb integrator!");
    if (!any_bsdf) return fail(RCV_E_SCENE, "No TwoStateBSDF found in the scene!");
  } else {
    // plugins/twostatevolpath.py:73-78
    if (!any_medium) return fail(RCV_E_SCENE, "No TwoStateMedium found in the scene!");
  }
  return RCV_OK;
}

extern "C" int rcv_scene_set_counting(rcv_scene *s, int32_t enabled) {
  RCV_REQUIRE(s && s->committed, "rcv_scene_set_counting: scene not committed");
  s->view.counters = enabled ? s->counters : nullptr;
  return RCV_OK;
}

extern "C" int rcv_scene_majorant_grid(const rcv_scene *s, int32_t medium, int32_t state, float *out,
                                       int64_t capacity, int32_t *res, void *stream) {
  RCV_REQUIRE(s && s->committed, "rcv_scene_majorant_grid: scene not committed");
  RCV_REQUIRE(res, "rcv_scene_majorant_grid: null res");
  RCV_REQUIRE(medium >= 0 && medium < int(s->media.size()), "rcv_scene_majorant_grid: bad medium %d", medium);
  RCV_REQUIRE(state == RCV_STATE_OLD || state == RCV_STATE_NEW, "rcv_scene_majorant_grid: bad state");
  const MediumInfo &m = s->media[medium];
  if (m.supergrid_offset < 0) {
    res[0] = res[1] = res[2] = 0;
    return RCV_OK;
  }
  const int64_t cells = int64_t(m.sg_res[0]) * m.sg_res[1] * m.sg_res[2];
  for (int k = 0; k < 3; ++k) res[k] = m.sg_res[k];
  if (!out) return RCV_OK;
  RCV_REQUIRE(capacity >= cells, "rcv_scene_majorant_grid: capacity %lld < %lld cells", (long long)capacity,
              (long long)cells);
  RCV_TRY_HIP(hipMemcpyAsync(out, s->supergrids + m.supergrid_offset + state * cells, cells * sizeof(float),
                             hipMemcpyDeviceToHost, as_stream(stream)),
              "rcv_scene_majorant_grid");
  RCV_TRY_HIP(hipStreamSynchronize(as_stream(stream)), "rcv_scene_majorant_grid: synchronize");
  return RCV_OK;
}

extern "C" int rcv_scene_read_counters(const rcv_scene *s, int64_t *out, int32_t n, void *stream) {
  RCV_REQUIRE(s && s->committed, "rcv_scene_read_counters: scene not committed");
  RCV_REQUIRE(out && n > 0 && n <= kNumCounters, "rcv_scene_read_counters: 1..%d counters", kNumCounters);
  unsigned long long tmp[kNumCounters];
  RCV_TRY_HIP(hipMemcpyAsync(tmp, s->counters, sizeof(tmp), hipMemcpyDeviceToHost, as_stream(stream)),
              "rcv_scene_read_counters");
  RCV_TRY_HIP(hipStreamSynchronize(as_stream(stream)), "rcv_scene_read_counters: synchronize");
  for (int i = 0; i < n; ++i) out[i] = int64_t(tmp[i]);
  return RCV_OK;
}
