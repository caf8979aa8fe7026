// Two-state volumetric path tracer (K3), its single-state form (the plain
// prbvolpath render used by the baseline method) and the path-replay adjoint
// (K5) on the NEW state.
//
// Follows plugins/twostatevolpath.py:141-479 and plugins/twostatemedium.py:
// 15-107 lane for lane: random numbers are drawn in the same order and only by
// the lanes the reference draws them for, weights are formed with the same
// expressions, and the loop-mask semantics (e.g. emitter hits are still
// gathered in the iteration where depth reaches max_depth, :253,284-308) are
// kept. NS = 2 carries (old, new) throughputs along one path; NS = 1 uses the
// NEW state only and reproduces the NS = 2 result of a scene whose two states
// are equal bit for bit (2a/(a+a) == 1 and (x+x)*0.5 == x in IEEE arithmetic).
//
// Extinction is grey (1-channel sigma_t grid times `scale`, as in the
// reference's experiments where spectral extinction is disabled), so the
// colour channel the reference samples for free-flight distances
// (:171-173) does not change any result; its random number is still drawn.
#pragma once

#include <type_traits>

#include "scene_device.hpp"

namespace rcv {

struct IntegArgs {
  int max_depth;
  int rr_depth;
  int hide_emitters;
  int flags;  // integrator-specific switches (volpathsimple: kSimpleNo*)
};

// Scatter log geometry: reduction tiles of a logged gradient, bins (one per
// tile, then one per seam between tiles t and t + 1), records per chunk.
constexpr int kLogMaxTiles = 16;
constexpr int kLogBins = 2 * kLogMaxTiles;
constexpr unsigned kLogChunk = 256;

struct GradTargets {
  float *g[kMaxParams];
  // Scatter log (the persistent volume adjoint): the replay pass appends
  // each scatter-add into g[log_param] as a record in log_rec (log_record)
  // instead of adding it. Records
  // are binned by the reduction tile (2^log_shift voxels) their corners
  // fall in: bin t for tile t, bin kLogMaxTiles + t for footprints across
  // tiles t and t + 1. A wavefront fills chunks of kLogChunk records, one
  // per bin at a time, drawn from a pool of log_chunks: log_ctl[0] counts
  // the chunks drawn, log_ctl[1 + b] those of bin b, listed in
  // log_list[b * log_chunks + k]; log_fill[c] is chunk c's record count.
  // log_reduce_kernel adds the records up. Null log_rec: scatter-adds are
  // atomics.
  int4 *log_rec;
  unsigned *log_ctl;
  unsigned *log_list;
  unsigned *log_fill;
  unsigned log_chunks;
  int log_shift;
  int log_param;
  __host__ __device__ void clear() {
    for (int i = 0; i < kMaxParams; ++i) g[i] = nullptr;
    log_rec = nullptr;
    log_ctl = nullptr;
    log_list = nullptr;
    log_fill = nullptr;
    log_chunks = 0;
    log_shift = 0;
    log_param = -1;
  }
};

// Event counters of a launch (scene counting mode): each call adds n per
// active lane of the wavefront with one atomic from its first active lane.
// With a null pointer (the timed launches) every call is a no-op, so counting
// costs no registers there.
struct LocalCounters {
  unsigned long long *g;
  __device__ void add(int k, unsigned n) const {
    if (!g) return;
    const uint64_t m = __ballot(1);
    if (__lane_id() == unsigned(__ffsll((unsigned long long)m) - 1))
      atomicAdd(g + k, (unsigned long long)(__popcll(m) * n));
  }
};

template <int NS>
__device__ constexpr int state_index(int k) {
  return NS == 2 ? k : RCV_STATE_NEW;
}

template <int NS>
struct MediumEv {
  bool valid;
  float t, mint, maj;
  V3 p;
  float st[NS];  // sigma_t
  float sn[NS];  // sigma_n = majorant - sigma_t
  V3 ss[NS];     // sigma_s = albedo * sigma_t
  V3 alb;        // NEW-state albedo (adjoint)
  GridFoot foot;
};

__device__ inline float majorant_of(const SceneView &S, int param, int state) {
  return key_to_float(reinterpret_cast<const uint32_t *>(S.majorants)[2 * param + state]);
}

// Ray/box slab test returning the parametric interval (may start behind the
// origin).
__device__ inline bool ray_aabb(const Ray &r, const float lo[3], const float hi[3], float &t0, float &t1) {
  t0 = -kInf;
  t1 = kInf;
  const float o[3] = {r.o.x, r.o.y, r.o.z};
  const float d[3] = {r.d.x, r.d.y, r.d.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float inv = 1.0f / d[k];
    float ta = (lo[k] - o[k]) * inv;
    float tb = (hi[k] - o[k]) * inv;
    if (ta > tb) {
      const float tmp = ta;
      ta = tb;
      tb = tmp;
    }
    t0 = ta > t0 ? ta : t0;
    t1 = tb < t1 ? tb : t1;
  }
  return t0 <= t1;
}

__device__ inline int floor_div_i(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Voxel range [i0, i1] that trilinear lookups anywhere in supergrid cell c of
// s cells along an axis of n voxels read from: the cell spans voxel
// coordinates [c n / s, (c + 1) n / s] and a lookup at u reads voxels
// floor(u - 1/2) and floor(u - 1/2) + 1 (clamped to the grid).
__device__ inline void supergrid_cell_voxels(int c, int n, int s, int &i0, int &i1) {
  i0 = max(floor_div_i(2 * c * n - s, 2 * s), 0);
  i1 = min(floor_div_i(2 * (c + 1) * n - s, 2 * s) + 1, n - 1);
}

// Majorant of supergrid cell `c` (x fastest) of an (sx, sy, sz) supergrid
// over an (nx, ny, nz) grid: the maximum over every voxel a point of the cell
// can interpolate from, so it bounds the interpolated grid in the cell.
__device__ inline float supergrid_cell_max(const float *__restrict__ grid, int nx, int ny, int nz, int sx, int sy,
                                           int sz, int c) {
  const int cx = c % sx, cy = (c / sx) % sy, cz = c / (sx * sy);
  int x0, x1, y0, y1, z0, z1;
  supergrid_cell_voxels(cx, nx, sx, x0, x1);
  supergrid_cell_voxels(cy, ny, sy, y0, y1);
  supergrid_cell_voxels(cz, nz, sz, z0, z1);
  float m = -kInf;
  for (int z = z0; z <= z1; ++z)
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) m = fmaxf(m, grid[(int64_t(z) * ny + y) * nx + x]);
  return m;
}

// Supergrid cell of a point given in cell space (unit cells), clamped to the
// grid; the product of the cell's majorant over the states in flight and the
// medium scale. NS == 1 reads the NEW state only.
template <int NS>
__device__ inline float supergrid_at(const SceneView &S, const MediumInfo &M, V3 q) {
  const int sx = M.sg_res[0], sy = M.sg_res[1], sz = M.sg_res[2];
  const int cx = int(fminf(fmaxf(floorf(q.x), 0.0f), float(sx - 1)));
  const int cy = int(fminf(fmaxf(floorf(q.y), 0.0f), float(sy - 1)));
  const int cz = int(fminf(fmaxf(floorf(q.z), 0.0f), float(sz - 1)));
  const int cells = sx * sy * sz;
  const float *g = S.supergrids + M.supergrid_offset + (cz * sy + cy) * sx + cx;
  float m = g[RCV_STATE_NEW * cells];
  if (NS == 2) m = fmaxf(m, g[RCV_STATE_OLD * cells]);
  return m * M.scale;
}

// Free-flight distance through a piecewise-constant majorant: a 3-D DDA
// (Amanatides & Woo 1987, "A Fast Voxel Traversal Algorithm") over the
// supergrid cells the ray crosses in [mint, maxt], accumulating optical depth
// until it reaches tau (plugins/twostatemedium.py:25-70). The ray is carried
// in cell space (each supergrid cell a unit cube), where its parameter t is
// the world ray's. Returns kInf when the ray leaves the medium first; sets
// `maj` to the majorant of the cell the collision lies in (:82-85).
template <int NS>
__device__ inline float supergrid_free_flight(const SceneView &S, const MediumInfo &M, const Ray &ray, float mint,
                                              float maxt, float tau, float &maj) {
  const float kx = float(M.sg_res[0]) / (M.hi[0] - M.lo[0]);
  const float ky = float(M.sg_res[1]) / (M.hi[1] - M.lo[1]);
  const float kz = float(M.sg_res[2]) / (M.hi[2] - M.lo[2]);
  const V3 O = v3((ray.o.x - M.lo[0]) * kx, (ray.o.y - M.lo[1]) * ky, (ray.o.z - M.lo[2]) * kz);
  const V3 D = v3(ray.d.x * kx, ray.d.y * ky, ray.d.z * kz);
  // Per axis: the parameter of the next cell face the ray meets (the upper
  // face of the entry cell for an increasing coordinate, its lower face for a
  // decreasing one) and the parameter step between faces. Every segment
  // [t, min(next faces)] then lies in one cell, so the majorant looked up at
  // its midpoint is the cell's.
  float tm[3], td[3];
  const float o3[3] = {O.x, O.y, O.z}, d3[3] = {D.x, D.y, D.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float res = float(M.sg_res[k]);
    const float cur = fminf(fmaxf(floorf(fmaf(mint, d3[k], o3[k])), 0.0f), res - 1.0f);
    const float step = d3[k] >= 0.0f ? 1.0f : -1.0f;
    const float nb = d3[k] >= 0.0f ? cur + 1.0f : cur;
    const bool moving = d3[k] != 0.0f;
    tm[k] = moving ? (nb - o3[k]) / d3[k] : kInf;
    td[k] = moving ? step / d3[k] : kInf;
  }
  float t = mint, tau_acc = 0.0f;
  bool reached = false, active = true;
  // A ray crosses at most sx + sy + sz faces inside the grid; the bound only
  // guards against non-finite input.
  const int guard = M.sg_res[0] + M.sg_res[1] + M.sg_res[2] + 4;
  for (int it = 0; active && it < guard; ++it) {
    const float t_next = fminf(fminf(tm[0], tm[1]), tm[2]);
    const bool a0 = tm[0] == t_next, a1 = !a0 && tm[1] == t_next, a2 = !a0 && !a1;
    const float th = 0.5f * (t + t_next);
    const float m = supergrid_at<NS>(S, M, v3(fmaf(th, D.x, O.x), fmaf(th, D.y, O.y), fmaf(th, D.z, O.z)));
    const float tau_next = tau_acc + m * (t_next - t);
    const float t_precise = t + (tau - tau_acc) / m;
    reached = m > 0.0f && t_precise < maxt && tau_next >= tau;
    t = reached ? t_precise : t_next;
    active = !reached && t_next < maxt;
    if (active) {
      tm[0] += a0 ? td[0] : 0.0f;
      tm[1] += a1 ? td[1] : 0.0f;
      tm[2] += a2 ? td[2] : 0.0f;
      tau_acc = tau_next;
    }
  }
  const float ts = reached ? t : kInf;
  maj = ts <= maxt ? supergrid_at<NS>(S, M, v3(fmaf(ts, D.x, O.x), fmaf(ts, D.y, O.y), fmaf(ts, D.z, O.z))) : 0.0f;
  return ts;
}

// Optical depth of a free flight, -log(1 - u) (plugins/twostatemedium.py:24).
__device__ inline float tau_of(float u) { return -logf(1.0f - u); }

// TwoStateMedium.sample_interaction_twostates (plugins/twostatemedium.py:
// 15-107): one global majorant max(maj_old, maj_new) (:71-78), or the
// supergrid DDA when the medium has a majorant supergrid (:25-70, 82-85).
template <int NS, int F>
__device__ inline void sample_medium(const SceneView &S, int mid, const Ray &ray, float u, MediumEv<NS> &ev) {
  const MediumInfo &M = S.media[mid];
  if ((F & kFeatHomog) && M.homogeneous) {
    // A homogeneous medium is unbounded (its interaction interval is [0,
    // ray.maxt]); its majorant is sigma_t itself, combined over the states
    // (:71-78), and its coefficients are constants.
    const ParamBind &P = S.params[M.sigma_param];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int s = state_index<NS>(k);
      ev.st[k] = M.scale * P.data[s][0];
      V3 alb;
      if (M.albedo_param >= 0) {
        const float *a = S.params[M.albedo_param].data[s];
        alb = v3(a[0], a[1], a[2]);
      } else {
        alb = v3(M.albedo[0], M.albedo[1], M.albedo[2]);
      }
      if (k == NS - 1) ev.alb = alb;
      ev.ss[k] = alb * ev.st[k];
    }
    const float maj = NS == 2 ? fmaxf(ev.st[0], ev.st[1]) : ev.st[0];
    const float t = tau_of(u) / maj;
    ev.valid = maj > 0.0f && maj < kInf && t <= ray.maxt;
    ev.mint = 0.0f;
    ev.maj = maj;
    ev.t = ev.valid ? t : kInf;
    ev.p = ray_at(ray, t);
#pragma unroll
    for (int k = 0; k < NS; ++k) ev.sn[k] = maj - ev.st[k];
    return;
  }
  float t0, t1;
  bool ok = ray_aabb(ray, M.lo, M.hi, t0, t1);
  ok = ok && (t0 < kInf || t1 < kInf);
  const float mint = ok ? fmaxf(0.0f, t0) : 0.0f;
  const float maxt = ok ? fminf(ray.maxt, t1) : kInf;
  const float scale = M.scale;
  const float tau = tau_of(u);
  float maj = 0.0f, t = kInf;
  if ((F & kFeatSupergrid) && M.supergrid_offset >= 0) {
    if (ok) t = supergrid_free_flight<NS>(S, M, ray, mint, maxt, tau, maj);
    // A collision is kept only where its cell's majorant is positive and
    // finite, the invariant of the single-majorant branch below.
    ok = ok && maj > 0.0f && maj < kInf;
  } else {
    maj = majorant_of(S, M.sigma_param, RCV_STATE_NEW);
    if (NS == 2) maj = fmaxf(maj, majorant_of(S, M.sigma_param, RCV_STATE_OLD));
    maj *= scale;
    // A medium whose majorant is not a positive finite number has no
    // collisions (t = inf for a zero majorant, as in the reference); this
    // also keeps every sampled step strictly forward.
    ok = ok && maj > 0.0f && maj < kInf;
    t = mint + tau / maj;
  }
  ev.valid = ok && (t <= maxt);
  ev.mint = mint;
  ev.maj = maj;
  ev.t = ev.valid ? t : kInf;
  ev.p = ray_at(ray, t);
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    ev.st[k] = 0.0f;
    ev.ss[k] = v3(0.f);
  }
  ev.alb = v3(0.f);
  if (ev.valid) {
    const ParamBind &P = S.params[M.sigma_param];
    grid_footprint(P, M, ev.p, ev.foot);
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int s = state_index<NS>(k);
      ev.st[k] = scale * grid_eval1(P.data[s], ev.foot);
      V3 alb;
      if ((F & kFeatAlbedoGrid) && M.albedo_param >= 0) {
        const ParamBind &A = S.params[M.albedo_param];
        GridFoot fa;
        grid_footprint(A, M, ev.p, fa);
        alb = grid_eval3(A.data[s], fa);
      } else {
        alb = v3(M.albedo[0], M.albedo[1], M.albedo[2]);
      }
      if (k == NS - 1) ev.alb = alb;
      ev.ss[k] = alb * ev.st[k];
    }
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) ev.sn[k] = maj - ev.st[k];
}

__device__ inline void scatter_grid(float *g, const GridFoot &f, float val) {
#pragma unroll
  for (int c = 0; c < 8; ++c) atomicAdd(g + f.idx(c), val * f.w(c));
}

__device__ inline void scatter_grid3(float *g, const GridFoot &f, V3 val) {
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float w = f.w(c);
    float *t = g + 3 * f.idx(c);
    atomicAdd(t + 0, val.x * w);
    atomicAdd(t + 1, val.y * w);
    atomicAdd(t + 2, val.z * w);
  }
}

__device__ inline bool null_bsdf(const SceneView &S, int shape) {
  const int b = S.shapes[shape].bsdf;
  return b >= 0 && S.bsdfs[b].type == RCV_BSDF_NULL;
}

// A recorded collision keeps the trilinear footprint of its density lookup
// instead of its point, so that reading it back needs no world-to-grid
// transform and no load of its medium's description: the three fractions
// and the collision's gradient coefficient in a float4, and one word packing
// the base voxel (kRecBaseBits bits), the +1 flag of each axis (clear where
// the footprint is clamped) and the medium. A footprint that does not pack
// (a grid of more than 2^26 voxels, a medium index above 7) overflows the
// record, and the shadow ray or path is tracked again instead.
constexpr int kRecBaseBits = 26;
constexpr unsigned kRecBaseMask = (1u << kRecBaseBits) - 1u;

__device__ inline bool pack_cell(const GridFoot &f, int medium, unsigned &cell) {
  if (unsigned(f.base) > kRecBaseMask || unsigned(medium) > 7u) return false;
  cell = unsigned(f.base) | (f.dx != 0 ? 1u << 26 : 0u) | (f.dy != 0 ? 1u << 27 : 0u) |
         (f.dz != 0 ? 1u << 28 : 0u) | (unsigned(medium) << 29);
  return true;
}

__device__ inline int cell_medium(unsigned cell) { return int(cell >> 29); }

// A scatter log record, 16 B: the packed footprint (pack_cell), the three
// fractions as 21-bit fixed point (round((2^21 - 1) t); a fraction lies in
// [0, 1]) in 63 bits, and the value.
constexpr float kLogFracScale = 2097151.0f;

__device__ inline int4 log_record(unsigned cell, const GridFoot &f, float val) {
  const unsigned long long q = (unsigned long long)__float2uint_rn(f.tx * kLogFracScale) |
                               ((unsigned long long)__float2uint_rn(f.ty * kLogFracScale) << 21) |
                               ((unsigned long long)__float2uint_rn(f.tz * kLogFracScale) << 42);
  return make_int4(int(cell), int(unsigned(q)), int(unsigned(q >> 32)), __float_as_int(val));
}

// Scatter log state of the lane's wavefront, in LDS word
// kLdsWordsPerThread - 1 of its lanes (free in every scene: past the stack,
// and past the NEE weights, which leave the last slot free): lane b holds
// the chunk the wavefront fills for bin b (-1: none), lane kLogBins + b
// that chunk's record count. Only the wavefront itself touches them.
static_assert(2 * kLogBins == 64, "scatter log state: two words per bin in a wavefront's 64 lanes");

__device__ inline int *wave_log_word(const LdsStack &L, int lane) {
  return L.base + (kLdsWordsPerThread - 1) * L.stride - int(threadIdx.x & 63u) + lane;
}

// Before the wavefront's first scatter: no chunk in any bin.
__device__ inline void log_begin(const GradTargets &G, const LdsStack &L) {
  const int lane = int(threadIdx.x & 63u);
  if (G.log_rec) *wave_log_word(L, lane) = lane < kLogBins ? -1 : 0;
}

// With the whole wavefront past its last scatter: the record counts of the
// chunks it still fills.
__device__ inline void log_end(const GradTargets &G, const LdsStack &L) {
  const int lane = int(threadIdx.x & 63u);
  if (G.log_rec && lane < kLogBins) {
    const int c = *wave_log_word(L, lane);
    if (c >= 0) G.log_fill[c] = unsigned(*wave_log_word(L, kLogBins + lane));
  }
}

// Scatter-add of val over footprint f into the gradient of param. When
// that gradient is logged the record goes to its bin's current chunk, at
// the slot an LDS atomic on the bin's count gives the lane. Lanes left
// without a slot (the chunk is full, or the bin has none yet) are rare:
// for the bin of the first of them the chunk is closed (its count stored)
// and the next one drawn from the pool with one global atomic, and those
// of any other bin add with atomics this once. Added with atomics
// otherwise, once the pool is used up, or for a footprint that does not
// pack.
__device__ inline void scatter_density(const GradTargets *G, const LdsStack &L, int param, float *g,
                                       const GridFoot &f, float val) {
  unsigned cell;
  if (G->log_rec == nullptr || param != G->log_param || !pack_cell(f, 0, cell)) {
    scatter_grid(g, f, val);
    return;
  }
  const int tile = f.base >> G->log_shift;
  const int bin = ((f.base + f.dx + f.dy + f.dz) >> G->log_shift) == tile ? tile : kLogMaxTiles + tile;
  const int4 rec = log_record(cell, f, val);
  const int chunk = *wave_log_word(L, bin);
  const unsigned r = unsigned(__hip_atomic_fetch_add(wave_log_word(L, kLogBins + bin), 1, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WAVEFRONT));
  const bool ok = chunk >= 0 && r < kLogChunk;
  if (ok) {
    const size_t at = size_t(chunk) * kLogChunk + r;
    G->log_rec[at] = rec;
  }
  const uint64_t bad = __ballot(!ok);
  if (bad == 0) return;
  const int lane = int(__lane_id());
  const int leader = __ffsll((unsigned long long)bad) - 1;
  const int b = __builtin_amdgcn_readlane(bin, leader);
  const uint64_t mb = __ballot(!ok && bin == b);
  int c = -1;
  if (lane == leader) {
    int *const cw = wave_log_word(L, b);
    int *const fw = wave_log_word(L, kLogBins + b);
    if (*cw >= 0) G->log_fill[*cw] = min(unsigned(*fw), kLogChunk);
    const unsigned id = atomicAdd(G->log_ctl, 1u);
    if (id < G->log_chunks) {
      c = int(id);
      G->log_list[size_t(b) * G->log_chunks + atomicAdd(G->log_ctl + 1 + b, 1u)] = id;
    }
    *cw = c;
    *fw = c >= 0 ? int(__popcll(mb)) : 0;
  }
  c = __builtin_amdgcn_readlane(c, leader);
  if (!ok) {
    if (bin == b && c >= 0) {
      const size_t at = size_t(c) * kLogChunk + unsigned(__popcll(mb & __lanemask_lt()));
      G->log_rec[at] = rec;
    } else {
      scatter_grid(g, f, val);
    }
  }
}

// The medium of the records being read back: its gradient (or tangent)
// buffer and the row and slice strides of its density grid, loaded again
// only when a record names another medium.
struct RecMedium {
  int med = -1;
  float *g = nullptr;
  int sy = 0, sz = 0;
  __device__ void select(const SceneView &S, const GradTargets *G, int m) {
    if (m == med) return;
    med = m;
    const int p = S.media[m].sigma_param;
    g = G->g[p];
    sy = S.params[p].nx;
    sz = S.params[p].nx * S.params[p].ny;
  }
  __device__ GridFoot foot(float4 e, unsigned cell) const {
    GridFoot f;
    f.base = int(cell & kRecBaseMask);
    f.dx = int((cell >> 26) & 1u);
    f.dy = ((cell >> 27) & 1u) ? sy : 0;
    f.dz = ((cell >> 28) & 1u) ? sz : 0;
    f.tx = e.x;
    f.ty = e.y;
    f.tz = e.z;
    return f;
  }
};

// The shadow-ray collisions of one NEE (footprint, -scale / sigma_n of the
// NEW state, 0 where sigma_n = 0) recorded by the replay pass, so that the
// transmittance gradient can be scattered once the NEE contribution is known
// without tracking the shadow ray a second time; n counts the records and is
// kNeeRecords + 1 once they overflow (more collisions than records, or a
// footprint that does not pack), and such a shadow ray is tracked again.
constexpr int kNeeRecords = 32;
struct NeeRecord {
  int n;
  float hseg;  // length of the shadow ray inside the homogeneous medium
  float4 e[kNeeRecords];
  unsigned cell[kNeeRecords];
};

// A shadow ray of sample_emitter_twostates in flight: its current segment,
// the distance tracked so far, its medium and the two transmittances.
template <int NS>
struct Shadow {
  Ray ray;
  float total;
  int medium;
  bool needs_int;
  float T[NS];
};

// Emitter direction without visibility (plugins/twostatevolpath.py:397-405)
// and the shadow ray towards it; returns the emitter's value / pdf (0 and
// ds.pdf = 0 when the sample is invalid: nothing to track).
template <int NS>
__device__ inline V3 shadow_begin(const SceneView &S, V3 ref_p, V3 ref_n, bool ref_surface, int medium, Pcg32 &rng,
                                  DirSample &ds, Shadow<NS> &sh) {
  const float u1 = rng.next_1d();
  const float u2 = rng.next_1d();
  const V3 ev = sample_emitter_direction(S, ref_p, u1, u2, ds);
  sh.ray.o = ref_surface ? offset_point(ref_p, ref_n, ds.d) : ref_p;
  sh.ray.d = ds.d;
  sh.ray.maxt = kInf;
  sh.total = 0.0f;
  sh.medium = medium;
  sh.needs_int = true;
#pragma unroll
  for (int k = 0; k < NS; ++k) sh.T[k] = 1.0f;
  return ev;
}

// One step of the ratio tracking of a shadow ray towards an emitter sample
// at distance `dist` (the body of the loop of :409-477): a collision in the
// medium (transmittance ratio sigma_n / majorant per state) or the next
// surface (null: crossed; otherwise the transmittance is 0). False once the
// shadow ray has reached the sample or both transmittances are 0. `si` holds
// the shadow ray's next surface between steps. GRAD: see nee_twostates.
template <int NS, bool GRAD, int F>
__device__ inline bool shadow_step(const SceneView &S, Shadow<NS> &sh, SurfInt &si, float dist, Pcg32 &rng,
                                   const LdsStack &stack, LocalCounters &cnt, float gscale, const GradTargets *G,
                                   NeeRecord *rec, float *hom_acc, float *fwd_sum) {
  const float remaining = dist * (1.0f - kShadowEps) - sh.total;
  sh.ray.maxt = remaining;
  if (!(remaining > 0.0f)) return false;
  if (sh.needs_int) {
    Hit h;
    trace<false>(S, sh.ray, h, stack);
    cnt.add(kCntTraversals, 1);
    si = make_si(S, sh.ray, h);
    sh.needs_int = false;
  }
  bool active_medium = sh.medium >= 0;
  bool active_surface = !active_medium;
  MediumEv<NS> mev;
  mev.valid = false;
  mev.t = kInf;
  float trm[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) trm[k] = 1.0f;
  if (active_medium) {
    const float u = rng.next_1d();
    sample_medium<NS, F>(S, sh.medium, sh.ray, u, mev);
    cnt.add(kCntNeeSteps, 1);
    if (si.t < mev.t) {
      mev.valid = false;
      mev.t = kInf;
    }
    if ((F & kFeatHomog) && S.media[sh.medium].homogeneous) {
      const float seg = fminf(fminf(remaining, si.t), si.t) - mev.mint;
#pragma unroll
      for (int k = 0; k < NS; ++k) trm[k] = expf(-(seg * mev.st[k]));
      mev.valid = false;
      mev.t = kInf;
      if (rec) rec->hseg += seg;
      const MediumInfo &M = S.media[sh.medium];
      if (GRAD && hom_acc && G->g[M.sigma_param]) *hom_acc += -gscale * M.scale * seg;
      if (GRAD && fwd_sum)
        if (const float *tv = G->g[M.sigma_param]) *fwd_sum += -M.scale * seg * tv[0];
    }
  }
  const bool escaped = active_medium && !mev.valid;
  active_medium = active_medium && mev.valid;
  if (active_medium) {
    sh.ray.o = mev.p;
    si.t = si.t - mev.t;
#pragma unroll
    for (int k = 0; k < NS; ++k) trm[k] = mev.maj != 0.0f ? mev.sn[k] / mev.maj : mev.sn[k];
    if (GRAD && (fwd_sum || gscale != 0.0f)) {
      const MediumInfo &M = S.media[sh.medium];
      float *g = G->g[M.sigma_param];
      if (g && mev.sn[NS - 1] != 0.0f) {
        if (fwd_sum) {
          *fwd_sum += -M.scale / mev.sn[NS - 1] * grid_eval1(g, mev.foot);
        } else {
          scatter_density(G, stack, M.sigma_param, g, mev.foot, -gscale * M.scale / mev.sn[NS - 1]);
          cnt.add(kCntGradAtomics, 8);
        }
      }
    }
    if (rec) {
      unsigned cell;
      if (rec->n < kNeeRecords && pack_cell(mev.foot, sh.medium, cell)) {
        const float sn = mev.sn[NS - 1];
        rec->e[rec->n] = make_float4(mev.foot.tx, mev.foot.ty, mev.foot.tz,
                                     sn != 0.0f ? -S.media[sh.medium].scale / sn : 0.0f);
        rec->cell[rec->n] = cell;
        rec->n++;
      } else {
        rec->n = kNeeRecords + 1;
      }
    }
  }
  active_surface = (active_surface || escaped) && si.valid && !active_medium;
  if (active_surface) {
    const float b = (!(F & kFeatSurfaces) || null_bsdf(S, si.shape)) ? 1.0f : 0.0f;
#pragma unroll
    for (int k = 0; k < NS; ++k) trm[k] = trm[k] * b;
  }
  bool any_t = false;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    sh.T[k] *= trm[k];
    any_t = any_t || sh.T[k] != 0.0f;
  }
  if (active_surface) {
    sh.ray = spawn_ray(si, sh.ray.d);
    sh.needs_int = true;
  }
  const bool active = (active_medium || active_surface) && any_t;
  if (!active) return false;
  sh.total += active_medium ? mev.t : si.t;
  if (active_surface) {
    const ShapeInfo bs = S.shapes[si.shape];
    if (bs.interior != bs.exterior) sh.medium = dot(si.n, sh.ray.d) > 0.0f ? bs.exterior : bs.interior;
  }
  return true;
}

// sample_emitter_twostates (plugins/twostatevolpath.py:387-479): emitter
// direction without visibility, then ratio tracking of both transmittances.
// GRAD = true replays the same random numbers and scatters
// d(log T)/d(density) * gscale into the density gradient.
//
// In a homogeneous medium (:422-431) the shadow ray advances straight to the
// next surface or the end of the segment with the analytic transmittance
// exp(-sigma_t d) of each state (TwoStateMedium.eval_tr_old/new,
// plugins/twostatemedium.py:109-115); GRAD adds d(log T)/d(value) * gscale =
// -scale d gscale to *hom_acc. In forward mode (fwd_sum set) the GRAD pass
// instead adds d(log T)/d(theta) . tangent (G holds the tangents) to
// *fwd_sum.
template <int NS, bool GRAD, int F>
__device__ inline void nee_twostates(const SceneView &S, V3 ref_p, V3 ref_n, bool ref_surface, int medium,
                                     Pcg32 &rng, V3 em[NS], DirSample &ds, const LdsStack &stack,
                                     LocalCounters &cnt, float gscale, const GradTargets *G,
                                     NeeRecord *rec = nullptr, float *hom_acc = nullptr,
                                     float *fwd_sum = nullptr) {
  Shadow<NS> sh;
  const V3 ev = shadow_begin<NS>(S, ref_p, ref_n, ref_surface, medium, rng, ds, sh);
#pragma unroll
  for (int k = 0; k < NS; ++k) em[k] = v3(0.f);
  if (ds.pdf == 0.0f) return;
  SurfInt si;
  si.valid = false;
  si.t = kInf;
  while (shadow_step<NS, GRAD, F>(S, sh, si, ds.dist, rng, stack, cnt, gscale, G, rec, hom_acc, fwd_sum)) {
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) em[k] = ev * sh.T[k];
}

struct AdjState {
  V3 dL;    // dLoss/dL of this camera sample
  V3 Lrem;  // radiance still to be gathered along the replayed path
  V3 dI;    // forward mode: d(radiance)/d(theta) . tangent of this sample
};

// The collisions of one primal pass of the adjoint: per collision the point
// and d log(weight)/d(density) per unit of remaining radiance (scale /
// sigma_t for a real collision, -scale / sigma_n for a null one), dLoss/dL .
// L gathered before the collision, and the medium. Once the path's radiance
// L is known, the collision's gradient is that coefficient times dLoss/dL .
// (L - the gathered part), which is the replay pass's dLoss/dL . Lrem at the
// same collision, so the path need not be traced a second time. A path with
// more than kTapeRecords collisions is replayed instead.
constexpr int kTapeRecords = 32;
struct PathTape {
  int n;
  float4 e[kTapeRecords];      // footprint fractions, coefficient
  float s[kTapeRecords];       // dLoss/dL . L gathered before the collision
  unsigned cell[kTapeRecords];  // packed footprint and medium (pack_cell)
};
struct NoTape {};

// Untaped adjoint: dLoss/dL . c of each NEE contribution c of the primal
// pass, in path order, in the lane's LDS slots (LdsStack::slot). The replay
// pass reads the value before it tracks the same shadow ray (same random
// numbers, same collisions), so it scatters the transmittance gradient while
// tracking: one shadow-ray tracking per pass and nothing recorded per
// collision. A path with more NEEs than slots tracks the later ones' shadow
// rays again in the replay. (Per-lane scratch instead of LDS cost the 64^3
// adjoint ~6 ms at a mid state of the optimisation and doubled it at the
// initial one: each slot's store and load went to memory.)
struct NeeWeights {
  int k;  // NEEs of the path so far
};
struct NoNeeWeights {};

// Deferred NEE (renders of scenes without surfaces, nee_start): after the
// lanes on their paths have advanced them, a wavefront steps only its shadow
// rays while more than this many of its lanes track one, then once more with
// all of them. Measured on the 64^3 volume workload against tracking each
// shadow ray at once, at mid and final states of the optimisation: the
// single-state render 25-27 % and the two-state render 23-24 % faster with
// 16 (8, 32 and 64: less), slower only at a uniform density, where every
// shadow ray ends at its first collision. The adjoint, whose path code (run
// on every trip) is heavier than a tracking step, was slower with every hold
// from 0 to 64 and tracks each shadow ray at once.
constexpr int kNeeHold = 16;

// A deferred NEE in flight lives in the lane's LDS slots (a render kernel
// uses them for nothing else), so that it holds no registers across bounce()
// calls: words 0-5 the shadow ray's origin and direction, 6 the distance
// tracked, 7 its medium, 8 whether its next surface must be traced, 9-10 the
// transmittances, 11 the emitter sample's distance, 12 its pdf, 13-15 its
// value / pdf (kNeeFlightWords words). Scenes with fewer slots (a deep BVH)
// take the kernels that track at once (Scene::features).

template <int NS>
__device__ inline void flight_put(const LdsStack &L, const Shadow<NS> &sh) {
  L.slot(0) = __float_as_int(sh.ray.o.x);
  L.slot(1) = __float_as_int(sh.ray.o.y);
  L.slot(2) = __float_as_int(sh.ray.o.z);
  L.slot(3) = __float_as_int(sh.ray.d.x);
  L.slot(4) = __float_as_int(sh.ray.d.y);
  L.slot(5) = __float_as_int(sh.ray.d.z);
  L.slot(6) = __float_as_int(sh.total);
  L.slot(7) = sh.medium;
  L.slot(8) = sh.needs_int ? 1 : 0;
#pragma unroll
  for (int k = 0; k < NS; ++k) L.slot(9 + k) = __float_as_int(sh.T[k]);
}

template <int NS>
__device__ inline void flight_get(const LdsStack &L, Shadow<NS> &sh) {
  sh.ray.o = v3(__int_as_float(L.slot(0)), __int_as_float(L.slot(1)), __int_as_float(L.slot(2)));
  sh.ray.d = v3(__int_as_float(L.slot(3)), __int_as_float(L.slot(4)), __int_as_float(L.slot(5)));
  sh.ray.maxt = kInf;
  sh.total = __int_as_float(L.slot(6));
  sh.medium = L.slot(7);
  sh.needs_int = L.slot(8) != 0;
#pragma unroll
  for (int k = 0; k < NS; ++k) sh.T[k] = __int_as_float(L.slot(9 + k));
}

// One-pass (taped) adjoint: for kernels whose media have a majorant
// supergrid and whose only medium gradient is the density's. Records live in
// per-lane scratch, which does not fit in cache, so a record written and
// read back from HBM costs more than tracking the collision again (an
// L2-resident grid lookup) when paths collide often, as under one global
// majorant in an optically thick medium; with a supergrid, tracking a
// collision again costs a DDA walk and paths collide less. The albedo grid's
// gradient needs the colour of the remaining radiance, so kernels with an
// albedo grid keep the replay pass. The mode is uniform per kernel: lanes of
// one wavefront in different modes execute both modes' work.
template <int F>
constexpr bool kTapeAdjoint = (F & kFeatSupergrid) && !(F & kFeatAlbedoGrid);

// Adjoint of the homogeneous medium's constants (sigma_t value, albedo):
// every event adds to per-lane sums, which a kernel reduces over its
// wavefront and adds to the gradient once, at its end (flush_homogeneous).
struct HomSum {
  float s;  // d/d(sigma_t value)
  V3 a;     // d/d(albedo)
};
struct NoHomSum {};

// All 64 lanes of the wavefront must be active.
__device__ inline void flush_homogeneous(const SceneView &S, const GradTargets &G, HomSum h) {
  for (int off = 32; off > 0; off >>= 1) {
    h.s += __shfl_xor(h.s, off, 64);
    h.a.x += __shfl_xor(h.a.x, off, 64);
    h.a.y += __shfl_xor(h.a.y, off, 64);
    h.a.z += __shfl_xor(h.a.z, off, 64);
  }
  if (__lane_id() == 0 && S.hom_medium >= 0) {
    const MediumInfo &M = S.media[S.hom_medium];
    if (float *g = G.g[M.sigma_param]) atomicAdd(g, h.s);
    if (M.albedo_param >= 0)
      if (float *g = G.g[M.albedo_param]) {
        atomicAdd(g + 0, h.a.x);
        atomicAdd(g + 1, h.a.y);
        atomicAdd(g + 2, h.a.z);
      }
  }
}

// One camera sample of the two-state volumetric path tracer, advanced one
// bounce per call (the body of the reference's loop, :175-380), so that a
// kernel can start a lane's next camera sample as soon as its path ends
// instead of waiting for the longest path of its wavefront. NS = 2: L[0] =
// old (H), L[1] = new (F). ADJ (NS = 1) is the adjoint of path-replay
// backpropagation: d(log weight)/d(theta) at each collision times the
// radiance still to come is scattered into the gradients. With a tape
// (kTapeAdjoint) the primal pass records the collisions and scatters the NEE
// transmittance gradients at once, and end_adjoint scatters the collisions';
// otherwise the replay pass (replay = true, the primal radiance in adj.Lrem)
// replays the same random numbers and scatters as it goes. FWD (with ADJ) is
// forward mode: the replay pass gathers the tangents (in G) where the
// adjoint scatters and adds d(radiance)/d(theta) . tangent to adj.dI, the
// transpose of the adjoint on the same samples.
template <int NS, bool ADJ, int F, bool FWD = false>
struct VolPath {
  static_assert(!FWD || ADJ, "forward mode replays paths as the adjoint does");
  static constexpr bool kTape = ADJ && !FWD && kTapeAdjoint<F>;
  Pcg32 rng;
  Ray ray;
  V3 thr[NS], L[NS];
  SurfInt si;
  V3 last_p;
  float last_pdf;
  int depth, medium;
  bool needs_int, specular_chain, replay, nee_grad;
  AdjState adj;
  std::conditional_t<kTape, PathTape, NoTape> tape;
  static constexpr bool kNeeInline = ADJ && !FWD && !kTape;
  std::conditional_t<kNeeInline, NeeWeights, NoNeeWeights> neew;
  // Renders of scenes without surfaces defer the ratio tracking of NEE
  // shadow rays to the following bounce() calls (nee_start).
  static constexpr bool kDeferNee = !ADJ && !(F & kFeatSurfaces);
  bool in_nee;
  std::conditional_t<ADJ && (F & kFeatHomog) != 0, HomSum, NoHomSum> hom;

  // Zeroes the homogeneous-medium sums (once per lane, before its samples).
  __device__ void init_hom() {
    if constexpr (ADJ && (F & kFeatHomog)) {
      hom.s = 0.0f;
      hom.a = v3(0.f);
    }
  }

  __device__ void begin(const Pcg32 &r, const Ray &camera, LocalCounters &cnt, bool is_replay) {
    rng = r;
    ray = camera;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      L[k] = v3(0.f);
      thr[k] = v3(1.f);
    }
    depth = 0;
    specular_chain = true;
    si.valid = false;
    si.t = kInf;
    si.shape = -1;
    si.p = v3(0.f);
    si.n = v3(0.f, 0.f, 1.f);
    si.u = si.v = 0.0f;
    needs_int = true;
    last_p = v3(0.f);
    last_pdf = 1.0f;
    medium = -1;
    replay = ADJ && is_replay;
    nee_grad = false;
    in_nee = false;
    (void)rng.next_1d();  // colour channel for free-flight sampling (:171-173)
    cnt.add(kCntPaths, 1);
  }

  // Adjoint: start the primal pass (is_replay false) or the replay pass of a
  // camera sample with dLoss/dL = dL and primal radiance Lp (read by the
  // replay only). NEE transmittance gradients do not depend on the radiance
  // still to come: a taped adjoint scatters them in the primal pass, an
  // untaped one in the replay.
  __device__ void begin_adjoint(const Pcg32 &r, const Ray &camera, LocalCounters &cnt, bool is_replay, V3 dL,
                                V3 Lp) {
    begin(r, camera, cnt, is_replay);
    if constexpr (ADJ) {
      adj.dL = dL;
      adj.Lrem = Lp;
      adj.dI = v3(0.f);
      nee_grad = kTape ? !is_replay : is_replay;
      if constexpr (kTape) tape.n = 0;
      if constexpr (kNeeInline) neew.k = 0;
    }
  }

  // Adjoint, once bounce() has returned false: true when the camera sample is
  // done (its taped collisions scattered), false when it needs its replay
  // pass (an untaped kernel, or a tape overflow).
  __device__ bool end_adjoint(const SceneView &S, LocalCounters &cnt, const GradTargets *G) {
    if constexpr (kTape) {
      if (replay || !any_nonzero(L[0])) return true;
      if (tape.n <= kTapeRecords) {
        const float total = dot(adj.dL, L[0]);
        RecMedium rm;
        for (int j = 0; j < tape.n; ++j) {
          const float lr = total - tape.s[j];
          if (lr != 0.0f) {
            const float4 e = tape.e[j];
            const unsigned cell = tape.cell[j];
            rm.select(S, G, cell_medium(cell));
            scatter_grid(rm.g, rm.foot(e, cell), lr * e.w);
            cnt.add(kCntGradAtomics, 8);
          }
        }
        return true;
      }
      return false;
    } else if constexpr (ADJ) {
      return replay || !any_nonzero(L[0]);
    }
    return true;
  }

  // The contribution c[k] = thr nw mis em of an emitter sample whose shadow
  // ray has been tracked (em: the emitter's value / pdf times the
  // transmittance; nw, nee_pdf: the BSDF or phase function towards it),
  // added to L (:331-344), and its adjoint or forward-mode terms. tsum, cs
  // (have_cs) and rec are what the tracking gathered for them; rng0 is the
  // random state at the NEE's start, with which a shadow ray is tracked again
  // when its gradient could not be taken while tracking.
  __device__ void nee_account(const SceneView &S, const LdsStack &stack, LocalCounters &cnt, const GradTargets *G,
                              const V3 em[NS], const V3 nw[NS], float nee_pdf, float ds_pdf, float tsum,
                              bool have_cs, uint64_t rng0, V3 ref_p, V3 ref_n, bool ref_surface, NeeRecord *rec) {
    const float mis_em = mis_weight(ds_pdf, nee_pdf);
    V3 c[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      c[k] = thr[k] * nw[k] * mis_em * em[k];
      L[k] += c[k];
    }
    if constexpr (FWD) {
      // the shadow ray's d(log T)/d(theta) . tangent times this NEE
      // contribution
      if (nee_grad && any_nonzero(c[0])) adj.dI += c[0] * tsum;
    }
    if constexpr (kNeeInline) {
      if (!replay) {
        if (neew.k < stack.nslots - 1) stack.slot(neew.k) = __float_as_int(dot(adj.dL, c[0]));
      } else if (!have_cs) {
        // past the slots: track the shadow ray again, now with its weight
        const float cs = dot(adj.dL, c[0]);
        if (cs != 0.0f && any_nonzero(c[0])) {
          Pcg32 r2 = rng;
          r2.state = rng0;
          V3 em2[NS];
          DirSample ds2;
          float *hacc = nullptr;
          if constexpr ((F & kFeatHomog) != 0) hacc = &hom.s;
          nee_twostates<NS, true, F>(S, ref_p, ref_n, ref_surface, medium, r2, em2, ds2, stack, cnt, cs, G, nullptr,
                                     hacc);
        }
      }
      ++neew.k;
    }
    if constexpr (kTape) {
      const float cs = dot(adj.dL, c[0]);
      if (nee_grad && cs != 0.0f && any_nonzero(c[0])) {
        if (rec->n <= kNeeRecords) {
          if constexpr ((F & kFeatHomog) != 0) {
            if (rec->hseg != 0.0f) hom.s += -cs * S.media[S.hom_medium].scale * rec->hseg;
          }
          // d(log T)/d(density) at each recorded shadow-ray collision
          RecMedium rm;
          for (int j = 0; j < rec->n; ++j) {
            const float4 e = rec->e[j];
            const unsigned cell = rec->cell[j];
            rm.select(S, G, cell_medium(cell));
            if (rm.g && e.w != 0.0f) {
              scatter_grid(rm.g, rm.foot(e, cell), cs * e.w);
              cnt.add(kCntGradAtomics, 8);
            }
          }
        } else {
          Pcg32 r2 = rng;
          r2.state = rng0;
          V3 em2[NS];
          DirSample ds2;
          float *hacc = nullptr;
          if constexpr ((F & kFeatHomog) != 0) hacc = &hom.s;
          nee_twostates<NS, true, F>(S, ref_p, ref_n, ref_surface, medium, r2, em2, ds2, stack, cnt, cs, G, nullptr,
                                     hacc);
        }
      }
    }
    if (ADJ && replay) adj.Lrem -= c[0];
  }

  // Deferred NEE: the emitter sample and shadow ray of a medium NEE at
  // ref_p. Its ratio tracking then advances one step per bounce() call
  // (lanes of a wavefront in a NEE and lanes on their paths advance
  // together, instead of the wavefront waiting for its longest shadow ray),
  // drawing the same random numbers in the same order as tracking it at
  // once; the path resumes once the NEE is accounted. The path's surface
  // record `si` is free until then (a real scattering event traces the new
  // direction again) and holds the shadow ray's.
  __device__ void nee_start(const SceneView &S, V3 ref_p, const LdsStack &stack, LocalCounters &cnt,
                            const GradTargets *G) {
    if constexpr (kDeferNee) {
      DirSample ds;
      Shadow<NS> sh;
      const V3 ev = shadow_begin<NS>(S, ref_p, v3(0.f), false, medium, rng, ds, sh);
      si.valid = false;
      si.t = kInf;
      if (ds.pdf == 0.0f) {
        nee_finish(S, stack, cnt, G, sh, 0.0f, ev);
      } else {
        flight_put<NS>(stack, sh);
        stack.slot(11) = __float_as_int(ds.dist);
        stack.slot(12) = __float_as_int(ds.pdf);
        stack.slot(13) = __float_as_int(ev.x);
        stack.slot(14) = __float_as_int(ev.y);
        stack.slot(15) = __float_as_int(ev.z);
        in_nee = true;
      }
    }
  }

  // Deferred NEE: one ratio-tracking step; accounts the NEE once the shadow
  // ray is done.
  __device__ void nee_advance(const SceneView &S, const LdsStack &stack, LocalCounters &cnt, const GradTargets *G) {
    if constexpr (kDeferNee) {
      Shadow<NS> sh;
      flight_get<NS>(stack, sh);
      const float dist = __int_as_float(stack.slot(11));
      if (shadow_step<NS, false, F>(S, sh, si, dist, rng, stack, cnt, 0.0f, G, nullptr, nullptr, nullptr)) {
        flight_put<NS>(stack, sh);
      } else {
        const V3 ev = v3(__int_as_float(stack.slot(13)), __int_as_float(stack.slot(14)),
                         __int_as_float(stack.slot(15)));
        nee_finish(S, stack, cnt, G, sh, __int_as_float(stack.slot(12)), ev);
      }
    }
  }

  __device__ void nee_finish(const SceneView &S, const LdsStack &stack, LocalCounters &cnt, const GradTargets *G,
                             const Shadow<NS> &sh, float pdf, V3 ev) {
    if constexpr (kDeferNee) {
      in_nee = false;
      V3 em[NS], nw[NS];
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        em[k] = pdf == 0.0f ? v3(0.f) : ev * sh.T[k];
        nw[k] = v3(kInv4Pi);
      }
      nee_account(S, stack, cnt, G, em, nw, kInv4Pi, pdf, 0.0f, false, 0, ray.o, v3(0.f), false, nullptr);
    }
  }

  // One bounce; false once the path has ended. With deferred NEE, the lanes
  // on their paths advance them first (a lane may start a NEE there), then
  // the lanes with a shadow ray in flight step it: alone while more than
  // kNeeHold of them track one, then once more together, so a shadow ray
  // that ends at its first collision costs no extra trip.
  __device__ bool bounce(const SceneView &S, const IntegArgs &I, const LdsStack &stack, LocalCounters &cnt,
                         const GradTargets *G) {
    bool live = true;
    if (!kDeferNee || !in_nee) live = path_step(S, I, stack, cnt, G);
    if constexpr (kDeferNee) {
      while (__popcll(__ballot(in_nee)) > kNeeHold) {
        if (in_nee) nee_advance(S, stack, cnt, G);
      }
      if (in_nee) nee_advance(S, stack, cnt, G);
    }
    return live;
  }

  // The body of the reference's loop for one bounce of the path.
  __device__ bool path_step(const SceneView &S, const IntegArgs &I, const LdsStack &stack, LocalCounters &cnt,
                            const GradTargets *G) {
    const int max_depth = I.max_depth < 0 ? 0x7fffffff : I.max_depth;
    // ---- Russian roulette (:182-188)
    bool nonzero = false;
  #pragma unroll
    for (int k = 0; k < NS; ++k) nonzero = nonzero || any_nonzero(thr[k]);
    if (!nonzero) return false;
    const V3 tm = NS == 2 ? vmax(thr[NS - 1], thr[0]) : thr[0];
    const float q = fminf(hmax(tm), 0.99f);  // eta = 1: no dielectric interfaces
    const bool perform_rr = depth > I.rr_depth;
    const float urr = rng.next_1d();
    if (perform_rr && !(urr < q)) return false;
    if (perform_rr) {
      const float rq = 1.0f / q;
  #pragma unroll
      for (int k = 0; k < NS; ++k) thr[k] = thr[k] * rq;
    }

    bool active_medium = medium >= 0;
    bool active_surface = !active_medium;
    const bool in_hom = (F & kFeatHomog) && active_medium && S.media[medium].homogeneous;
    // delta tracking with null collisions unless no medium of the scene is
    // heterogeneous (handle_null_scattering of upstream prbvolpath)
    const bool null_mode = !(F & kFeatHomog) || S.handle_null;

    // ---- free flight with the combined majorant (:193-206)
    MediumEv<NS> mev;
    mev.valid = false;
    mev.t = kInf;
    if (active_medium) {
      const float u = rng.next_1d();
      sample_medium<NS, F>(S, medium, ray, u, mev);
      cnt.add(kCntCollisions, 1);
      // a homogeneous medium's collision bounds the surface search (:207)
      if (in_hom && mev.valid) ray.maxt = mev.t;
      if (needs_int) {
        Hit h;
        trace<false>(S, ray, h, stack);
        cnt.add(kCntTraversals, 1);
        si = make_si(S, ray, h);
      }
      needs_int = false;
      if (si.t < mev.t) {
        mev.valid = false;
        mev.t = kInf;
      }
    }
    const bool escaped = active_medium && !mev.valid;
    active_medium = active_medium && mev.valid;

    bool act_null = false, act_scatter = false;
    V3 w[NS];
  #pragma unroll
    for (int k = 0; k < NS; ++k) w[k] = v3(1.f);
    // ---- without null scattering: transmittance over the NEW state's
    // free-flight pdf (:212-213, 232-247)
    float ff_t = 0.0f;
    if ((F & kFeatHomog) && !null_mode && (escaped || active_medium)) {
      ff_t = fminf(mev.t, si.t) - mev.mint;
      const float sig = mev.st[NS - 1];
      const float tr = expf(-(ff_t * sig));
      const float pdf = si.t < mev.t ? tr : tr * sig;
  #pragma unroll
      for (int k = 0; k < NS; ++k) {
        const float r = pdf > 0.0f ? expf(-(ff_t * mev.st[k])) / pdf : 0.0f;
        if (escaped)
          thr[k] = thr[k] * r;
        else
          w[k] = w[k] * r;
      }
      act_scatter = active_medium;
    }
    // ---- null / real decision with the averaged probability (:219-231)
    if (active_medium && null_mode) {
      const float maj = mev.maj;
      float p;
      if (NS == 2)
        p = maj != 0.0f ? (mev.st[1] / maj + mev.st[0] / maj) * 0.5f : 0.0f;
      else
        p = maj != 0.0f ? mev.st[0] / maj : 0.0f;
      const float un = rng.next_1d();
      act_null = un >= p;
      act_scatter = !act_null;
      if (act_null) cnt.add(kCntNullEvents, 1);
      if (NS == 2 && act_null) {
        const float den = mev.sn[1] + mev.sn[0];
        w[1] = w[1] * (2.0f * mev.sn[1] / den);
        w[0] = w[0] * (2.0f * mev.sn[0] / den);
      }
    }
    if (act_scatter) {
      depth += 1;
      last_p = mev.p;
    }
    const bool alive = depth < max_depth;  // (:253)
    act_scatter = act_scatter && alive;
    if (act_null) {
      ray.o = mev.p;
      si.t = si.t - mev.t;
    }
    if (act_scatter) {
      cnt.add(kCntRealScatter, 1);
      if (!null_mode) {
  #pragma unroll
        for (int k = 0; k < NS; ++k) w[k] = w[k] * mev.ss[k];  // (:261-263)
      } else if (NS == 2) {
        const float den = mev.st[1] + mev.st[0];
        w[1] = w[1] * (2.0f * mev.ss[1] / den);
        w[0] = w[0] * (2.0f * mev.ss[0] / den);
      } else {
        w[0] = w[0] * (mev.ss[0] / mev.st[0]);
      }
    }
    if (active_medium) {
  #pragma unroll
      for (int k = 0; k < NS; ++k) thr[k] *= w[k];
    }
    if constexpr (kTape) {
      if (!replay && (act_null || act_scatter)) {
        // record the collision (its gradient needs the path's radiance)
        const MediumInfo &M = S.media[medium];
        if (G->g[M.sigma_param] && !(act_null && mev.sn[0] == 0.0f)) {
          unsigned cell;
          if (tape.n < kTapeRecords && pack_cell(mev.foot, medium, cell)) {
            const float coef = act_null ? -M.scale / mev.sn[0] : M.scale / mev.st[0];
            tape.e[tape.n] = make_float4(mev.foot.tx, mev.foot.ty, mev.foot.tz, coef);
            tape.s[tape.n] = dot(adj.dL, L[0]);
            tape.cell[tape.n] = cell;
            tape.n++;
          } else {
            tape.n = kTapeRecords + 1;
          }
        }
      }
    }
    if constexpr (ADJ && (F & kFeatHomog) != 0) {
      if (replay && in_hom && (escaped || active_medium)) {
        // d log(weight)/d(value) (sigma_t = scale value) of this segment's
        // weight times the radiance it multiplies: null collision -scale /
        // sigma_n, real collision +scale / sigma_t; without null scattering
        // exp(-sigma_t t) / pdf gives -scale t, and a real collision's
        // sigma_s (= albedo sigma_t) +scale / sigma_t.
        const MediumInfo &M = S.media[medium];
        const float lr = dot(adj.dL, adj.Lrem);
        float coef = 0.0f;
        if (null_mode) {
          if (act_null)
            coef = mev.sn[0] != 0.0f ? -M.scale / mev.sn[0] : 0.0f;
          else if (act_scatter)
            coef = M.scale / mev.st[0];
        } else {
          coef = -ff_t * M.scale + (act_scatter ? M.scale / mev.st[0] : 0.0f);
        }
        if constexpr (FWD) {
          if (const float *tv = G->g[M.sigma_param]) adj.dI += adj.Lrem * (coef * tv[0]);
          if (act_scatter && M.albedo_param >= 0)
            if (const float *ta = G->g[M.albedo_param])
              adj.dI += adj.Lrem * v3(mev.alb.x != 0.f ? ta[0] / mev.alb.x : 0.f,
                                      mev.alb.y != 0.f ? ta[1] / mev.alb.y : 0.f,
                                      mev.alb.z != 0.f ? ta[2] / mev.alb.z : 0.f);
        } else {
          hom.s += lr * coef;
          if (act_scatter && M.albedo_param >= 0) {
            const V3 dl = adj.dL * adj.Lrem;
            hom.a += v3(mev.alb.x != 0.f ? dl.x / mev.alb.x : 0.f, mev.alb.y != 0.f ? dl.y / mev.alb.y : 0.f,
                        mev.alb.z != 0.f ? dl.z / mev.alb.z : 0.f);
          }
        }
      }
    }
    if (ADJ && replay && !in_hom && (act_null || act_scatter)) {
      // d log(weight) / d theta at this collision times the radiance that
      // the weight multiplies (everything gathered after this point).
      const MediumInfo &M = S.media[medium];
      float *gd = G->g[M.sigma_param];
      const float lr = dot(adj.dL, adj.Lrem);
      if constexpr (FWD) {
        // the tangent gathered over the same footprint
        if (gd && !(act_null && mev.sn[0] == 0.0f)) {
          const float coef = act_null ? -M.scale / mev.sn[0] : M.scale / mev.st[0];
          adj.dI += adj.Lrem * (coef * grid_eval1(gd, mev.foot));
        }
        if ((F & kFeatAlbedoGrid) && act_scatter && M.albedo_param >= 0 && G->g[M.albedo_param]) {
          GridFoot fa;
          grid_footprint(S.params[M.albedo_param], M, mev.p, fa);
          const V3 ta = grid_eval3(G->g[M.albedo_param], fa);
          adj.dI += adj.Lrem * v3(mev.alb.x != 0.f ? ta.x / mev.alb.x : 0.f, mev.alb.y != 0.f ? ta.y / mev.alb.y : 0.f,
                                  mev.alb.z != 0.f ? ta.z / mev.alb.z : 0.f);
        }
      } else if (gd && lr != 0.0f) {
        if (act_null) {
          if (mev.sn[0] != 0.0f) scatter_density(G, stack, M.sigma_param, gd, mev.foot, -lr * M.scale / mev.sn[0]);
        } else {
          scatter_density(G, stack, M.sigma_param, gd, mev.foot, lr * M.scale / mev.st[0]);
        }
        cnt.add(kCntGradAtomics, 8);
      }
      if (!FWD && (F & kFeatAlbedoGrid) && act_scatter && M.albedo_param >= 0 && G->g[M.albedo_param]) {
        const V3 dl = adj.dL * adj.Lrem;
        const V3 ga = v3(mev.alb.x != 0.f ? dl.x / mev.alb.x : 0.f, mev.alb.y != 0.f ? dl.y / mev.alb.y : 0.f,
                         mev.alb.z != 0.f ? dl.z / mev.alb.z : 0.f);
        GridFoot fa;
        grid_footprint(S.params[M.albedo_param], M, mev.p, fa);
        scatter_grid3(G->g[M.albedo_param], fa, ga);
        cnt.add(kCntGradAtomics, 24);
      }
    }

    // ---- phase sampling (:268-281), isotropic
    if (act_scatter) {
      (void)rng.next_1d();
      const float a = rng.next_1d();
      const float b = rng.next_1d();
      const V3 wo = square_to_uniform_sphere(a, b);
      ray.o = mev.p;
      ray.d = wo;
      ray.maxt = kInf;
      needs_int = true;
      last_pdf = kInv4Pi;
    }

    // ---- surface interactions and emitter hits (:284-308)
    active_surface = active_surface || escaped;
    if (active_surface && needs_int) {
      Hit h;
      trace<false>(S, ray, h, stack);
      cnt.add(kCntTraversals, 1);
      si = make_si(S, ray, h);
    }
    if (active_surface) {
      const int em = si_emitter(S, si);
      if (em >= 0 && !(depth == 0 && I.hide_emitters)) {
        const bool count_direct = depth == 0 || specular_chain;
        const DirSample ds = dir_sample_from_hit(S, si, last_p, ray.d);
        const float em_pdf = pdf_emitter_direction(S, ds);
        const V3 emitted = emitter_eval(S, em, si, ray.d);
        const float mis = mis_weight(last_pdf, em_pdf);
  #pragma unroll
        for (int k = 0; k < NS; ++k) {
          const V3 c = count_direct ? thr[k] * emitted : thr[k] * mis * emitted;
          L[k] += c;
          if (ADJ && replay) adj.Lrem -= c;
        }
      }
    }
    active_surface = active_surface && si.valid;
    BsdfInfo bsdf;
    bsdf.type = -1;
    bool smooth = false;
    Frame fr;
    V3 wi = v3(0.f, 0.f, 1.f);
    if (active_surface) {
      const int b = S.shapes[si.shape].bsdf;
      if (b >= 0) bsdf = S.bsdfs[b];
      smooth = (F & kFeatSurfaces) && bsdf.type == RCV_BSDF_DIFFUSE;
      fr = make_frame(si.n);
      wi = fr.to_local(-ray.d);
    }

    // ---- emitter sampling (:315-344)
    const bool active_e_surface = active_surface && smooth && (depth + 1 < max_depth);
    const bool sample_emitters = act_scatter && S.media[medium].sample_emitters != 0;
    specular_chain = specular_chain && !act_scatter;
    specular_chain = specular_chain || (act_scatter && !sample_emitters);
    const bool active_e_medium = act_scatter && sample_emitters;
    if (active_e_surface || active_e_medium) {
      const V3 ref_p = active_e_medium ? mev.p : si.p;
      if constexpr (kDeferNee) {
        nee_start(S, ref_p, stack, cnt, G);  // no surfaces: a medium NEE
      } else {
        const uint64_t rng0 = rng.state;
        // Untaped adjoint, replay pass: this NEE's dLoss/dL . c from the
        // primal pass, with which the tracking below scatters the
        // transmittance gradient.
        float cs_in = 0.0f;
        bool have_cs = false;
        if constexpr (kNeeInline) {
          if (replay && neew.k < stack.nslots - 1) {
            cs_in = __int_as_float(stack.slot(neew.k));
            have_cs = true;
          }
        }
        float tsum = 0.0f;  // forward mode: d(log T)/d(theta) . tangent along the shadow ray
        NeeRecord rec;      // taped adjoint: the shadow ray's collisions
        if constexpr (kTape) {
          rec.n = 0;
          rec.hseg = 0.0f;
        }
        V3 em[NS];
        DirSample ds;
        if constexpr (ADJ) {
          // one instantiation for both passes, so that they compute the same
          // transmittance from the same random numbers
          float *hacc = nullptr;
          if constexpr (kNeeInline && (F & kFeatHomog) != 0) hacc = have_cs ? &hom.s : nullptr;
          nee_twostates<NS, true, F>(S, ref_p, si.n, active_e_surface, medium, rng, em, ds, stack, cnt, cs_in, G,
                                     kTape && nee_grad ? &rec : nullptr, hacc, FWD && nee_grad ? &tsum : nullptr);
        } else {
          nee_twostates<NS, false, F>(S, ref_p, si.n, active_e_surface, medium, rng, em, ds, stack, cnt, 0.f, G);
        }
        V3 nw[NS];
        float nee_pdf;
        if (active_e_surface) {
          const V3 wo = fr.to_local(ds.d);
  #pragma unroll
          for (int k = 0; k < NS; ++k)
            nw[k] = diffuse_eval(diffuse_reflectance(S, bsdf, state_index<NS>(k), si.u, si.v), wi, wo);
          nee_pdf = diffuse_pdf(wi, wo);
        } else {
  #pragma unroll
          for (int k = 0; k < NS; ++k) nw[k] = v3(kInv4Pi);
          nee_pdf = kInv4Pi;
        }
        nee_account(S, stack, cnt, G, em, nw, nee_pdf, ds.pdf, tsum, have_cs, rng0, ref_p, si.n, active_e_surface,
                    &rec);
      }
    }

    // ---- BSDF sampling (:347-380)
    BsdfSample bs;
    bs.pdf = 0.0f;
    bs.is_null = false;
    bs.weight = v3(0.f);
    bs.wo = v3(0.f, 0.f, 1.f);
    V3 rho_new = v3(0.f);
    if (active_surface) {
      (void)rng.next_1d();
      const float a = rng.next_1d();
      const float b = rng.next_1d();
      if (bsdf.type == RCV_BSDF_NULL) {
        bs.wo = -wi;
        bs.pdf = 1.0f;
        bs.weight = v3(1.f);
        bs.is_null = true;
      } else if ((F & kFeatSurfaces) && bsdf.type == RCV_BSDF_DIFFUSE && wi.z > 0.0f) {
        bs.wo = square_to_cosine_hemisphere(a, b);
        bs.pdf = diffuse_pdf(wi, bs.wo);
        rho_new = diffuse_reflectance(S, bsdf, state_index<NS>(NS - 1), si.u, si.v);
        bs.weight = bs.pdf > 0.0f ? rho_new : v3(0.f);
      }
      active_surface = bs.pdf > 0.0f;
    }
    if (active_surface) {
      if (NS == 2) {
        // old state: f_old / pdf unless the sampled lobe is delta (null)
        V3 w_old;
        if (bs.is_null || !(F & kFeatSurfaces)) {
          w_old = bs.weight;
        } else {
          const V3 f_old = diffuse_eval(diffuse_reflectance(S, bsdf, RCV_STATE_OLD, si.u, si.v), wi, bs.wo);
          w_old = f_old / bs.pdf;
        }
        thr[0] *= w_old;
      }
      thr[NS - 1] *= bs.weight;
      const V3 d = bs.is_null ? ray.d : fr.to_world(bs.wo);
      ray = spawn_ray(si, d);
      needs_int = true;
      if (!bs.is_null) {
        depth += 1;
        last_p = si.p;
        last_pdf = bs.pdf;
        cnt.add(kCntSurfaceHits, 1);
      }
      if (!bs.is_null) specular_chain = false;  // diffuse lobe is smooth (:376-377)
      const ShapeInfo sh = S.shapes[si.shape];
      if (sh.interior != sh.exterior) medium = dot(si.n, ray.d) > 0.0f ? sh.exterior : sh.interior;
    }
    return alive && (active_surface || active_medium);
  }
};

// The camera samples [s0, s1) of one pixel; their radiance is added to acc in
// sample order. A lane starts its next sample in the bounce after its path
// ends.
template <int NS, int F>
__device__ void volpath_render_samples(const SceneView &S, const IntegArgs &I, const SensorInfo &sen, uint32_t seed,
                                       int px, int py, int64_t pixel, int spp, int s0, int s1, V3 acc[NS],
                                       const LdsStack &stack, LocalCounters &cnt) {
  VolPath<NS, false, F> t;
  bool live = false;
  int s = s0;
  while (true) {
    if (!live) {
      if (s >= s1) break;
      Pcg32 rng = lane_rng(seed, uint32_t(pixel * spp + s));
      const Ray r = camera_ray(sen, px, py, rng);
      t.begin(rng, r, cnt, false);
    }
    live = t.bounce(S, I, stack, cnt, nullptr);
    if (!live) {
#pragma unroll
      for (int k = 0; k < NS; ++k) acc[k] += t.L[k];
      ++s;
    }
  }
}

// Path-replay adjoint of the camera samples [s0, s1) of one pixel (NEW
// state): per sample a primal pass for its radiance L that records its
// collisions (scattered once L is known), or, for untaped kernels and tape
// overflows, the primal pass and then, when L is not zero, the replay pass
// with the same random numbers. dimage/dL = 1/spp (box filter) is folded
// into dL.
template <int F>
__device__ void volpath_adjoint_samples(const SceneView &S, const IntegArgs &I, const SensorInfo &sen, uint32_t seed,
                                        int px, int py, int64_t pixel, int spp, int s0, int s1, V3 dL,
                                        const GradTargets *G, const LdsStack &stack, LocalCounters &cnt,
                                        HomSum *hom_acc) {
  VolPath<1, true, F> t;
  t.L[0] = v3(0.f);  // read (unused) when the first sample starts
  t.init_hom();
  bool live = false, want_replay = false;
  int s = s0;
  while (true) {
    if (!live) {
      if (s >= s1) break;
      Pcg32 rng = lane_rng(seed, uint32_t(pixel * spp + s));
      const Ray r = camera_ray(sen, px, py, rng);
      t.begin_adjoint(rng, r, cnt, want_replay, dL, t.L[0]);
    }
    live = t.bounce(S, I, stack, cnt, G);
    if (!live) {
      want_replay = !t.end_adjoint(S, cnt, G);
      if (!want_replay) ++s;
    }
  }
  if constexpr ((F & kFeatHomog) != 0) {
    hom_acc->s += t.hom.s;
    hom_acc->a += t.hom.a;
  }
}

// Forward mode of the camera samples [s0, s1) of one pixel (NEW state): per
// sample the primal pass for its radiance L and, when L is not zero, the
// replay pass with the same random numbers, which gathers the tangents (in
// G) where the adjoint scatters; adds d(L)/d(theta) . tangent of each sample
// to acc.
template <int F>
__device__ void volpath_forward_samples(const SceneView &S, const IntegArgs &I, const SensorInfo &sen, uint32_t seed,
                                        int px, int py, int64_t pixel, int spp, int s0, int s1,
                                        const GradTargets *G, const LdsStack &stack, LocalCounters &cnt, V3 &acc) {
  VolPath<1, true, F, true> t;
  t.L[0] = v3(0.f);  // read (unused) when the first sample starts
  t.init_hom();
  bool live = false, want_replay = false;
  int s = s0;
  while (true) {
    if (!live) {
      if (s >= s1) break;
      Pcg32 rng = lane_rng(seed, uint32_t(pixel * spp + s));
      const Ray r = camera_ray(sen, px, py, rng);
      t.begin_adjoint(rng, r, cnt, want_replay, v3(0.f), t.L[0]);
    }
    live = t.bounce(S, I, stack, cnt, G);
    if (!live) {
      if (t.replay) acc += t.adj.dI;
      want_replay = !t.end_adjoint(S, cnt, G);
      if (!want_replay) ++s;
    }
  }
}

}  // namespace rcv
