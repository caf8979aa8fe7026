// Device-side scene services used by the render kernels: camera rays of the
// (batch) perspective sensor, BVH traversal with a per-thread stack in LDS,
// surface interactions, diffuse / null BSDFs, area and constant emitters,
// bitmap textures and density grids.
//
// These are the upstream services the two-state plugins call (sample_rays,
// ray_intersect, sample_emitter_direction, pdf_emitter_direction, BSDF
// sample/eval/pdf, grid lookups; plugins/twostatepath.py:197,272-316,
// plugins/twostatevolpath.py:201,286-300,393,415, plugins/twostatemedium.py:
// 22,73-94). Their conventions are this library's own (box reconstruction
// filter, cell-centred trilinear grids with clamping, bilinear textures with
// repeat wrapping) and are stated where they are defined.
#pragma once

#include "../../include/rcv.h"
#include "device.hpp"
#include "scene_internal.hpp"
#include "scene_types.hpp"

namespace rcv {

struct Ray {
  V3 o, d;
  float maxt;
};

RCV_HD V3 ray_at(const Ray &r, float t) { return r.o + r.d * t; }

struct Hit {
  float t;
  int slot;  // BVH primitive slot, -1: miss
  float b1, b2;
};

// Surface interaction (geometric normal used as the shading normal).
struct SurfInt {
  bool valid;
  float t;
  V3 p, n;
  float u, v;
  int shape;
};

// ------------------------------------------------------------------ camera
// Film pixel (px, py) of a batch sensor; the view is px / (film_w / n_views).
// Sample position inside the pixel is jittered by one 2-D sample (box filter:
// the sample is splatted to its own pixel only). Camera space: x = left,
// y = up, z = forward; film u grows to the right, v downwards.
__device__ inline Ray camera_ray(const SensorInfo &sen, int px, int py, Pcg32 &rng) {
  const int view_w = sen.film_w / sen.n_views;
  const int view = px / view_w;
  const int lx = px - view * view_w;
  const float jx = rng.next_1d();
  const float jy = rng.next_1d();
  const float fx = (float(lx) + jx) / float(view_w);
  const float fy = (float(py) + jy) / float(sen.film_h);
  const ViewInfo &vw = sen.views[view];
  const float aspect = float(view_w) / float(sen.film_h);
  const float cx = (1.0f - 2.0f * fx) * vw.tan_half;
  const float cy = (1.0f - 2.0f * fy) * vw.tan_half / aspect;
  const V3 left = v3(vw.left[0], vw.left[1], vw.left[2]);
  const V3 up = v3(vw.up[0], vw.up[1], vw.up[2]);
  const V3 fwd = v3(vw.fwd[0], vw.fwd[1], vw.fwd[2]);
  Ray r;
  r.o = v3(vw.origin[0], vw.origin[1], vw.origin[2]);
  r.d = normalize(left * cx + up * cy + fwd);
  r.maxt = kInf;
  return r;
}

// ------------------------------------------------------------ primitives
// Moller-Trumbore ray/triangle test (1997).
__device__ inline bool hit_triangle(const float4 *tris, int slot, const Ray &r, float tmax, float &t,
                                    float &b1, float &b2) {
  const float4 a = tris[3 * slot + 0];
  const float4 b = tris[3 * slot + 1];
  const float4 c = tris[3 * slot + 2];
  const V3 v0 = v3(a.x, a.y, a.z), e1 = v3(b.x, b.y, b.z), e2 = v3(c.x, c.y, c.z);
  const V3 pv = cross(r.d, e2);
  const float det = dot(e1, pv);
  if (det == 0.0f) return false;
  const float inv = 1.0f / det;
  const V3 tv = r.o - v0;
  const float u = dot(tv, pv) * inv;
  if (!(u >= 0.0f && u <= 1.0f)) return false;
  const V3 qv = cross(tv, e1);
  const float v = dot(r.d, qv) * inv;
  if (!(v >= 0.0f && u + v <= 1.0f)) return false;
  const float tt = dot(e2, qv) * inv;
  if (!(tt > 0.0f && tt < tmax)) return false;
  t = tt;
  b1 = u;
  b2 = v;
  return true;
}

// Ray/oriented-box test: slabs of [-1, 1]^3 in the box's object space, where
// the ray keeps its parameter t. Returns the first crossing with t in (0,
// tmax) -- the entry face from outside, the exit face from inside -- as face
// 2k (+k) or 2k + 1 (-k). A NaN slab bound (ray parallel to and on a face
// plane) leaves the interval unchanged, as in hit_box.
__device__ inline bool hit_boxprim(const BoxPrim &B, const Ray &r, float tmax, float &t, int &face) {
  const float *m = B.inv;
  const float o[3] = {fmaf(m[0], r.o.x, fmaf(m[1], r.o.y, fmaf(m[2], r.o.z, m[3]))),
                      fmaf(m[4], r.o.x, fmaf(m[5], r.o.y, fmaf(m[6], r.o.z, m[7]))),
                      fmaf(m[8], r.o.x, fmaf(m[9], r.o.y, fmaf(m[10], r.o.z, m[11])))};
  const float d[3] = {fmaf(m[0], r.d.x, fmaf(m[1], r.d.y, m[2] * r.d.z)),
                      fmaf(m[4], r.d.x, fmaf(m[5], r.d.y, m[6] * r.d.z)),
                      fmaf(m[8], r.d.x, fmaf(m[9], r.d.y, m[10] * r.d.z))};
  float t0 = -kInf, t1 = kInf;
  int f0 = -1, f1 = -1;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float inv = 1.0f / d[k];
    float ta = (-1.0f - o[k]) * inv, tb = (1.0f - o[k]) * inv;
    int fa = 2 * k + 1, fb = 2 * k;
    if (ta > tb) {
      const float tmp = ta;
      ta = tb;
      tb = tmp;
      fa = 2 * k;
      fb = 2 * k + 1;
    }
    if (ta > t0) {
      t0 = ta;
      f0 = fa;
    }
    if (tb < t1) {
      t1 = tb;
      f1 = fb;
    }
  }
  if (!(t0 <= t1)) return false;
  if (t0 > 0.0f && t0 < tmax && f0 >= 0) {
    t = t0;
    face = f0;
    return true;
  }
  if (t1 > 0.0f && t1 < tmax && f1 >= 0) {
    t = t1;
    face = f1;
    return true;
  }
  return false;
}

// Ray/sphere test in the numerically robust form of Haines et al., "Precision
// improvements for ray/sphere intersection" (Ray Tracing Gems, 2019).
__device__ inline bool hit_sphere(float4 s, const Ray &r, float tmax, float &t) {
  const V3 c = v3(s.x, s.y, s.z);
  const V3 f = r.o - c;
  const float bp = -dot(f, r.d);
  const V3 l = f + r.d * bp;
  const float disc = s.w * s.w - dot(l, l);
  if (disc < 0.0f) return false;
  const float cc = dot(f, f) - s.w * s.w;
  const float q = bp + copysignf(sqrtf(disc), bp);
  float t0 = cc / q, t1 = q;
  if (t0 > t1) {
    const float tmp = t0;
    t0 = t1;
    t1 = tmp;
  }
  if (t0 > 0.0f && t0 < tmax) {
    t = t0;
    return true;
  }
  if (t1 > 0.0f && t1 < tmax) {
    t = t1;
    return true;
  }
  return false;
}

__device__ inline bool hit_box(const BVHNode &n, V3 o, V3 inv_d, float tmax, float &tnear) {
  float t0 = 0.0f, t1 = tmax;
  const float lo[3] = {n.lo[0], n.lo[1], n.lo[2]};
  const float hi[3] = {n.hi[0], n.hi[1], n.hi[2]};
  const float oo[3] = {o.x, o.y, o.z};
  const float id[3] = {inv_d.x, inv_d.y, inv_d.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float ta = (lo[k] - oo[k]) * id[k];
    float tb = (hi[k] - oo[k]) * id[k];
    if (ta > tb) {
      const float tmp = ta;
      ta = tb;
      tb = tmp;
    }
    // NaN (0 * inf on a slab plane) leaves the interval unchanged.
    t0 = ta > t0 ? ta : t0;
    t1 = tb < t1 ? tb : t1;
  }
  tnear = t0;
  return t0 <= t1;
}

// Per-thread traversal stack in LDS, strided by the block size so that the
// 64 lanes of a wave hit 64 distinct banks, and after it the thread's
// `nslots` words of per-lane storage (slot(j), same stride).
struct LdsStack {
  int *base;
  int stride;
  int *slots;
  int nslots;
  __device__ int &operator[](int i) const { return base[i * stride]; }
  __device__ int &slot(int j) const { return slots[j * stride]; }
};

// The thread's view of a launch's dynamic LDS (kLdsWordsPerThread words per
// thread) for scene S: its stack, then its slots.
__device__ inline LdsStack lds_of(const SceneView &S, int *lds) {
  const int tid = int(threadIdx.x), stride = int(blockDim.x);
  return LdsStack{lds + tid, stride, lds + S.lds_stack * stride + tid, S.lds_slots};
}

// Closest hit (any_hit = false) or occlusion test (any_hit = true) against
// [0, r.maxt). Stack depth never exceeds the BVH depth (SceneView::lds_stack).
template <bool ANY_HIT>
__device__ inline bool trace(const SceneView &S, const Ray &r, Hit &h, const LdsStack &stack) {
  const V3 inv_d = v3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
  float tbest = r.maxt;
  h.slot = -1;
  h.t = kInf;
  int sp = 0;
  int node = 0;
  float tn;
  if (!hit_box(S.nodes[0], r.o, inv_d, tbest, tn)) return false;
  while (true) {
    const BVHNode n = S.nodes[node];
    if (n.count > 0) {
      for (int i = 0; i < n.count; ++i) {
        const int slot = n.first + i;
        const int pid = S.prim_ids[slot];
        float t, b1 = 0.f, b2 = 0.f;
        bool hit;
        if (pid < S.n_tris) {
          hit = hit_triangle(S.tris, slot, r, tbest, t, b1, b2);
        } else if (pid < S.n_tris + S.n_sph) {
          hit = hit_sphere(S.spheres[pid - S.n_tris], r, tbest, t);
        } else {
          int face = 0;
          hit = hit_boxprim(S.boxes[pid - S.n_tris - S.n_sph], r, tbest, t, face);
          b1 = float(face);
        }
        if (hit) {
          tbest = t;
          h.t = t;
          h.slot = slot;
          h.b1 = b1;
          h.b2 = b2;
          if (ANY_HIT) return true;
        }
      }
    } else {
      const BVHNode a = S.nodes[n.first];
      const BVHNode b = S.nodes[n.first + 1];
      float ta, tb;
      const bool ha = hit_box(a, r.o, inv_d, tbest, ta);
      const bool hb = hit_box(b, r.o, inv_d, tbest, tb);
      if (ha && hb) {
        const bool a_first = ta <= tb;
        stack[sp++] = a_first ? n.first + 1 : n.first;
        node = a_first ? n.first : n.first + 1;
        continue;
      }
      if (ha) {
        node = n.first;
        continue;
      }
      if (hb) {
        node = n.first + 1;
        continue;
      }
    }
    // pop, skipping nodes that the current closest hit has culled
    bool found = false;
    while (sp > 0) {
      const int cand = stack[--sp];
      if (hit_box(S.nodes[cand], r.o, inv_d, tbest, tn)) {
        node = cand;
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  return h.slot >= 0;
}

__device__ inline SurfInt make_si(const SceneView &S, const Ray &r, const Hit &h) {
  SurfInt si;
  si.valid = h.slot >= 0;
  si.t = si.valid ? h.t : kInf;
  si.u = si.v = 0.0f;
  si.shape = -1;
  si.p = v3(0.f);
  si.n = v3(0.f, 0.f, 1.f);
  if (!si.valid) return si;
  const int pid = S.prim_ids[h.slot];
  if (pid < S.n_tris) {
    const float4 a = S.tris[3 * h.slot + 0];
    const float4 b = S.tris[3 * h.slot + 1];
    const float4 c = S.tris[3 * h.slot + 2];
    const V3 e1 = v3(b.x, b.y, b.z), e2 = v3(c.x, c.y, c.z);
    si.p = v3(a.x, a.y, a.z) + e1 * h.b1 + e2 * h.b2;
    si.n = normalize(cross(e1, e2));
    si.shape = __builtin_bit_cast(int, a.w);
    const float2 uv0 = S.tri_uv[3 * h.slot], uv1 = S.tri_uv[3 * h.slot + 1], uv2 = S.tri_uv[3 * h.slot + 2];
    const float b0 = 1.0f - h.b1 - h.b2;
    si.u = uv0.x * b0 + uv1.x * h.b1 + uv2.x * h.b2;
    si.v = uv0.y * b0 + uv1.y * h.b1 + uv2.y * h.b2;
  } else if (pid >= S.n_tris + S.n_sph) {
    // box: the face's plane point and normal; uv as the 'cube' mesh lays it
    // out per face (two triangles per face, corners (0,0) (1,0) (1,1) (0,1))
    const BoxPrim &B = S.boxes[pid - S.n_tris - S.n_sph];
    const int face = int(h.b1);
    si.p = ray_at(r, h.t);
    si.n = v3(B.n[3 * face], B.n[3 * face + 1], B.n[3 * face + 2]);
    si.shape = B.shape;
    const float *m = B.inv;
    const float qx = 0.5f * (fmaf(m[0], si.p.x, fmaf(m[1], si.p.y, fmaf(m[2], si.p.z, m[3]))) + 1.0f);
    const float qy = 0.5f * (fmaf(m[4], si.p.x, fmaf(m[5], si.p.y, fmaf(m[6], si.p.z, m[7]))) + 1.0f);
    const float qz = 0.5f * (fmaf(m[8], si.p.x, fmaf(m[9], si.p.y, fmaf(m[10], si.p.z, m[11]))) + 1.0f);
    const int k = face >> 1;
    const bool neg = face & 1;
    // +x: (y, z)  -x: (z, y)  +y: (z, x)  -y: (x, z)  +z: (x, y)  -z: (y, x)
    const float a0 = k == 0 ? qy : (k == 1 ? qz : qx);
    const float a1 = k == 0 ? qz : (k == 1 ? qx : qy);
    si.u = neg ? a1 : a0;
    si.v = neg ? a0 : a1;
  } else {
    const int s = pid - S.n_tris;
    const float4 sp = S.spheres[s];
    const V3 c = v3(sp.x, sp.y, sp.z);
    const V3 p = ray_at(r, h.t);
    si.n = normalize(p - c);
    si.p = c + si.n * sp.w;
    si.shape = S.sphere_shape[s];
    si.u = (atan2f(si.n.y, si.n.x) + kPi) * (0.5f * kInvPi);
    si.v = acosf(fminf(1.0f, fmaxf(-1.0f, si.n.z))) * kInvPi;
  }
  return si;
}

__device__ inline Ray spawn_ray(const SurfInt &si, V3 d) {
  Ray r;
  r.o = offset_point(si.p, si.n, d);
  r.d = d;
  r.maxt = kInf;
  return r;
}

// --------------------------------------------------------------- textures
// Bilinear lookup of an (ny, nx, 3) texture with repeat wrapping; texel
// (row j, column i) is centred at uv = ((i + 0.5)/nx, (j + 0.5)/ny).
struct TexFoot {
  int idx[4];
  float w[4];
};

__device__ inline void texture_footprint(const ParamBind &P, float u, float v, TexFoot &f) {
  const float x = u * float(P.nx) - 0.5f;
  const float y = v * float(P.ny) - 0.5f;
  const float fx0 = floorf(x), fy0 = floorf(y);
  const float tx = x - fx0, ty = y - fy0;
  auto wrap = [](int i, int n) {
    int m = i % n;
    return m < 0 ? m + n : m;
  };
  const int x0 = wrap(int(fx0), P.nx), x1 = wrap(int(fx0) + 1, P.nx);
  const int y0 = wrap(int(fy0), P.ny), y1 = wrap(int(fy0) + 1, P.ny);
  f.idx[0] = y0 * P.nx + x0;
  f.idx[1] = y0 * P.nx + x1;
  f.idx[2] = y1 * P.nx + x0;
  f.idx[3] = y1 * P.nx + x1;
  f.w[0] = (1.0f - tx) * (1.0f - ty);
  f.w[1] = tx * (1.0f - ty);
  f.w[2] = (1.0f - tx) * ty;
  f.w[3] = tx * ty;
}

__device__ inline V3 texture_eval(const float *data, const TexFoot &f) {
  V3 r = v3(0.f);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float *t = data + 3 * f.idx[k];
    r += v3(t[0], t[1], t[2]) * f.w[k];
  }
  return r;
}

// ------------------------------------------------------------------ grids
// Trilinear lookup of an (nz, ny, nx, ch) grid spanning [lo, hi]; voxel
// (i, j, k) is centred at lo + (i + 0.5, j + 0.5, k + 0.5) / n * (hi - lo);
// indices are clamped to the grid.
struct GridFoot {
  int base;            // index of corner (x0, y0, z0)
  int dx, dy, dz;      // offsets to the +1 corners (0 where clamped)
  float tx, ty, tz;
  __device__ float w(int c) const {
    return ((c & 1) ? tx : 1.0f - tx) * ((c & 2) ? ty : 1.0f - ty) * ((c & 4) ? tz : 1.0f - tz);
  }
  __device__ int idx(int c) const { return base + ((c & 1) ? dx : 0) + ((c & 2) ? dy : 0) + ((c & 4) ? dz : 0); }
};

__device__ inline void grid_footprint(const ParamBind &P, const MediumInfo &M, V3 p, GridFoot &g) {
  const float lx = (p.x - M.lo[0]) / (M.hi[0] - M.lo[0]);
  const float ly = (p.y - M.lo[1]) / (M.hi[1] - M.lo[1]);
  const float lz = (p.z - M.lo[2]) / (M.hi[2] - M.lo[2]);
  const float x = lx * float(P.nx) - 0.5f, y = ly * float(P.ny) - 0.5f, z = lz * float(P.nz) - 0.5f;
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  g.tx = x - fx;
  g.ty = y - fy;
  g.tz = z - fz;
  const int ix = int(fx), iy = int(fy), iz = int(fz);
  const int x0 = min(max(ix, 0), P.nx - 1), x1 = min(max(ix + 1, 0), P.nx - 1);
  const int y0 = min(max(iy, 0), P.ny - 1), y1 = min(max(iy + 1, 0), P.ny - 1);
  const int z0 = min(max(iz, 0), P.nz - 1), z1 = min(max(iz + 1, 0), P.nz - 1);
  g.base = (z0 * P.ny + y0) * P.nx + x0;
  g.dx = x1 - x0;
  g.dy = (y1 - y0) * P.nx;
  g.dz = (z1 - z0) * P.nx * P.ny;
}

__device__ inline float grid_eval1(const float *data, const GridFoot &g) {
  // Interpolate along x, then y, then z.
  const float c000 = data[g.base], c100 = data[g.base + g.dx];
  const float c010 = data[g.base + g.dy], c110 = data[g.base + g.dy + g.dx];
  const float c001 = data[g.base + g.dz], c101 = data[g.base + g.dz + g.dx];
  const float c011 = data[g.base + g.dz + g.dy], c111 = data[g.base + g.dz + g.dy + g.dx];
  const float x00 = c000 + (c100 - c000) * g.tx, x10 = c010 + (c110 - c010) * g.tx;
  const float x01 = c001 + (c101 - c001) * g.tx, x11 = c011 + (c111 - c011) * g.tx;
  const float y0 = x00 + (x10 - x00) * g.ty, y1 = x01 + (x11 - x01) * g.ty;
  return y0 + (y1 - y0) * g.tz;
}

__device__ inline V3 grid_eval3(const float *data, const GridFoot &g) {
  V3 r = v3(0.f);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float *t = data + 3 * g.idx(c);
    r += v3(t[0], t[1], t[2]) * g.w(c);
  }
  return r;
}

// --------------------------------------------------------------- emitters
struct DirSample {
  V3 p, n, d;
  float dist, pdf;
  int emitter;
};

__device__ inline V3 emitter_rgb(const EmitterInfo &e) { return v3(e.rgb[0], e.rgb[1], e.rgb[2]); }

// Radiance leaving an emitter surface towards -ray.d (area: front side only)
// or arriving from the environment for a miss.
__device__ inline V3 emitter_eval(const SceneView &S, int em, const SurfInt &si, V3 ray_d) {
  if (em < 0) return v3(0.f);
  const EmitterInfo e = S.emitters[em];
  if (e.type == RCV_EMITTER_CONSTANT) return emitter_rgb(e);
  return dot(si.n, ray_d) < 0.0f ? emitter_rgb(e) : v3(0.f);
}

// Emitter hit by `si` (environment for misses), -1 if none.
__device__ inline int si_emitter(const SceneView &S, const SurfInt &si) {
  // The environment id is taken as a value (readfirstlane of a uniform
  // kernel argument): a select between a pointer into the shape table and a
  // pointer to the argument would copy the argument to scratch memory.
  const int env = __builtin_amdgcn_readfirstlane(S.env_emitter);
  if (!si.valid) return env;
  return S.shapes[si.shape].emitter;
}

// Solid-angle density of choosing direction ds.d towards emitter ds.emitter
// from ref (uniform emitter selection).
__device__ inline float pdf_emitter_direction(const SceneView &S, const DirSample &ds) {
  if (ds.emitter < 0) return 0.0f;
  const EmitterInfo e = S.emitters[ds.emitter];
  const float sel = 1.0f / float(S.n_emitters);
  if (e.type == RCV_EMITTER_CONSTANT) return kInv4Pi * sel;
  const float dp = dot(ds.d, ds.n);
  if (!(dp < 0.0f)) return 0.0f;
  return e.inv_area * ds.dist * ds.dist / fabsf(dp) * sel;
}

// DirectionSample3f(scene, si, ref) for a surface interaction reached by
// `ray_d` from reference point ref_p.
__device__ inline DirSample dir_sample_from_hit(const SceneView &S, const SurfInt &si, V3 ref_p, V3 ray_d) {
  DirSample ds;
  ds.emitter = si_emitter(S, si);
  ds.p = si.p;
  ds.n = si.n;
  if (si.valid) {
    const V3 dv = si.p - ref_p;
    ds.dist = norm(dv);
    ds.d = dv / ds.dist;
  } else {
    ds.d = ray_d;
    ds.dist = kInf;
  }
  ds.pdf = 0.0f;
  return ds;
}

// Samples an emitter direction from reference point ref_p (uniform emitter
// selection, sample component u1 reused). Returns radiance / pdf (zero when
// the sampled point faces away); visibility is NOT tested.
__device__ inline V3 sample_emitter_direction(const SceneView &S, V3 ref_p, float u1, float u2, DirSample &ds) {
  ds.emitter = -1;
  ds.pdf = 0.0f;
  ds.dist = 0.0f;
  ds.d = v3(0.f, 0.f, 1.f);
  ds.p = ref_p;
  ds.n = v3(0.f, 0.f, 1.f);
  const int n_em = S.n_emitters;
  if (n_em == 0) return v3(0.f);
  int idx = min(int(u1 * float(n_em)), n_em - 1);
  u1 = u1 * float(n_em) - float(idx);
  const EmitterInfo e = S.emitters[idx];
  const float sel = 1.0f / float(n_em);
  ds.emitter = idx;
  if (e.type == RCV_EMITTER_CONSTANT) {
    const V3 d = square_to_uniform_sphere(u1, u2);
    const V3 c = v3(S.bs_center[0], S.bs_center[1], S.bs_center[2]);
    const float radius = fmaxf(S.bs_radius, norm(ref_p - c));
    ds.dist = 2.0f * radius;
    ds.d = d;
    ds.p = ref_p + d * ds.dist;
    ds.n = -d;
    ds.pdf = kInv4Pi * sel;
    return emitter_rgb(e) * (4.0f * kPi * float(n_em));
  }
  // area: choose a triangle by area (sample reuse), then a uniform point on it
  int lo = 0, hi = e.tri_count - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (S.em_cdf[e.tri_begin + mid] < u1)
      lo = mid + 1;
    else
      hi = mid;
  }
  const float c0 = lo > 0 ? S.em_cdf[e.tri_begin + lo - 1] : 0.0f;
  const float c1 = S.em_cdf[e.tri_begin + lo];
  u1 = fminf(fmaxf((u1 - c0) / fmaxf(c1 - c0, 1e-30f), 0.0f), 0.99999994f);
  const int slot = S.em_tris[e.tri_begin + lo];
  const float4 a = S.tris[3 * slot + 0], b = S.tris[3 * slot + 1], c = S.tris[3 * slot + 2];
  const V3 e1 = v3(b.x, b.y, b.z), e2 = v3(c.x, c.y, c.z);
  const float su = sqrtf(1.0f - u1);
  const float bb1 = 1.0f - su, bb2 = su * u2;
  ds.p = v3(a.x, a.y, a.z) + e1 * bb1 + e2 * bb2;
  ds.n = normalize(cross(e1, e2));
  const V3 dv = ds.p - ref_p;
  const float d2 = dot(dv, dv);
  ds.dist = sqrtf(d2);
  ds.d = dv / ds.dist;
  const float dp = dot(ds.d, ds.n);
  ds.pdf = e.inv_area * d2 / fabsf(dp) * sel;
  // a point on the emitter's back or in its plane is not a sample: pdf 0,
  // as Mitsuba reports it (the callers drop lanes with pdf 0,
  // plugins/twostatevolpath.py:395-397)
  if (!(dp < 0.0f) || !(ds.pdf > 0.0f) || !(ds.pdf < kInf)) {
    ds.pdf = 0.0f;
    return v3(0.f);
  }
  return emitter_rgb(e) / ds.pdf;
}

// ------------------------------------------------------------------ BSDFs
struct BsdfSample {
  V3 wo;        // local
  float pdf;
  V3 weight;    // f * cos / pdf
  bool is_null;
};

// Reflectance of a diffuse BSDF in state `state` at uv.
__device__ inline V3 diffuse_reflectance(const SceneView &S, const BsdfInfo &b, int state, float u, float v) {
  if (b.texture_param < 0) return v3(b.rgb[0], b.rgb[1], b.rgb[2]);
  const ParamBind &P = S.params[b.texture_param];
  TexFoot f;
  texture_footprint(P, u, v, f);
  return texture_eval(P.data[state], f);
}

// Diffuse eval (f * cos) and pdf for local directions wi, wo.
__device__ inline V3 diffuse_eval(V3 rho, V3 wi, V3 wo) {
  if (!(wi.z > 0.0f && wo.z > 0.0f)) return v3(0.f);
  return rho * (kInvPi * wo.z);
}
__device__ inline float diffuse_pdf(V3 wi, V3 wo) {
  if (!(wi.z > 0.0f && wo.z > 0.0f)) return 0.0f;
  return kInvPi * wo.z;
}

}  // namespace rcv
