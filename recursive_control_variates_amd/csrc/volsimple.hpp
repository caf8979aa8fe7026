// Simplified volumetric path tracer with differential ratio tracking (DRT)
// for its adjoint: the `volpathsimple` integrator (plugins/volpathsimple.py:
// 23-668), the reference's adjoint for volume optimisations
// (experiments.py:131). NEW state only (it renders one state; a two-state
// medium is read through its NEW state, as the adapter delegates,
// plugins/twostatemedium.py:117-184).
//
// Scene restrictions as in the reference (:24-30): one medium inside a
// closed null-BSDF boundary, no other surfaces, the environment as the only
// emitter (checked at launch).
//
// Primal (Simple<false>): per path, cross the boundary, then repeat: Russian
// roulette; delta tracking until a real collision or the boundary
// (sample_real_interaction, :315-367); albedo; NEE with ratio tracking
// (:370-445); isotropic phase sampling; finally the environment with MIS.
//
// Adjoint (Simple<true>), the same random stream as the primal plus a
// second stream for gradient-only decisions (:108-121), with L the primal
// radiance still to be gathered (`Lrem`):
//  * transmittance: after each segment, d/dsigma_t at 4 uniform points of
//    the segment weighted by -dL . Lrem * length / 4 (:559-583);
//  * in-scattering, collision term: at a real collision, d(sigma_t albedo)
//    weighted by mis * dL . Lrem / (sigma_t albedo) with mis = s^2 / (1 + s^2),
//    s = sigma_t (:150-180);
//  * in-scattering, DRT term: one segment per path chosen by a reservoir over
//    the segments with the throughput as weight (:743-779); along it a
//    distance t' drawn with density proportional to the majorant
//    transmittance, weighted by a ratio-tracking estimate of T(t') over its
//    density, and d(sigma_t albedo)(t') weighted by 1 / (1 + s^2) times the
//    incident radiance of a recursive path from t' (:480-556, 586-633);
//  * NEE transmittance: the ratio-tracking replay of plugins/
//    volpathsimple.py:450-477 (-dL . c scale / sigma_n per step).
#pragma once

#include "volpath.hpp"

namespace rcv {

// rcv_integrator_desc.flags for volpathsimple: set bits disable a feature
// (0 = the reference's defaults, plugins/volpathsimple.py:33-48).
constexpr int kSimpleNoNee = RCV_SIMPLE_NO_NEE;
constexpr int kSimpleNoDrt = RCV_SIMPLE_NO_DRT;
constexpr int kSimpleNoDrtSubsampling = RCV_SIMPLE_NO_DRT_SUBSAMPLING;
constexpr int kSimpleNoDrtMis = RCV_SIMPLE_NO_DRT_MIS;

// Tentative collisions with the majorant until a real one (accepted with
// probability sigma_t / majorant) or the end of the segment
// (sample_real_interaction, plugins/volpathsimple.py:315-367). `ev.t` is
// measured from the segment's origin.
template <int F>
__device__ inline void simple_real_collision(const SceneView &S, int medium, Ray ray, Pcg32 &rng,
                                             LocalCounters &cnt, MediumEv<1> &ev) {
  float run_t = 0.0f;
  while (true) {
    sample_medium<1, F>(S, medium, ray, rng.next_1d(), ev);
    cnt.add(kCntCollisions, 1);
    if (!ev.valid) return;
    const float r = ev.maj != 0.0f ? ev.st[0] / ev.maj : 0.0f;
    if (!(rng.next_1d() >= r)) {
      ev.t += run_t;
      return;
    }
    cnt.add(kCntNullEvents, 1);
    ray.o = ev.p;
    ray.maxt -= ev.t;
    run_t += ev.t;
  }
}

// Ratio-tracking transmittance over [0, tmax] of `ray` inside the medium
// (estimate_transmittance, plugins/volpathsimple.py:450-491). GRAD: scatter
// -adj * scale / sigma_n per tentative collision (the replay of :476-482).
// rec (adjoint, first pass): the collisions that replay would scatter at
// (point, sigma_n; sigma_n = 0 where it would not), so the caller can scatter
// them once adj is known.
template <int F, bool GRAD>
__device__ inline float simple_ratio_tracking(const SceneView &S, int medium, Ray ray, float tmax, Pcg32 &rng,
                                              LocalCounters &cnt, float adj, const GradTargets *G,
                                              NeeRecord *rec = nullptr) {
  float T = 1.0f;
  ray.maxt = tmax;
  while (true) {
    MediumEv<1> ev;
    sample_medium<1, F>(S, medium, ray, rng.next_1d(), ev);
    cnt.add(kCntNeeSteps, 1);
    if (!ev.valid) break;
    const float tr = ev.maj != 0.0f ? ev.sn[0] / ev.maj : ev.sn[0];
    if (GRAD && tr > 0.0f) {
      const MediumInfo &M = S.media[medium];
      float *g = G->g[M.sigma_param];
      if (g && adj != 0.0f) {
        scatter_grid(g, ev.foot, -adj * M.scale / ev.sn[0]);
        cnt.add(kCntGradAtomics, 8);
      }
    }
    if (rec) {
      unsigned cell;
      if (rec->n < kNeeRecords && pack_cell(ev.foot, medium, cell)) {
        rec->e[rec->n] = make_float4(ev.foot.tx, ev.foot.ty, ev.foot.tz,
                                     tr > 0.0f ? -S.media[medium].scale / ev.sn[0] : 0.0f);
        rec->cell[rec->n] = cell;
        rec->n++;
      } else {
        rec->n = kNeeRecords + 1;
      }
    }
    T *= tr;
    ray.o = ev.p;
    ray.maxt -= ev.t;
    if (T == 0.0f) break;
  }
  return T;
}

// Emitter sample from a medium point, then the transmittance to the medium
// boundary (sample_emitter, plugins/volpathsimple.py:415-445). Returns the
// emitter value over the direction pdf times the transmittance.
template <int F, bool GRAD>
__device__ inline V3 simple_sample_emitter(const SceneView &S, int medium, V3 p, Pcg32 &rng, DirSample &ds,
                                           const LdsStack &stack, LocalCounters &cnt, float adj,
                                           const GradTargets *G, NeeRecord *rec = nullptr) {
  const float u1 = rng.next_1d();
  const float u2 = rng.next_1d();
  V3 ev = sample_emitter_direction(S, p, u1, u2, ds);
  if (ds.pdf == 0.0f) return v3(0.f);
  Ray r;
  r.o = p;
  r.d = ds.d;
  r.maxt = kInf;
  Hit h;
  trace<false>(S, r, h, stack);
  cnt.add(kCntTraversals, 1);
  const SurfInt si = make_si(S, r, h);
  if (!si.valid) return v3(0.f);
  return ev * simple_ratio_tracking<F, GRAD>(S, medium, r, si.t, rng, cnt, adj, G, rec);
}

// Where a path starts: from the camera (cross the medium boundary first,
// reach_medium, :302-312) or inside the medium after a DRT collision
// (a recursive ray restored from a path state, :72-84, 618-633).
struct SimpleStart {
  bool from_camera;
  int depth;
  float last_pdf;
  SurfInt si;  // boundary ahead of a resumed ray
};

__device__ inline SimpleStart simple_camera_start() {
  SimpleStart st;
  st.from_camera = true;
  st.depth = 0;
  st.last_pdf = 1.0f;
  st.si.valid = false;
  st.si.t = kInf;
  st.si.shape = -1;
  st.si.p = v3(0.f);
  st.si.n = v3(0.f, 0.f, 1.f);
  st.si.u = st.si.v = 0.0f;
  return st;
}

// Seed of the gradient-only stream of an adjoint launch: TEA(seed, 2) (the
// reference seeds it from a number drawn from the main stream,
// plugins/volpathsimple.py:108-120; any stream independent of the main one
// serves, and a launch-level seed keeps it reproducible).
__host__ __device__ inline uint32_t simple_alt_seed(uint32_t seed) {
  uint32_t a = seed, b = 2u;
  tea(a, b);
  return a;
}

struct DrtReservoir {
  V3 wsum, cur;  // sum of weights seen; weight of the kept segment
  bool valid;
  int depth;
  Ray ray;
  SurfInt si;
};

template <int F>
__device__ V3 simple_recursive_li(const SceneView &S, const IntegArgs &I, int medium, V3 p, int depth,
                                  Pcg32 &alt, const LdsStack &stack, LocalCounters &cnt);
template <int F>
__device__ void simple_drt(const SceneView &S, const IntegArgs &I, int medium, Ray ray, const SurfInt &si,
                           int depth, V3 adjoint, Pcg32 &alt, const LdsStack &stack, LocalCounters &cnt,
                           const GradTargets *G);

// One camera sample (or resumed path) of volpathsimple. ADJ = false: returns
// the radiance. ADJ = true: adj->Lrem holds the primal radiance on entry;
// the gradients are scattered into G using the `alt` stream.
template <bool ADJ, int F>
__device__ V3 simple_sample(const SceneView &S, const IntegArgs &I, int medium, Pcg32 &rng, Ray ray,
                            const SimpleStart &start, const LdsStack &stack, LocalCounters &cnt, AdjState *adj,
                            Pcg32 *alt, const GradTargets *G) {
  const bool use_nee = !(I.flags & kSimpleNoNee);
  const bool use_drt = !(I.flags & kSimpleNoDrt);
  const bool drt_sub = !(I.flags & kSimpleNoDrtSubsampling);
  const bool drt_mis = !(I.flags & kSimpleNoDrtMis);
  const int max_depth = I.max_depth < 0 ? 0x7fffffff : I.max_depth;
  const MediumInfo &M = S.media[medium];
  V3 result = v3(0.f);  // primal radiance (ADJ: adj->Lrem plays this role)
  V3 thr = v3(1.f);
  int depth = start.depth;
  bool active = true, escaped = false, has_scattered = !start.from_camera;
  float last_pdf = start.last_pdf;
  SurfInt si = start.si;
  if (start.from_camera) {
    cnt.add(kCntPaths, 1);
    (void)rng.next_1d();  // colour channel (:86)
    Hit h;
    trace<false>(S, ray, h, stack);
    cnt.add(kCntTraversals, 1);
    const SurfInt s0 = make_si(S, ray, h);
    escaped = !s0.valid;
    active = s0.valid;
    si = s0;
    if (active) {
      ray = spawn_ray(s0, ray.d);
      trace<false>(S, ray, h, stack);
      cnt.add(kCntTraversals, 1);
      si = make_si(S, ray, h);
      active = si.valid;
      ray.maxt = si.valid ? si.t : kInf;
    }
  } else {
    active = depth < max_depth;
    ray.maxt = si.valid ? si.t : kInf;
  }
  V3 last_p = ray.o;
  (void)rng.next_1d();  // seed of the gradient-only stream (:108)
  DrtReservoir res;
  res.wsum = v3(0.f);
  res.cur = v3(0.f);
  res.valid = false;
  res.depth = 0;

  while (active) {
    // ---- Russian roulette (:127-132)
    const float q = fminf(hmax(thr), 0.99f);
    const bool perform_rr = depth > I.rr_depth;
    const float urr = rng.next_1d();
    if (!any_nonzero(thr) || (perform_rr && !(urr < q))) break;
    if (perform_rr) thr = thr * (1.0f / q);

    // ---- real collision or boundary (:137-145)
    MediumEv<1> mev;
    simple_real_collision<F>(S, medium, ray, rng, cnt, mev);
    const bool did_scatter = mev.valid;
    const bool did_escape = !mev.valid;
    has_scattered = has_scattered || did_scatter;
    const V3 albedo = did_scatter ? mev.alb : v3(1.f);

    if (ADJ) {
      // ---- in-scattering gradients (:150-180)
      if (use_drt) {
        if (drt_sub) {
          // reservoir over the segments, weight = throughput (:743-766)
          res.wsum += thr;
          const float u = alt->next_1d();
          const V3 ratio = v3(thr.x / res.wsum.x, thr.y / res.wsum.y, thr.z / res.wsum.z);
          if (u <= (ratio.x + ratio.y + ratio.z) * (1.0f / 3.0f)) {
            res.cur = thr;
            res.valid = true;
            res.depth = depth;
            res.ray = ray;
            res.si = si;
          }
        }
      }
      if (use_drt && !drt_sub) {
        // DRT on every segment (:502-556 without the reservoir)
        simple_drt<F>(S, I, medium, ray, si, depth, adj->dL * thr, *alt, stack, cnt, G);
      }
      {
        // collision term (sampling density sigma_t T), MIS-weighted against
        // DRT when both run
        if (did_scatter && (!use_drt || drt_mis) && mev.st[0] > 0.0f) {
          const float s2 = mev.st[0] * mev.st[0];
          const float mis = use_drt ? s2 / (1.0f + s2) : 1.0f;
          const V3 li = v3(adj->Lrem.x / fmaxf(1e-8f, albedo.x), adj->Lrem.y / fmaxf(1e-8f, albedo.y),
                           adj->Lrem.z / fmaxf(1e-8f, albedo.z));
          const V3 dl = adj->dL * li;
          float *gd = G->g[M.sigma_param];
          if (gd) {
            scatter_grid(gd, mev.foot, mis * dot(dl, albedo) * M.scale / mev.st[0]);
            cnt.add(kCntGradAtomics, 8);
          }
          if ((F & kFeatAlbedoGrid) && M.albedo_param >= 0 && G->g[M.albedo_param]) {
            GridFoot fa;
            grid_footprint(S.params[M.albedo_param], M, mev.p, fa);
            scatter_grid3(G->g[M.albedo_param], fa, dl * mis);
            cnt.add(kCntGradAtomics, 24);
          }
        }
      }
      // ---- transmittance gradients: 4 uniform points of the segment (:559-583)
      {
        const float interval = did_escape ? si.t : mev.t;
        const float lr = dot(adj->dL, adj->Lrem);
        float *gd = G->g[M.sigma_param];
        const ParamBind &P = S.params[M.sigma_param];
        for (int j = 0; j < 4; ++j) {
          const float tj = alt->next_1d() * interval;
          if (gd && lr != 0.0f && interval < kInf) {
            GridFoot f;
            grid_footprint(P, M, ray_at(ray, tj), f);
            scatter_grid(gd, f, -lr * M.scale * interval * 0.25f);
            cnt.add(kCntGradAtomics, 8);
          }
        }
      }
    }

    thr = thr * albedo;
    if (did_scatter) {
      depth += 1;
      cnt.add(kCntRealScatter, 1);
    }
    active = did_scatter && depth < max_depth;

    // ---- NEE (:191-201, 370-413)
    if (use_nee && did_scatter && active) {
      const Pcg32 nee_rng = rng;
      DirSample ds;
      NeeRecord rec;
      rec.n = 0;
      const V3 em = simple_sample_emitter<F, false>(S, medium, mev.p, rng, ds, stack, cnt, 0.0f, G,
                                                    ADJ ? &rec : nullptr);
      const float mis = mis_weight(ds.pdf, kInv4Pi);
      const V3 c = thr * em * (kInv4Pi * mis);
      if (ADJ) {
        adj->Lrem -= c;
        const float cs = dot(adj->dL, c);
        if (cs != 0.0f) {
          Pcg32 r2 = nee_rng;
          DirSample ds2;
          if (rec.n <= kNeeRecords) {
            // the replay's scatter-adds, from the recorded collisions
            RecMedium rm;
            rm.select(S, G, medium);
            for (int j = 0; rm.g && j < rec.n; ++j) {
              const float4 e = rec.e[j];
              if (e.w != 0.0f) {
                scatter_grid(rm.g, rm.foot(e, rec.cell[j]), cs * e.w);
                cnt.add(kCntGradAtomics, 8);
              }
            }
          } else {
            (void)simple_sample_emitter<F, true>(S, medium, mev.p, r2, ds2, stack, cnt, cs, G);
          }
        }
      } else {
        result += c;
      }
    }

    // ---- phase sampling, isotropic (:203-210)
    if (did_scatter) {
      (void)rng.next_1d();
      const float a = rng.next_1d();
      const float b = rng.next_1d();
      ray.o = mev.p;
      ray.d = square_to_uniform_sphere(a, b);
      last_p = mev.p;
      last_pdf = kInv4Pi;
    }

    // ---- next boundary; escaped rays cross it (:218-236)
    if (did_scatter || did_escape) {
      ray.maxt = kInf;
      Hit h;
      trace<false>(S, ray, h, stack);
      cnt.add(kCntTraversals, 1);
      si = make_si(S, ray, h);
      ray.maxt = si.valid ? si.t : kInf;
    }
    if (did_scatter && !si.valid) active = false;  // accidental escape: absorbed
    if (did_escape) {
      if (si.valid) ray = spawn_ray(si, ray.d);
      escaped = true;
    }
  }

  if (ADJ) {
    // ---- the reservoir's segment (:240-250)
    if (use_drt && drt_sub && res.valid) {
      const float dm = (res.cur.x + res.cur.y + res.cur.z) * (1.0f / 3.0f);
      const float wm = (res.wsum.x + res.wsum.y + res.wsum.z) * (1.0f / 3.0f);
      const V3 sw = dm != 0.0f ? res.cur * (wm / dm) : v3(0.f);
      simple_drt<F>(S, I, medium, res.ray, res.si, res.depth, sw * adj->dL, *alt, stack, cnt, G);
    }
    return adj->Lrem;
  }

  // ---- environment (:253-281)
  if (escaped) {
    ray.maxt = kInf;
    Hit h;
    trace<false>(S, ray, h, stack);
    cnt.add(kCntTraversals, 1);
    const SurfInt se = make_si(S, ray, h);
    const int em = si_emitter(S, se);
    if (em >= 0 && !(depth <= 0 && I.hide_emitters)) {
      // MIS against emitter sampling only when NEE runs (:262-275)
      const DirSample ds = dir_sample_from_hit(S, se, last_p, ray.d);
      const float em_pdf = has_scattered ? pdf_emitter_direction(S, ds) : 0.0f;
      const float mis = use_nee ? mis_weight(last_pdf, em_pdf) : 1.0f;
      result += thr * emitter_eval(S, em, se, ray.d) * mis;
    }
  }
  return result;
}

// In-scattering gradient along one segment by differential ratio tracking
// (backpropagate_scattering_drt, plugins/volpathsimple.py:502-556): a
// distance s in the medium part of the segment with density proportional to
// the majorant transmittance, p(s) = m e^{-m s} / (1 - e^{-m L}); a
// ratio-tracking estimate of T(s) over [0, s]; weight T / p, so
// E[weight f(s)] = int_0^L T(s) f(s) ds; f = d(sigma_t albedo) times the
// incident radiance of a recursive path from s, MIS weight 1 / (1 + sigma_t^2)
// against the collision term. (The upstream distance sampler of the DRT is in
// the absent Mitsuba fork; this is the published estimator, with the
// medium's single majorant.)
template <int F>
__device__ void simple_drt(const SceneView &S, const IntegArgs &I, int medium, Ray ray, const SurfInt &si,
                           int depth, V3 adjoint, Pcg32 &alt, const LdsStack &stack, LocalCounters &cnt,
                           const GradTargets *G) {
  const MediumInfo &M = S.media[medium];
  ray.maxt = si.valid ? si.t : kInf;
  float t0, t1;
  const bool hit = ray_aabb(ray, M.lo, M.hi, t0, t1);
  const float mint = fmaxf(0.0f, t0);
  const float maxt = fminf(ray.maxt, t1);
  const float maj = majorant_of(S, M.sigma_param, RCV_STATE_NEW) * M.scale;
  const float u = alt.next_1d();
  const float len = maxt - mint;
  if (!hit || !(len > 0.0f) || !(len < kInf) || !(maj > 0.0f) || !(maj < kInf)) return;
  const float norm = -expm1f(-maj * len);
  const float s = -log1pf(-u * norm) / maj;
  const float pdf = maj * expf(-maj * s) / norm;
  Ray seg;
  seg.o = ray_at(ray, mint);
  seg.d = ray.d;
  seg.maxt = kInf;
  const float T = simple_ratio_tracking<F, false>(S, medium, seg, s, alt, cnt, 0.0f, G);
  const float weight = pdf > 0.0f ? T / pdf : 0.0f;
  const V3 p = ray_at(ray, mint + s);
  const ParamBind &P = S.params[M.sigma_param];
  GridFoot foot;
  grid_footprint(P, M, p, foot);
  const float st = M.scale * grid_eval1(P.data[RCV_STATE_NEW], foot);
  V3 alb;
  GridFoot fa;
  if ((F & kFeatAlbedoGrid) && M.albedo_param >= 0) {
    const ParamBind &A = S.params[M.albedo_param];
    grid_footprint(A, M, p, fa);
    alb = grid_eval3(A.data[RCV_STATE_NEW], fa);
  } else {
    alb = v3(M.albedo[0], M.albedo[1], M.albedo[2]);
  }
  const V3 li = simple_recursive_li<F>(S, I, medium, p, depth, alt, stack, cnt);
  const float mis = (I.flags & kSimpleNoDrtMis) ? 1.0f : 1.0f / (1.0f + st * st);
  const V3 a = adjoint * li * (mis * weight);
  float *gd = G->g[M.sigma_param];
  if (gd && weight != 0.0f) {
    scatter_grid(gd, foot, dot(a, alb) * M.scale);
    cnt.add(kCntGradAtomics, 8);
  }
  if ((F & kFeatAlbedoGrid) && M.albedo_param >= 0 && G->g[M.albedo_param] && weight != 0.0f) {
    scatter_grid3(G->g[M.albedo_param], fa, a * st);
    cnt.add(kCntGradAtomics, 24);
  }
}

// Incident radiance at a medium point for DRT (sample_recursive,
// plugins/volpathsimple.py:586-633): NEE from the point (throughput 1), then
// a phase-sampled recursive path resumed at depth + 1.
template <int F>
__device__ V3 simple_recursive_li(const SceneView &S, const IntegArgs &I, int medium, V3 p, int depth,
                                  Pcg32 &alt, const LdsStack &stack, LocalCounters &cnt) {
  V3 result = v3(0.f);
  if (!(I.flags & kSimpleNoNee)) {
    DirSample ds;
    const V3 em = simple_sample_emitter<F, false>(S, medium, p, alt, ds, stack, cnt, 0.0f, nullptr);
    result += em * (kInv4Pi * mis_weight(ds.pdf, kInv4Pi));
  }
  (void)alt.next_1d();
  const float a = alt.next_1d();
  const float b = alt.next_1d();
  Ray rec;
  rec.o = p;
  rec.d = square_to_uniform_sphere(a, b);
  rec.maxt = kInf;
  Hit h;
  trace<false>(S, rec, h, stack);
  cnt.add(kCntTraversals, 1);
  SimpleStart st;
  st.from_camera = false;
  st.depth = depth + 1;
  st.last_pdf = kInv4Pi;
  st.si = make_si(S, rec, h);
  result += simple_sample<false, F>(S, I, medium, alt, rec, st, stack, cnt, nullptr, nullptr, nullptr);
  return result;
}

}  // namespace rcv
