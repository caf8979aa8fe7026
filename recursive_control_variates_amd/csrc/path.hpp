// Two-state surface path tracer (K1), its single-state form (plain prb) and
// the path-replay adjoint for bitmap-texture reflectances.
//
// Follows TwoStatePRBIntegrator.sample_twostates (plugins/twostatepath.py:
// 225-364): one path, two throughputs; emission with MIS against the previous
// BSDF sample (:281-290); emitter sampling with a shadow ray, the NEW BSDF's
// value and pdf and the OLD BSDF's value under one MIS weight (:292-312);
// BSDF sampled from NEW, OLD weighted by f_old/pdf unless the lobe is delta
// (:314-333); Russian roulette on the NEW throughput for depth >= rr_depth
// applied to both (:340-355). Random numbers per vertex: 2 (NEE), 1 + 2
// (BSDF), 1 (RR), drawn by every active lane as in the reference loop.
#pragma once

#include "volpath.hpp"

namespace rcv {

__device__ inline void scatter_texture(float *g, const TexFoot &f, V3 val) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float *t = g + 3 * f.idx[k];
    atomicAdd(t + 0, val.x * f.w[k]);
    atomicAdd(t + 1, val.y * f.w[k]);
    atomicAdd(t + 2, val.z * f.w[k]);
  }
}

// Shadow ray from a surface point towards p (Interaction::spawn_ray_to):
// offset origin, direction re-derived from the offset origin.
__device__ inline Ray spawn_ray_to(const SurfInt &si, V3 p) {
  Ray r;
  r.o = offset_point(si.p, si.n, p - si.p);
  V3 d = p - r.o;
  const float dist = norm(d);
  r.d = d / dist;
  r.maxt = dist * (1.0f - kShadowEps);
  return r;
}

// REPLAY: the adjoint's replay pass (adj->Lrem starts at the primal
// radiance) scattering into G; with FWD the replay gathers the tangents in G
// instead and adds d(L)/d(theta) . tangent to adj->dI.
template <int NS, bool REPLAY, bool FWD = false>
__device__ void path_sample(const SceneView &S, const IntegArgs &I, Pcg32 &rng, Ray ray, V3 L[NS],
                            const LdsStack &stack, LocalCounters &cnt, AdjState *adj, const GradTargets *G) {
  V3 beta[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    L[k] = v3(0.f);
    beta[k] = v3(1.f);
  }
  float eta = 1.0f;
  int depth = 0;
  V3 prev_p = v3(0.f);
  float prev_pdf = 1.0f;
  bool prev_delta = true;
  const int max_depth = I.max_depth < 0 ? 0x7fffffff : I.max_depth;
  cnt.add(kCntPaths, 1);

  while (true) {
    Hit h;
    trace<false>(S, ray, h, stack);
    cnt.add(kCntTraversals, 1);
    const SurfInt si = make_si(S, ray, h);

    // ---- direct emission (:281-290)
    {
      const DirSample ds = dir_sample_from_hit(S, si, prev_p, ray.d);
      const float em_pdf = prev_delta ? 0.0f : pdf_emitter_direction(S, ds);
      const float mis = mis_weight(prev_pdf, em_pdf);
      const V3 le = emitter_eval(S, ds.emitter, si, ray.d);
      if (any_nonzero(le)) {
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          const V3 c = beta[k] * mis * le;
          L[k] += c;
          if (REPLAY) adj->Lrem -= c;
        }
      }
    }

    bool active_next = (depth + 1 < max_depth) && si.valid;
    BsdfInfo bsdf;
    bsdf.type = -1;
    bsdf.texture_param = -1;
    Frame fr;
    V3 wi = v3(0.f, 0.f, 1.f);
    if (si.valid) {
      const int b = S.shapes[si.shape].bsdf;
      if (b >= 0) bsdf = S.bsdfs[b];
      fr = make_frame(si.n);
      wi = fr.to_local(-ray.d);
    }
    const bool smooth = bsdf.type == RCV_BSDF_DIFFUSE;
    V3 rho[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) rho[k] = v3(0.f);
    TexFoot tf;
    if (active_next && smooth) {
#pragma unroll
      for (int k = 0; k < NS; ++k) rho[k] = diffuse_reflectance(S, bsdf, state_index<NS>(k), si.u, si.v);
      if (REPLAY && bsdf.texture_param >= 0) texture_footprint(S.params[bsdf.texture_param], si.u, si.v, tf);
    }

    // ---- emitter sampling with visibility (:292-312)
    {
      const float u1 = rng.next_1d();
      const float u2 = rng.next_1d();
      bool active_em = active_next && smooth;
      V3 c[NS];
#pragma unroll
      for (int k = 0; k < NS; ++k) c[k] = v3(0.f);
      if (active_em) {
        DirSample ds;
        V3 em_w = sample_emitter_direction(S, si.p, u1, u2, ds);
        active_em = ds.pdf != 0.0f;
        if (active_em && any_nonzero(em_w)) {
          const Ray sr = spawn_ray_to(si, ds.p);
          Hit sh;
          cnt.add(kCntTraversals, 1);
          if (trace<true>(S, sr, sh, stack)) em_w = v3(0.f);
        }
        if (active_em && any_nonzero(em_w)) {
          const V3 wo = fr.to_local(ds.d);
          const float pdf = diffuse_pdf(wi, wo);
          const float mis_em = mis_weight(ds.pdf, pdf);
#pragma unroll
          for (int k = 0; k < NS; ++k) c[k] = beta[k] * mis_em * diffuse_eval(rho[k], wi, wo) * em_w;
        }
      }
      if (FWD && active_next && smooth && bsdf.texture_param >= 0) {
        // the texture tangent at uv, gathered where the adjoint scatters
        if (const float *tt = G->g[bsdf.texture_param]) {
          const V3 tv = texture_eval(tt, tf);
          const V3 r = rho[NS - 1];
          adj->dI += adj->Lrem * v3(r.x > 0.f ? tv.x / r.x : 0.f, r.y > 0.f ? tv.y / r.y : 0.f,
                                    r.z > 0.f ? tv.z / r.z : 0.f);
        }
      } else if (REPLAY && active_next && smooth && bsdf.texture_param >= 0) {
        // NEE at this vertex and everything after it carry rho_new(uv).
        float *g = G->g[bsdf.texture_param];
        if (g && any_nonzero(adj->Lrem)) {
          const V3 dl = adj->dL * adj->Lrem;
          const V3 r = rho[NS - 1];
          const V3 gv = v3(r.x > 0.f ? dl.x / r.x : 0.f, r.y > 0.f ? dl.y / r.y : 0.f, r.z > 0.f ? dl.z / r.z : 0.f);
          scatter_texture(g, tf, gv);
          cnt.add(kCntGradAtomics, 12);
        }
      }
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        L[k] += c[k];
        if (REPLAY) adj->Lrem -= c[k];
      }
    }

    // ---- BSDF sampling (:316-338)
    {
      (void)rng.next_1d();
      const float a = rng.next_1d();
      const float b = rng.next_1d();
      BsdfSample bs;
      bs.pdf = 0.0f;
      bs.weight = v3(0.f);
      bs.is_null = false;
      bs.wo = v3(0.f, 0.f, 1.f);
      if (active_next) {
        if (bsdf.type == RCV_BSDF_NULL) {
          bs.wo = -wi;
          bs.pdf = 1.0f;
          bs.weight = v3(1.f);
          bs.is_null = true;
        } else if (smooth && wi.z > 0.0f) {
          bs.wo = square_to_cosine_hemisphere(a, b);
          bs.pdf = diffuse_pdf(wi, bs.wo);
          bs.weight = bs.pdf > 0.0f ? rho[NS - 1] : v3(0.f);
        }
      }
      prev_delta = bs.is_null;
      if (si.valid) {
        const V3 d = bs.is_null ? ray.d : fr.to_world(bs.wo);
        ray = spawn_ray(si, d);
      }
      if (NS == 2) {
        V3 w_old;
        if (bs.is_null) {
          w_old = bs.weight;
        } else {
          const V3 f_old = active_next && smooth ? diffuse_eval(rho[0], wi, bs.wo) : v3(0.f);
          w_old = bs.pdf > 0.0f ? f_old / bs.pdf : v3(0.f);
        }
        beta[0] *= w_old;
      }
      beta[NS - 1] *= bs.weight;
      prev_p = si.p;
      prev_pdf = bs.pdf;
    }

    // ---- stopping criterion (:342-358)
    const float beta_max = hmax(beta[NS - 1]);
    active_next = active_next && beta_max != 0.0f;
    const float rr_prob = fminf(beta_max * (eta * eta), 0.95f);
    const bool rr_active = depth >= I.rr_depth;
    if (rr_active) {
      const float r = 1.0f / rr_prob;
#pragma unroll
      for (int k = 0; k < NS; ++k) beta[k] = beta[k] * r;
    }
    const float urr = rng.next_1d();
    active_next = active_next && (!rr_active || urr < rr_prob);
    if (si.valid) {
      depth += 1;
      cnt.add(kCntSurfaceHits, 1);
    }
    if (!active_next) break;
  }
}

}  // namespace rcv
