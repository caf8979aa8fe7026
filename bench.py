#!/usr/bin/env python3
"""Benchmark of the recursive-control-variate inverse-rendering step on one
MI355X.

Workload (the heterogeneous-volume configuration): 64^3 float32 density grid
in the unit cube (target: 8 Gaussian blobs from seed 1, sigma_t scale 20),
albedo 0.8, isotropic phase, constant environment 1, 4 views of 256x256 on one
batch film, 32 spp primal and 32 spp adjoint, max_depth 16, grid initialised at
0.1, Adam lr 0.01, L1 loss, method cv_ps (two-state render + CV combine).

A step = scene update (two-state pointer rotation) + two-state render + fused
CV combine + loss/gradient + path-replay adjoint + Adam. ``value`` is
Msamples/s = (W*H*spp primal paths + W*H*spp_grad adjoint paths) / step time.
The JSON line also carries p50/p99 step ms, time-to-param-MSE 1e-3 for cv_ps
and the plain (baseline) loop, the roofline of the dominant kernel and a CPU
baseline (the float32 CPU restatement on host cores, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mse-steps", type=int, default=200, help="iterations per method for time-to-MSE (0: skip)")
    ap.add_argument("--mse-threshold", type=float, default=1e-3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--spp", type=int, default=32)
    ap.add_argument("--max-depth", type=int, default=16)
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--adjoint", choices=("prb", "volpathsimple"), default="prb",
                    help="adjoint integrator of the step: path replay of the primal integrator, or "
                         "volpathsimple's differential-ratio-tracking adjoint (the reference's choice for its "
                         "volume experiments, experiments.py:131)")
    ap.add_argument("--mode", choices=("step", "batch", "serve", "diff"), default="step",
                    help="step: the volume optimisation step (default line); batch: 64 concurrent jobs in one "
                         "graph-replayed iteration; serve: continuous batching under Poisson step requests; "
                         "diff: adjoints at 128^3 density + albedo, 4 x 512^2, with finite-difference checks")
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--profile-target", action="store_true",
                    help="run only: 3 workload steps, one two-state render, one adjoint and a 256 MiB copy "
                         "(the calibration kernel), for counter collection")
    return ap.parse_args()


def profile_target(args):
    import torch
    st = Stepper(args, "cv_ps")
    for _ in range(3):
        st.step()
    torch.cuda.synchronize()
    st.base.render_twostates(st.sc, 0, 5000, args.spp)
    st.params.update(st.opt)
    g = torch.full(st.sc.film_shape(0), 1e-6, device="cuda")
    st.adj.adjoint(st.sc, st.params, g, 0, 6000, args.spp)
    x = torch.ones(64 * 1024 * 1024, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    log("[bench] profile target done")


def build_problem(args, method="cv_ps"):
    import numpy as np
    import torch
    import recursive_control_variates_amd as rcv
    from recursive_control_variates_amd import scenes
    target = scenes.blob_grid(args.grid, 1)
    init = np.full_like(target, 0.1)
    sc = rcv.load_dict(scenes.volume_scene_dict(init, twostate=(method == "cv_ps"), film_res=args.res,
                                                spp=args.spp))
    itype = "twostateprbvolpath" if method == "cv_ps" else "prbvolpath"
    integ = rcv.load_dict({"type": itype, "max_depth": args.max_depth, "rr_depth": args.max_depth})
    return sc, integ, torch.from_numpy(target).cuda(), torch.from_numpy(init).cuda()


def reference_image(args, sc, integ, target):
    from recursive_control_variates_amd.optimize import render_reference
    return render_reference({"ref_passes": 4, "spp_ref": 128}, sc, integ, target={
        "object.interior_medium.sigma_t.data": target})


class Stepper:
    """One optimisation step of the workload, as the reference driver runs it
    (optimize.py:104-129) with the meta integrator."""

    def __init__(self, args, method="cv_ps", img_ref=None):
        import torch
        from recursive_control_variates_amd import scenes
        from recursive_control_variates_amd.cv_integrator import MetaIntegrator
        from recursive_control_variates_amd.optim import Adam, l1_loss
        from recursive_control_variates_amd.scene import traverse
        self.args = args
        self.sc, self.base, self.target, init = build_problem(args, method)
        self.key = scenes.VOLUME_KEY
        self.img_ref = img_ref if img_ref is not None else reference_image(args, self.sc, self.base, self.target)
        import recursive_control_variates_amd as rcv
        self.adj = self.base if getattr(args, "adjoint", "prb") == "prb" else rcv.load_dict(
            {"type": "volpathsimple", "max_depth": args.max_depth, "rr_depth": args.max_depth})
        self.meta = MetaIntegrator({"method": method, "integrator": self.base, "adjoint_integrator": self.adj})
        self.params = traverse(self.sc)
        self.opt = Adam(lr=0.01)
        self.opt.set_clamp(self.key, 0.0, 1e9)
        for _ in range(2 if method == "cv_ps" else 1):      # utils.py:62-87
            self.opt[self.key] = init
            self.params.update(self.opt)
        self.loss_fn = l1_loss
        self.i = 0
        self.torch = torch

    def step(self):
        from recursive_control_variates_amd.rendering import render
        self.params.update(self.opt)
        img = render(self.sc, self.params, integrator=self.meta, seed=self.i, spp=self.args.spp,
                     spp_grad=self.args.spp)
        loss = self.loss_fn(img, self.img_ref)
        loss.backward()
        self.opt.step()
        self.i += 1

    def param_mse(self):
        d = self.opt[self.key].detach() - self.target
        return float((d * d).mean())


def time_steps(st, steps, warmup):
    import numpy as np
    import torch
    for _ in range(warmup):
        st.step()
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t0 = time.perf_counter()
    evs[0].record()
    for k in range(steps):
        st.step()
        evs[k + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = np.array([evs[k].elapsed_time(evs[k + 1]) for k in range(steps)])
    return wall, per


def _event_time(fn, reps):
    """Average duration (ms) of fn's launches, HIP events on the launch stream."""
    import torch
    s = torch.cuda.current_stream()
    ts = []
    for r in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn(r)
        b.record(s)
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    return sum(ts) / len(ts)


def _traffic(name):
    tpath = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                return json.load(f).get("bytes_per_launch")
        except Exception:
            return None
    return None


def kernel_roofline(args, st):
    """Rooflines of the step's two render kernels, each timed alone with HIP
    events on the launch stream; event counts come from device counters in a
    separate launch so the counting atomics stay out of the timing. The
    adjoint is the dominant kernel of a step."""
    import torch
    sc, integ, adj = st.sc, st.base, st.adj
    n_pix = args.res * args.res * 4
    # two-state render: 2 states x 8 trilinear corners x 4 B per tentative
    # collision and per ratio-tracking step; 2 images x 12 B per pixel
    integ.render_twostates(sc, 0, 999, args.spp)
    torch.cuda.synchronize()
    r_ms = _event_time(lambda r: integ.render_twostates(sc, 0, 1000 + r, args.spp), args.kernel_reps)
    sc.set_counting(True)
    integ.render_twostates(sc, 0, 1000, args.spp)
    rc = sc.read_counters()
    r_bytes = 64 * (rc["collisions"] + rc["nee_steps"]) + 24 * n_pix
    # adjoint: 1 state x 8 corners x 4 B per lookup in the primal, replay and
    # NEE-gradient passes; 4 B per gradient scatter-add; 12 B per pixel read
    gimg = torch.full(sc.film_shape(0), 1.0 / (n_pix * 3), device="cuda")
    st.params.update(st.opt)
    sc.set_counting(False)
    adj.adjoint(sc, st.params, gimg, 0, 77, args.spp)
    torch.cuda.synchronize()
    a_ms = _event_time(lambda r: adj.adjoint(sc, st.params, gimg, 0, 2000 + r, args.spp), args.kernel_reps)
    sc.set_counting(True)
    adj.adjoint(sc, st.params, gimg, 0, 2000, args.spp)
    ac = sc.read_counters()
    sc.set_counting(False)
    a_bytes = 32 * (ac["collisions"] + ac["nee_steps"]) + 4 * ac["grad_atomics"] + 12 * n_pix
    out = {}
    akern = ("adjoint_persistent_kernel<VOLPATH> (path replay)" if adj is integ
             else "adjoint_sample_persistent_kernel<VOLPATH_SIMPLE> (differential ratio tracking)")
    for name, ms, by, cnt, kern in (("volume_adjoint", a_ms, a_bytes, ac, akern),
                                    ("volume_twostate", r_ms, r_bytes, rc,
                                     "render_persistent_kernel<2, VOLPATH> (two-state delta tracking)")):
        ach = by / (ms * 1e-3) / 1e9
        out[name] = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": _traffic(name), "kernel": kern,
                     "kernel_ms": round(ms, 4), "alg_bytes_per_launch": by, "counters": cnt}
    return out


def supergrid_compare(args, st, factor=8):
    """The step's two kernels at the current parameter state with a majorant
    supergrid of factor ``factor`` (ceil(grid / factor)^3 cells, DDA free
    flights) against the single global majorant: event-timed durations and
    tentative-collision counts from the device counters. Same samples, so the
    difference is the tracker's work."""
    import torch
    sc, integ, adj = st.sc, st.base, st.adj
    n_pix = args.res * args.res * 4
    gimg = torch.full(sc.film_shape(0), 1.0 / (n_pix * 3), device="cuda")
    st.params.update(st.opt)
    out = {}
    for f in (0, factor):
        sc.set_majorant_factor(0, f)
        integ.render_twostates(sc, 0, 999, args.spp)
        torch.cuda.synchronize()
        r_ms = _event_time(lambda r: integ.render_twostates(sc, 0, 3000 + r, args.spp), args.kernel_reps)
        a_ms = _event_time(lambda r: adj.adjoint(sc, st.params, gimg, 0, 4000 + r, args.spp), args.kernel_reps)
        sc.set_counting(True)
        integ.render_twostates(sc, 0, 3000, args.spp)
        rc = sc.read_counters()
        adj.adjoint(sc, st.params, gimg, 0, 4000, args.spp)
        ac = sc.read_counters()
        sc.set_counting(False)
        out[f"factor_{f}"] = {"twostate_ms": round(r_ms, 3), "adjoint_ms": round(a_ms, 3),
                              "twostate_collisions": rc["collisions"], "twostate_null_events": rc["null_events"],
                              "adjoint_collisions": ac["collisions"], "adjoint_nee_steps": ac["nee_steps"]}
    sc.set_majorant_factor(0, 0)
    return out


MSE_LEVELS = (1e-3, 2e-2, 1.5e-2, 1.25e-2)


def time_to_mse(args, img_ref):
    """Wall-clock until the unscaled-grid MSE reaches each level, for cv_ps and
    the plain loop (MSE checked every 5 iterations; the check synchronises)."""
    import torch
    levels = sorted(set(MSE_LEVELS) | {args.mse_threshold}, reverse=True)
    out = {}
    for method in ("cv_ps", "baseline"):
        st = Stepper(args, method, img_ref=img_ref)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reached = {}
        mse = st.param_mse()
        best = mse
        for i in range(args.mse_steps):
            st.step()
            if (i + 1) % 5 == 0 or i + 1 == args.mse_steps:
                mse = st.param_mse()
                best = min(best, mse)
                for lv in levels:
                    if lv not in reached and mse <= lv:
                        reached[lv] = (round(time.perf_counter() - t0, 3), i + 1)
        total = time.perf_counter() - t0
        out[method] = {"seconds": reached.get(args.mse_threshold, (None,))[0],
                       "by_level": {f"{lv:g}": {"seconds": reached[lv][0], "steps": reached[lv][1]}
                                    if lv in reached else None for lv in levels},
                       "final_mse": mse, "best_mse": best, "run_seconds": round(total, 3),
                       "ms_per_iteration": round(total * 1e3 / max(1, args.mse_steps), 3),
                       "iterations": args.mse_steps}
        log(f"[bench] time-to-mse {method}: {out[method]}")
        del st
        torch.cuda.empty_cache()
    ratios = {}
    for lv in levels:
        a = out["cv_ps"]["by_level"][f"{lv:g}"]
        b = out["baseline"]["by_level"][f"{lv:g}"]
        ratios[f"{lv:g}"] = round(b["seconds"] / a["seconds"], 3) if (a and b and a["seconds"]) else None
    out["speedup_cv_vs_plain"] = ratios
    return out


def cpu_baseline(args):
    """The float32 CPU restatement (oracle) timed on host cores on a bounded
    slice of the workload's pixels: two-state primal paths per second."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import render_ref as R
    from recursive_control_variates_amd import scene_desc, scenes
    target = scenes.blob_grid(args.grid, 1)
    init = np.full_like(target, 0.1)
    d = scene_desc.describe(scenes.volume_scene_dict(init, film_res=args.res, spp=args.spp))
    threads = min(16, os.cpu_count() or 1)
    n_pix_total = d.sensors[0].film_width * d.sensors[0].film_height
    # middle band of the film (rays through the medium), growing until the budget is used;
    # once the band is the whole film, further whole-film passes (new seeds) are added
    n = 256
    spent = 0.0
    while True:
        b = n_pix_total // 2 - n // 2
        t0 = time.perf_counter()
        R.render(d, 1, args.max_depth, args.max_depth, 7, args.spp, params_old=[target], params_new=[init],
                 precision=32, n_threads=threads, pixel_range=(b, b + n))
        dt = time.perf_counter() - t0
        spent += dt
        if dt > args.cpu_seconds / 3 or spent > args.cpu_seconds or n >= n_pix_total:
            break
        n = min(n_pix_total, n * 4)
    passes = 1
    while n >= n_pix_total and dt < args.cpu_seconds / 3 and spent < args.cpu_seconds:
        t0 = time.perf_counter()
        R.render(d, 1, args.max_depth, args.max_depth, 7 + passes, args.spp, params_old=[target],
                 params_new=[init], precision=32, n_threads=threads, pixel_range=(0, n))
        step = time.perf_counter() - t0
        dt += step
        spent += step
        passes += 1
    msps = passes * n * args.spp / dt / 1e6
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(msps, 4), "unit": "Msamples/s (two-state primal paths)", "cores": threads,
            "kind": "port", "sample": f"{passes} pass(es) of {n} pixels x {args.spp} spp of the 4x256^2 film "
                                      f"({'whole film' if n >= n_pix_total else 'central band'}), "
                                      f"two-state primal, float32, {dt:.2f}s on {cpu}"}


def bench_batch(args):
    """Config-4 style: ``--jobs`` volume optimisations (per-job seeds 100.., image
    side U{64..512}, grid U{32,64,128}, spp U{4..64}, max_depth U{8..32}), cv_ps,
    advanced together by one hipGraph-replayed batched iteration, after an
    autotuner sweep of block size x queue grab; compared with the plain per-job
    GPU loop (baseline method, one job at a time through the single-job path)."""
    import numpy as np
    import torch
    from recursive_control_variates_amd.batch import BatchOptimizer, autotune
    from recursive_control_variates_amd.cv_integrator import MetaIntegrator
    from recursive_control_variates_amd.optim import Adam, l1_loss
    from recursive_control_variates_amd.rendering import render
    from recursive_control_variates_amd.scene import traverse
    from recursive_control_variates_amd.serve import make_job
    import recursive_control_variates_amd as rcv
    seeds = list(range(100, 100 + args.jobs))
    t0 = time.perf_counter()
    made = [make_job(s, "volume", "cv_ps") for s in seeds]
    log(f"[bench] {len(made)} jobs set up in {time.perf_counter() - t0:.1f}s")
    paths = sum((m["side"] ** 2) * m["spp"] * 2 for _, m in made)
    small = lambda: [make_job(s, "volume", "cv_ps")[0] for s in seeds[:8]]
    best, table = autotune(small, blocks=(64, 128, 256, 512), grabs=(64, 128, 256), steps=2)
    log(f"[bench] autotune best {best}: {table}")
    opt = BatchOptimizer([sp for sp, _ in made], use_graph=True, block_size=best[0], queue_grab=best[1])
    opt.step(2)
    opt.synchronize()
    t = time.perf_counter()
    K = max(1, args.steps // 4)
    opt.step(K)
    opt.synchronize()
    batched_ms = (time.perf_counter() - t) * 1e3 / K
    log(f"[bench] batched iteration of {len(seeds)} jobs: {batched_ms:.1f} ms")
    opt.close()
    # plain per-job loop: each job one baseline iteration through the single-job path
    per_job = []
    for s in seeds:
        spec, m = make_job(s, "volume", "baseline")
        base = rcv.load_dict(dict(spec.integrator))
        meta = MetaIntegrator({"method": "baseline", "integrator": base})
        p = traverse(spec.scene)
        ad = Adam(lr=spec.lr)
        ad.set_clamp(spec.key, *spec.clamp)
        ad[spec.key] = torch.as_tensor(spec.init).cuda()
        p.update(ad)

        def one(i):
            p.update(ad)
            img = render(spec.scene, p, integrator=meta, seed=i, spp=spec.spp, spp_grad=spec.spp)
            l1_loss(img, spec.img_ref).backward()
            ad.step()
        one(0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        one(1)
        one(2)
        torch.cuda.synchronize()
        per_job.append((time.perf_counter() - t) * 1e3 / 2)
    seq_ms = float(sum(per_job))
    return {"metric": "batched cv_ps iteration of heterogeneous volume jobs vs plain per-job GPU loop",
            "jobs": len(seeds), "batched_ms_per_iteration": round(batched_ms, 3),
            "plain_per_job_loop_ms_per_iteration": round(seq_ms, 3),
            "speedup": round(seq_ms / batched_ms, 3), "msamples_per_s": round(paths / batched_ms / 1e3, 3),
            "autotune": {"best": list(best), "ms": {f"{b}x{g}": round(v, 3) for (b, g), v in table.items()}},
            "hip_graph": True, "dtype": "f32", "data": "synthetic (blob targets, seeds 100..)"}


def bench_serve(args):
    """Config-5 style: Poisson step requests at 200/s over a 70% surface /
    30% volume job mix (sizes as config 4, 100-500 iterations per job), HBM
    admission budget 200 GB, continuous batching."""
    from recursive_control_variates_amd.serve import run_simulation
    offered = run_simulation(duration=args.serve_seconds, rate=200.0, surface_fraction=0.7, target_jobs=16,
                             iterations=(100, 500), budget_bytes=200e9, seed=0, log=log)
    capacity = offered["served_steps"] / offered["duration_s"]
    log(f"[bench] offered 200 steps/s: served {capacity:.1f} steps/s")
    # below saturation the latency is the service time, not the queue length
    rate = max(1.0, 0.7 * capacity)
    below = run_simulation(duration=args.serve_seconds, rate=rate, surface_fraction=0.7, target_jobs=16,
                           iterations=(100, 500), budget_bytes=200e9, seed=1, log=log)
    return {"metric": "served optimisation steps: p50/p99 step latency and Msamples/s",
            "offered_200_per_s": offered, "capacity_steps_per_s": round(capacity, 2),
            "at_70pct_capacity": below}


def bench_diff(args):
    """Config-3 style: 128^3 density + 128^3 x 3 albedo grid, 4 views of
    512^2, 32 spp, max_depth 16. Times the two-state render and both adjoints
    (path replay; volpathsimple's DRT adjoint) for d/d(density) and
    d/d(albedo) at once, and checks each adjoint against central finite
    differences of two-state renders (shared paths) along random directions
    supported on 32 seed-chosen voxels of each grid."""
    import numpy as np
    import torch
    import recursive_control_variates_amd as rcv
    from recursive_control_variates_amd import scenes
    res, grid, spp = 512, 128, 32
    dens = scenes.blob_grid(grid, 1)
    alb = scenes.albedo_grid(grid, 1)
    sc = rcv.load_dict(scenes.volume_scene_dict(dens, film_res=res, spp=spp, albedo_grid_data=alb))
    p = rcv.traverse(sc)
    dk, ak = scenes.VOLUME_KEY, scenes.ALBEDO_KEY
    ts = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 16, "rr_depth": 16})
    prb = rcv.load_dict({"type": "prbvolpath", "max_depth": 16, "rr_depth": 16})
    simple = rcv.load_dict({"type": "volpathsimple", "max_depth": 16, "rr_depth": 16})
    n_pix = res * res * 4
    rng = np.random.default_rng(7)
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (res, 4 * res, 3)).astype(np.float32) / (n_pix * 3)).cuda()
    xd = torch.from_numpy(dens).cuda().requires_grad_(True)
    xa = torch.from_numpy(alb).cuda().requires_grad_(True)
    p[dk] = xd
    p[ak] = xa
    p.update()
    p.update()
    out = {"metric": "config 3: two-state render and adjoints at 128^3 density + albedo grid, 4 x 512^2, 32 spp",
           "samples_per_launch": n_pix * spp}
    ts.render_twostates(sc, 0, 1, spp)
    torch.cuda.synchronize()
    out["twostate_ms"] = round(_event_time(lambda r: ts.render_twostates(sc, 0, 10 + r, spp), 3), 3)
    for name, integ in (("prb", prb), ("volpathsimple", simple)):
        integ.adjoint(sc, p, gimg, 0, 2, spp)
        torch.cuda.synchronize()
        out[f"adjoint_{name}_ms"] = round(_event_time(lambda r: integ.adjoint(sc, p, gimg, 0, 20 + r, spp), 3), 3)
    # the same launches with the k = 8 majorant supergrid (16^3 cells of local majorants)
    sc.set_majorant_factor(0, 8)
    ts.render_twostates(sc, 0, 1, spp)
    prb.adjoint(sc, p, gimg, 0, 2, spp)
    torch.cuda.synchronize()
    out["supergrid_8"] = {
        "twostate_ms": round(_event_time(lambda r: ts.render_twostates(sc, 0, 10 + r, spp), 3), 3),
        "adjoint_prb_ms": round(_event_time(lambda r: prb.adjoint(sc, p, gimg, 0, 20 + r, spp), 3), 3)}
    sc.set_majorant_factor(0, 0)
    log(f"[bench] diff timings {out}")
    # finite differences along directions on 32 seed-chosen voxels. The
    # two-state render of (theta + eps delta, theta - eps delta) and the
    # path-replay adjoint at theta trace the same paths for the same seed
    # (common random numbers), so their per-seed values agree to float
    # precision; volpathsimple's DRT estimator draws extra random numbers and
    # is compared in expectation ([mean, standard error] over the seeds).
    fd = {}
    seeds = 8
    for key, base in ((dk, dens), (ak, alb)):
        flat = base.reshape(-1, base.shape[-1])
        strong = np.flatnonzero(dens.reshape(-1) > 0.3 * dens.max())
        vox = np.random.default_rng(11).choice(strong, 32, replace=False)
        rows = []
        for d_i in range(2):
            delta = np.zeros_like(flat)
            delta[vox] = np.random.default_rng(100 + d_i).uniform(-1, 1, (32, flat.shape[1]))
            delta = delta.reshape(base.shape).astype(np.float32)
            dt = torch.from_numpy(delta).cuda().double()
            eps = 1e-3 * float(np.abs(base).max())
            f_vals, a_vals = [], {"prb": [], "volpathsimple": []}
            for s in range(seeds):
                p[key] = torch.from_numpy(base + eps * delta).cuda()
                p.update()
                p[key] = torch.from_numpy(base - eps * delta).cuda()
                p.update()
                H, F = ts.render_twostates(sc, 0, 500 + s, spp)
                f_vals.append(float(((H.double() - F.double()) * gimg.double()).sum()) / (2 * eps))
                # back to old = new = theta: update() shifts only assigned keys
                # (parameters_changed of the modified objects), so assign twice
                x = torch.from_numpy(base).cuda().requires_grad_(True)
                for _ in range(2):
                    p[key] = x
                    p.update()
                for name, integ, sd in (("prb", prb, 500 + s), ("volpathsimple", simple, 700 + s)):
                    g = integ.adjoint(sc, p, gimg, 0, sd, spp)[key]
                    a_vals[name].append(float((g.double() * dt).sum()))
            f_vals = np.array(f_vals)
            se = lambda v: float(np.std(v, ddof=1) / np.sqrt(len(v)))
            prb_v = np.array(a_vals["prb"])
            rows.append({"fd": [float(f_vals.mean()), se(f_vals)],
                         **{n: [float(np.mean(v)), se(v)] for n, v in a_vals.items()},
                         "prb_same_seed_max_rel_diff": float(np.max(np.abs(f_vals - prb_v)) /
                                                             max(abs(float(f_vals.mean())), 1e-30))})
        fd[key] = rows
    out["finite_differences"] = fd
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    log(f"[bench] device {torch.cuda.get_device_name(0)}")
    if args.profile_target:
        profile_target(args)
        return
    if args.mode == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.mode == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if args.mode == "diff":
        print(json.dumps(bench_diff(args)), flush=True)
        return
    st = Stepper(args, "cv_ps")
    log("[bench] reference image rendered")
    wall, per = time_steps(st, args.steps, args.warmup)
    ms = wall * 1e3 / args.steps
    n_pix = args.res * args.res * 4
    samples = n_pix * args.spp * 2
    msps = samples / (ms * 1e-3) / 1e6
    log(f"[bench] {args.steps} steps: {ms:.3f} ms/step, {msps:.1f} Msamples/s")
    roof = kernel_roofline(args, st)
    log(f"[bench] roofline {roof}")
    sg = supergrid_compare(args, st)
    log(f"[bench] supergrid {sg}")
    img_ref = st.img_ref
    del st
    ttm = time_to_mse(args, img_ref) if args.mse_steps > 0 else None
    cpu = cpu_baseline(args) if args.cpu_seconds > 0 else None
    log(f"[bench] cpu baseline {cpu}")
    out = {
        "metric": "Msamples/s (two-state primal + adjoint paths) of a cv_ps volume optimisation step; "
                  "also time-to-param-MSE 1e-3 (CV vs plain) and p50/p99 step ms",
        "value": round(msps, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (procedural blob-grid target, seed 1)",
        "config": {"workload": f"heterogeneous volume {args.grid}^3, 4 views {args.res}x{args.res}, "
                               f"spp {args.spp}/{args.spp}, max_depth {args.max_depth}, cv_ps, Adam lr 0.01, L1, "
                               f"adjoint {args.adjoint}",
                   "global_batch": n_pix * args.spp, "parallelism": "single GPU"},
        "p50_step_ms": round(float(np.percentile(per, 50)), 4),
        "p99_step_ms": round(float(np.percentile(per, 99)), 4),
        "time_to_param_mse": ttm,
        "roofline": {k: v for k, v in roof["volume_adjoint"].items()
                     if k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_detail": roof,
        "majorant_supergrid": sg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
