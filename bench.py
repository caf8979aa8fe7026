#!/usr/bin/env python3
"""Benchmark of the recursive-control-variate inverse-rendering step on one
MI355X.

Workload (the heterogeneous-volume configuration): 64^3 float32 density grid
in the unit cube (target: 8 Gaussian blobs from seed 1, sigma_t scale 20),
albedo 0.8, isotropic phase, constant environment 1, 4 views of 256x256 on one
batch film, 32 spp primal and 32 spp adjoint, max_depth 16, grid initialised at
0.1, Adam lr 0.01, L1 loss, method cv_ps (two-state render + CV combine).

A step = scene update (two-state pointer rotation) + two-state render + fused
CV combine + loss/gradient + path-replay adjoint + Adam. The workload is the
whole 200-iteration optimisation: ``--steps K`` times its last K steps
(bracketed by device synchronisations), the ``--warmup W`` steps before them
run untimed, as does any earlier part of the run (none with the defaults
K = 183, W = 17: the first 17 steps are left out as optimize.py:105 leaves
them out of its runtime figures). ``value`` is Msamples/s = (W*H*spp primal
paths + W*H*spp_grad adjoint paths) per timed step; p50/p99 come from HIP
events around each timed step.

The same line carries: time-to-param-MSE for cv_ps and the plain loop (both
run for the whole optimisation, MSE recorded after every step, clock = HIP
events from the start of step 1), the rooflines of the step's two render
kernels (HIP events around every launch during the run; algorithmic bytes
from counting launches replayed with the inputs of sampled timed steps), a
CPU baseline (the float32 CPU restatement of the same two launches on the
host's cores, bounded sample) and, as extra sections, the surface-texture,
large-grid, batched and served workloads.
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
    ap.add_argument("--steps", type=int, default=200 - SKIP_STEPS,
                    help="timed steps: the last STEPS iterations of the optimisation")
    ap.add_argument("--warmup", type=int, default=SKIP_STEPS,
                    help="untimed steps right before the timed ones")
    ap.add_argument("--iterations", type=int, default=200,
                    help="optimisation length (at least warmup + steps)")
    ap.add_argument("--mse-threshold", type=float, default=1e-3)
    ap.add_argument("--reachable-threshold", type=float, default=1.25e-2,
                    help="an MSE level both methods reach within the run (the 64^3 workload's MSE has a floor)")
    ap.add_argument("--count-samples", type=int, default=12,
                    help="timed steps whose launches are replayed in counting mode for the rooflines")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--extras", default="surface,large_grid,batch,serve",
                    help="extra workload sections of the default line (comma list; empty: none)")
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--spp", type=int, default=32)
    ap.add_argument("--max-depth", type=int, default=16)
    ap.add_argument("--majorant-factor", default="0",
                    help="majorant supergrid of the volume workload: 0 = one global majorant, k > 0 = cells of k^3 "
                         "voxels, 'auto' = the reference's rule for use_majorant_supergrid (utils.py:27-43)")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--adjoint", choices=("prb", "volpathsimple"), default="prb",
                    help="adjoint integrator of the step: path replay of the primal integrator, or "
                         "volpathsimple's differential-ratio-tracking adjoint (the reference's choice for its "
                         "volume experiments, experiments.py:131)")
    ap.add_argument("--mode", choices=("step", "batch", "serve", "diff"), default="step",
                    help="step: the volume optimisation (default line); batch: 64 concurrent jobs in one "
                         "graph-replayed iteration; serve: continuous batching under Poisson step requests; "
                         "diff: adjoints at 128^3 density + albedo, 4 x 512^2, with finite-difference checks")
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--autotune", choices=("full", "short"), default="full")
    ap.add_argument("--profile-target", action="store_true",
                    help="run only the cv_ps optimisation (ITERATIONS steps) and a 256 MiB copy (the "
                         "calibration kernel), for counter collection")
    return ap.parse_args()


def profile_target(args):
    import torch
    st = Stepper(args, "cv_ps")
    torch.cuda.synchronize()
    log("[bench] profile target: reference image done")
    for i in range(args.iterations):
        st.step()
        if (i + 1) % 50 == 0:
            torch.cuda.synchronize()
            log(f"[bench] profile target: step {i + 1}")
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
    mf = getattr(args, "majorant_factor", "0")
    if mf == "auto":
        from recursive_control_variates_amd.optimize import adjust_majorant_res_factor
        adjust_majorant_res_factor(sc, init.shape)
    elif int(mf) > 0:
        sc.set_majorant_factor(0, int(mf))
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


SKIP_STEPS = 17   # optimize.py:105 leaves a run's first 17 steps out of its runtime figures


class _AdjointCapture:
    """Keeps the inputs of the adjoint launches of chosen steps (image
    gradient, seed, spp) so that their event counts can be measured after the
    run in separate counting launches at the same parameter state."""

    def __init__(self, integ, want):
        self.integ = integ
        self.want = set(want)
        self.step = -1
        self.saved = {}
        self._orig = integ.adjoint
        integ.adjoint = self._adjoint

    def _adjoint(self, scene, params, grad_in, sensor=0, seed=0, spp=0):
        if self.step in self.want:
            self.saved[self.step] = dict(grad=grad_in.detach().clone(), seed=seed, spp=spp)
        return self._orig(scene, params, grad_in, sensor, seed, spp)

    def close(self):
        del self.integ.adjoint


def run_optimisation(args, method, img_ref, n_iter, window=None, sample=()):
    """Runs ``n_iter`` iterations of the workload's optimisation with a HIP
    event at every step boundary, HIP events around every adjoint and
    two-state launch (on their launch stream), and the parameter MSE after
    every step written to a device buffer that is read once at the end.
    ``window`` = (first, count) brackets steps [first, first + count)
    (0-based) by device synchronisations and times them on the host clock.
    The launch inputs of the steps in ``sample`` and the parameter states
    around them are kept for counting launches afterwards."""
    import numpy as np
    import torch
    st = Stepper(args, method, img_ref=img_ref)
    st.base.time_launches(True)
    if st.adj is not st.base:
        st.adj.time_launches(True)
    adj_ev, ts_ev = st.adj.adjoint_events, st.base.twostate_events
    want = set(sample)
    cap = _AdjointCapture(st.adj, want)
    snaps = {}
    mse = torch.zeros(n_iter, dtype=torch.float32, device="cuda")
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(n_iter + 1)]
    wall = None
    torch.cuda.synchronize()
    evs[0].record()
    for i in range(n_iter):
        if window and i == window[0]:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        if i in want or (i + 1) in want:
            snaps[i] = st.opt[st.key].detach().clone()     # the NEW state of step i
        cap.step = i
        st.step()
        evs[i + 1].record()
        d = st.opt[st.key].detach() - st.target
        mse[i] = (d * d).mean()
        if window and i == window[0] + window[1] - 1:
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
    torch.cuda.synchronize()
    cap.close()
    run = dict(per=np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(n_iter)]),
               cum=np.array([evs[0].elapsed_time(evs[i + 1]) for i in range(n_iter)]),
               mse=mse.double().cpu().numpy(), wall=wall, adj_ms=np.array(adj_ev.elapsed_ms()),
               ts_ms=np.array(ts_ev.elapsed_ms()) if method == "cv_ps" else None, snaps=snaps,
               captured=cap.saved, method=method)
    st.base.time_launches(False)
    st.adj.time_launches(False)
    return st, run


def count_bytes(args, st, run):
    """Algorithmic bytes of the adjoint and two-state launches of the
    sampled steps, from device event counters in counting launches with the
    same inputs at the same parameter state:
    two-state: 2 states x 8 trilinear corners x 4 B per tentative collision
    and per ratio-tracking step, 2 images x 12 B per pixel;
    adjoint: 1 state x 8 corners x 4 B per lookup in the primal and replay
    passes (main path and NEE ratio tracking), 4 B per gradient scatter-add,
    12 B per pixel of the image gradient read."""
    sc, p, key = st.sc, st.params, st.key
    n_pix = sc.film_shape(0)[0] * sc.film_shape(0)[1]
    rows = {}
    for i in sorted(run["captured"]):
        cap = run["captured"][i]
        new = run["snaps"][i]
        old = run["snaps"].get(i - 1, new)
        row = {}
        if run["method"] == "cv_ps":
            p[st.key] = old
            p.update()
            p[st.key] = new
            p.update()
            sc.set_counting(True)
            st.base.render_twostates(sc, 0, i, args.spp)
            rc = sc.read_counters()
            sc.set_counting(False)
            row["twostate_bytes"] = 64 * (rc["collisions"] + rc["nee_steps"]) + 24 * n_pix
            row["twostate_counters"] = rc
        x = new.clone().requires_grad_(True)
        for _ in range(2):
            p[key] = x
            p.update()
        sc.set_counting(True)
        st.adj.adjoint(sc, p, cap["grad"], 0, cap["seed"], cap["spp"])
        ac = sc.read_counters()
        sc.set_counting(False)
        row["adjoint_bytes"] = 32 * (ac["collisions"] + ac["nee_steps"]) + 4 * ac["grad_atomics"] + 12 * n_pix
        row["adjoint_counters"] = ac
        rows[i] = row
    return rows


def _traffic(name):
    """HBM bytes per launch of a kernel from the committed PMC summary, with
    the state it was measured at."""
    tpath = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                d = json.load(f)
            return d.get("bytes_per_launch"), d.get("state")
        except Exception:
            return None, None
    return None, None


def rooflines(run, counts, window):
    """Rooflines of the step's two render kernels over the sampled steps of
    the timed window: algorithmic bytes of those launches over their
    event-timed durations in the run."""
    import numpy as np
    idx = sorted(counts)
    first, count = window
    out = {}
    kinds = [("volume_adjoint", "adj_ms", "adjoint_bytes",
              "adjoint_persistent_kernel<VOLPATH> + log_reduce_kernel (path replay and its scatter-log "
              "reduction, one rcv_render_backward call)")]
    if run["ts_ms"] is not None:
        kinds.append(("volume_twostate", "ts_ms", "twostate_bytes",
                      "render_persistent_kernel<2, VOLPATH> (two-state delta tracking)"))
    for name, mkey, bkey, kern in kinds:
        ms = np.array([run[mkey][i] for i in idx])
        by = np.array([counts[i][bkey] for i in idx], np.float64)
        ach = by.sum() / (ms.sum() * 1e-3) / 1e9
        traffic, tstate = _traffic(name)
        out[name] = {"bound": "hbm", "achieved": round(float(ach), 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(float(ach) / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_state": tstate,
                     "kernel": kern, "sampled_steps": [i + 1 for i in idx],
                     "kernel_ms_sampled_mean": round(float(ms.mean()), 4),
                     "kernel_ms_sampled": [round(float(x), 3) for x in ms],
                     "events_sampled": [{k: counts[i][bkey.replace("bytes", "counters")][k]
                                         for k in ("collisions", "nee_steps", "grad_atomics", "null_events",
                                                   "real_scatter")} for i in idx],
                     "kernel_ms_window_mean": round(float(run[mkey][first:first + count].mean()), 4),
                     "alg_bytes_per_launch_mean": float(by.mean())}
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
    x = st.opt[st.key].detach().clone().requires_grad_(True)
    for _ in range(2):
        st.params[st.key] = x
        st.params.update()
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
    # where the adjoint's time goes at this state: the same launch without
    # gradient targets (primal and replay passes, no scatter-adds) and the
    # plain render of the same samples
    x = st.opt[st.key].detach().clone()
    for _ in range(2):
        st.params[st.key] = x
        st.params.update()
    out["adjoint_without_scatter_ms"] = round(_event_time(
        lambda r: adj.adjoint(sc, st.params, gimg, 0, 4000 + r, args.spp), args.kernel_reps), 3)
    out["plain_render_same_samples_ms"] = round(_event_time(
        lambda r: adj.render(sc, 0, 4000 + r, args.spp), args.kernel_reps), 3)
    out["state"] = "the parameters after the last step of the cv_ps run"
    return out


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


def mse_levels(args):
    return sorted({args.mse_threshold, args.reachable_threshold, 2e-2, 1.5e-2}, reverse=True)


def time_to_levels(run, levels, mask_mse=None):
    """First step whose parameter MSE is at or below each level and the
    elapsed time from the start of the first step to the end of that step
    (HIP events on the launch stream)."""
    mse = run["mse"] if mask_mse is None else mask_mse
    out = {}
    for lv in levels:
        hit = [i for i in range(len(mse)) if mse[i] <= lv]
        out[f"{lv:g}"] = ({"seconds": round(float(run["cum"][hit[0]]) / 1e3, 3), "steps": hit[0] + 1}
                          if hit else None)
    return out


def _host_cores():
    """Cores this process may run on: its CPU affinity, capped by a cgroup
    CPU quota when one is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            n = min(n, max(1, int(q) // int(per)))
    except (OSError, ValueError):
        pass
    return n


def cpu_baseline(args, state):
    """The float32 CPU restatement (oracle) of the step's two launches — the
    two-state primal render and the path-replay adjoint of the density — on
    every core this process may use, at a parameter state and image gradient
    taken from the GPU run, over every s-th film row (s halves until about
    ``cpu_seconds`` are spent): (primal + adjoint paths) per second, the
    unit of ``value``."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import render_ref as R
    from recursive_control_variates_amd import scene_desc, scenes
    old, new, gimg, seed, seed_grad, step = state
    d = scene_desc.describe(scenes.volume_scene_dict(new, film_res=args.res, spp=args.spp))
    threads = _host_cores()
    W, H = d.sensors[0].film_width, d.sensors[0].film_height
    stride, spent = 64, 0.0
    while True:
        rows = list(range(stride // 2, H, stride))
        t0 = time.perf_counter()
        for r in rows:
            R.render(d, 1, args.max_depth, args.max_depth, seed, args.spp, params_old=[old], params_new=[new],
                     precision=32, n_threads=threads, pixel_range=(r * W, (r + 1) * W))
        tp = time.perf_counter() - t0
        g = np.zeros_like(gimg)
        g[rows] = gimg[rows]
        t0 = time.perf_counter()
        R.adjoint(d, args.max_depth, args.max_depth, seed_grad, args.spp, g, param=0, params=[new], precision=32,
                  n_threads=threads, integrator_type=1)
        ta = time.perf_counter() - t0
        spent += tp + ta
        if stride == 1 or tp + ta > args.cpu_seconds / 3 or spent > args.cpu_seconds:
            break
        stride //= 2
    n = len(rows) * W
    msps = 2 * n * args.spp / (tp + ta) / 1e6
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(msps, 4), "unit": "Msamples/s (two-state primal + adjoint paths)", "cores": threads,
            "kind": "port",
            "sample": f"{len(rows)} of {H} rows (every {stride}th) of the 4 x {args.res}^2 film x {args.spp} spp, "
                      f"two-state primal {tp:.2f}s + density adjoint {ta:.2f}s, float32, at the parameters and "
                      f"image gradient of step {step + 1} of the GPU cv_ps run, {threads} threads on {cpu}",
            "primal_msamples_per_s": round(n * args.spp / tp / 1e6, 4),
            "adjoint_msamples_per_s": round(n * args.spp / ta / 1e6, 4)}


def observable_texels(res=64):
    """Back-wall texels of the Cornell box whose centre the camera sees (the
    sphere hides the others): texel (row j, column i) lies at
    ((i + 0.5) / res * 2 - 1, (j + 0.5) / res * 2 - 1, -1); a texel is
    observable when the segment from the camera at (0, 0, 3.9) misses the
    sphere of radius 0.3 at (0.35, -0.7, 0.2). The wall lies inside the
    field of view."""
    import numpy as np
    c = (np.arange(res) + 0.5) / res * 2 - 1
    yy, xx = np.meshgrid(c, c, indexing="ij")
    p = np.stack([xx, yy, np.full_like(xx, -1.0)], -1)
    o = np.array([0.0, 0.0, 3.9])
    d = p - o
    L = np.linalg.norm(d, axis=-1, keepdims=True)
    d = d / L
    oc = o - np.array([0.35, -0.7, 0.2])
    b = (d * oc).sum(-1)
    disc = b * b - (oc * oc).sum() + 0.3 ** 2
    t = -b - np.sqrt(np.maximum(disc, 0))
    hit = (disc > 0) & (t > 0) & (t < L[..., 0])
    return ~hit


def bench_surface(args):
    """The surface workload: 128 x 128 textured Cornell box, 64 x 64 x 3
    back-wall texture (target U[0.1, 0.9] from seed 0, initial 0.5),
    max_depth 4, 16 spp primal and adjoint, Adam lr 0.02, L2 image loss, 50
    iterations of cv_ps and of the plain loop: time to texture MSE 1e-3 over
    all texels and over the texels the camera sees, per-step ms."""
    import numpy as np
    import torch
    import recursive_control_variates_amd as rcv
    from recursive_control_variates_amd import scenes
    from recursive_control_variates_amd.cv_integrator import MetaIntegrator
    from recursive_control_variates_amd.optim import Adam, l2_loss
    from recursive_control_variates_amd.optimize import render_reference
    from recursive_control_variates_amd.rendering import render
    target = scenes.texture_target(64, 0)
    init = np.full_like(target, 0.5)
    obs = torch.from_numpy(observable_texels(64)).cuda()
    tgt = torch.from_numpy(target).cuda()
    levels = (3e-2, 2e-2, 1.5e-2, 1e-2, 5e-3, 1e-3)
    out = {"workload": "Cornell box 128^2, 64^2 x 3 texture, max_depth 4, spp 16/16, Adam lr 0.02, L2, "
                       "50 iterations", "observable_texels": int(obs.sum()), "texels": 64 * 64}
    img_ref = None
    for method in ("cv_ps", "baseline"):
        sc = rcv.load_dict(scenes.cornell_scene_dict(init, twostate=(method == "cv_ps"), film_res=128, spp=16))
        base = rcv.load_dict({"type": "twostateprb" if method == "cv_ps" else "prb", "max_depth": 4})
        if img_ref is None:
            plain = rcv.load_dict({"type": "prb", "max_depth": 4})
            img_ref = render_reference({"ref_passes": 4, "spp_ref": 1024}, sc, plain,
                                       target={scenes.TEXTURE_KEY: tgt})
        meta = MetaIntegrator({"method": method, "integrator": base})
        p = rcv.traverse(sc)
        opt = Adam(lr=0.02)
        opt.set_clamp(scenes.TEXTURE_KEY, 0.0, 1.0)
        for _ in range(2 if method == "cv_ps" else 1):
            opt[scenes.TEXTURE_KEY] = torch.from_numpy(init).cuda()
            p.update(opt)
        n = 50
        mse = torch.zeros(n, device="cuda")
        mse_obs = torch.zeros(n, device="cuda")
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(n + 1)]
        torch.cuda.synchronize()
        evs[0].record()
        for i in range(n):
            p.update(opt)
            img = render(sc, p, integrator=meta, seed=i, spp=16, spp_grad=16)
            l2_loss(img, img_ref).backward()
            opt.step()
            evs[i + 1].record()
            d2 = (opt[scenes.TEXTURE_KEY].detach() - tgt) ** 2
            mse[i] = d2.mean()
            mse_obs[i] = d2[obs].mean()
        torch.cuda.synchronize()
        run = {"cum": np.array([evs[0].elapsed_time(evs[i + 1]) for i in range(n)]),
               "per": np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(n)]),
               "mse": mse.double().cpu().numpy()}
        mo = mse_obs.double().cpu().numpy()
        out[method] = {"seconds_to_1e-3": time_to_levels(run, [1e-3])["0.001"],
                       "seconds_to_1e-3_observable": time_to_levels(run, [1e-3], mo)["0.001"],
                       "by_level": time_to_levels(run, levels),
                       "by_level_observable": time_to_levels(run, levels, mo),
                       "final_mse": float(run["mse"][-1]), "final_mse_observable": float(mo[-1]),
                       "ms_per_step_after_skip": round(float(run["per"][SKIP_STEPS:].mean()), 4),
                       "p50_step_ms_after_skip": round(float(np.percentile(run["per"][SKIP_STEPS:], 50)), 4),
                       "p99_step_ms_after_skip": round(float(np.percentile(run["per"][SKIP_STEPS:], 99)), 4),
                       "run_seconds": round(float(run["cum"][-1]) / 1e3, 4)}
        log(f"[bench] surface {method}: {out[method]}")
    return out


def bench_large_grid(args):
    """Timings at 128^3 density + 128^3 x 3 albedo grid, 4 views of 512^2,
    32 spp, max_depth 16: the two-state render and both adjoints, with one
    global majorant and with the k = 8 supergrid."""
    import torch
    import recursive_control_variates_amd as rcv
    from recursive_control_variates_amd import scenes
    res, grid, spp = 512, 128, 32
    dens = scenes.blob_grid(grid, 1)
    alb = scenes.albedo_grid(grid, 1)
    sc = rcv.load_dict(scenes.volume_scene_dict(dens, film_res=res, spp=spp, albedo_grid_data=alb))
    p = rcv.traverse(sc)
    ts = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 16, "rr_depth": 16})
    prb = rcv.load_dict({"type": "prbvolpath", "max_depth": 16, "rr_depth": 16})
    simple = rcv.load_dict({"type": "volpathsimple", "max_depth": 16, "rr_depth": 16})
    n_pix = res * res * 4
    gimg = torch.full((res, 4 * res, 3), 1.0 / (n_pix * 3), device="cuda")
    p[scenes.VOLUME_KEY] = torch.from_numpy(dens).cuda().requires_grad_(True)
    p[scenes.ALBEDO_KEY] = torch.from_numpy(alb).cuda().requires_grad_(True)
    p.update()
    p.update()
    out = {"workload": "128^3 density + 128^3 x 3 albedo, 4 x 512^2, 32 spp, max_depth 16, d/d(density) and "
                       "d/d(albedo) per adjoint launch", "samples_per_launch": n_pix * spp}
    for f in (0, 8):
        sc.set_majorant_factor(0, f)
        ts.render_twostates(sc, 0, 1, spp)
        prb.adjoint(sc, p, gimg, 0, 2, spp)
        torch.cuda.synchronize()
        row = {"twostate_ms": round(_event_time(lambda r: ts.render_twostates(sc, 0, 10 + r, spp), 3), 3),
               "adjoint_prb_ms": round(_event_time(lambda r: prb.adjoint(sc, p, gimg, 0, 20 + r, spp), 3), 3)}
        if f == 0:
            simple.adjoint(sc, p, gimg, 0, 2, spp)
            row["adjoint_volpathsimple_ms"] = round(
                _event_time(lambda r: simple.adjoint(sc, p, gimg, 0, 20 + r, spp), 3), 3)
        out[f"majorant_factor_{f}"] = row
    sc.set_majorant_factor(0, 0)
    log(f"[bench] large grid {out}")
    return out


def bench_batch(args, short=False):
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
    short = short or args.autotune == "short"
    blocks, grabs = ((64, 256), (64, 128)) if short else ((64, 128, 256, 512), (64, 128, 256))
    best, table = autotune(small, blocks=blocks, grabs=grabs, steps=2)
    log(f"[bench] autotune best {best}: {table}")
    opt = BatchOptimizer([sp for sp, _ in made], use_graph=True, block_size=best[0], queue_grab=best[1])
    opt.step(2)
    opt.synchronize()
    t = time.perf_counter()
    K = 5 if short else 10
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


def bench_serve(args, seconds=None):
    """Config-5 style: Poisson step requests at 200/s over a 70% surface /
    30% volume job mix (sizes as config 4, 100-500 iterations per job), HBM
    admission budget 200 GB, continuous batching."""
    from recursive_control_variates_amd.serve import run_simulation
    seconds = seconds or args.serve_seconds
    offered = run_simulation(duration=seconds, rate=200.0, surface_fraction=0.7, target_jobs=16,
                             iterations=(100, 500), budget_bytes=200e9, seed=0, log=log)
    capacity = offered["served_steps"] / offered["duration_s"]
    log(f"[bench] offered 200 steps/s: served {capacity:.1f} steps/s")
    # below saturation the latency is the service time, not the queue length
    rate = max(1.0, 0.7 * capacity)
    below = run_simulation(duration=seconds, rate=rate, surface_fraction=0.7, target_jobs=16,
                           iterations=(100, 500), budget_bytes=200e9, seed=1, log=log)
    return {"metric": "served optimisation steps: p50/p99 step latency and Msamples/s", "seconds": seconds,
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
    K, W = args.steps, args.warmup
    n_iter = max(args.iterations, K + W)
    first = n_iter - K                       # 0-based index of the first timed step
    skip = min(SKIP_STEPS, first)
    pick = np.unique(np.linspace(first, n_iter - 1, max(1, args.count_samples)).round().astype(int)).tolist()
    sc_, integ_, target_, _ = build_problem(args, "cv_ps")
    img_ref = reference_image(args, sc_, integ_, target_)
    del sc_, integ_, target_
    log("[bench] reference image rendered")
    runs, states = {}, {}
    for method in ("cv_ps", "baseline"):
        st, run = run_optimisation(args, method, img_ref, n_iter, window=(first, K) if method == "cv_ps" else None,
                                   sample=pick if method == "cv_ps" else ())
        log(f"[bench] {method}: {n_iter} iterations, {run['cum'][-1] / 1e3:.2f} s; final MSE {run['mse'][-1]:.5f}")
        runs[method] = run
        if method == "cv_ps":
            counts = count_bytes(args, st, run)
            mid = pick[len(pick) // 2]
            states["cpu"] = (run["snaps"].get(mid - 1, run["snaps"][mid]).cpu().numpy(),
                             run["snaps"][mid].cpu().numpy(), run["captured"][mid]["grad"].cpu().numpy(),
                             mid, run["captured"][mid]["seed"], mid)
            sg = supergrid_compare(args, st)
            log(f"[bench] supergrid {sg}")
        del st
        torch.cuda.empty_cache()
    cv = runs["cv_ps"]
    n_pix = args.res * args.res * 4
    samples = n_pix * args.spp * 2
    ms = cv["wall"] * 1e3 / K
    msps = samples / (ms * 1e-3) / 1e6
    per = cv["per"][first:]
    log(f"[bench] timed steps {first + 1}..{n_iter}: {ms:.3f} ms/step, {msps:.1f} Msamples/s")
    roof = rooflines(cv, counts, (first, K))
    log(f"[bench] roofline {roof}")
    levels = mse_levels(args)
    ttm = {"threshold": args.mse_threshold, "reachable_threshold": args.reachable_threshold,
           "clock": "HIP events from the start of step 1 to the end of the step that reaches the level"}
    for method, run in runs.items():
        lv = time_to_levels(run, levels)
        hit = lv[f"{args.mse_threshold:g}"]
        rh = lv[f"{args.reachable_threshold:g}"]
        ttm[method] = {"seconds": hit["seconds"] if hit else None,
                       "seconds_to_reachable": rh["seconds"] if rh else None,
                       "by_level": lv, "final_mse": float(run["mse"][-1]), "best_mse": float(run["mse"].min()),
                       "iterations": n_iter, "run_seconds": round(float(run["cum"][-1]) / 1e3, 3),
                       "ms_per_step_after_skip": round(float(run["per"][skip:].mean()), 4),
                       "p50_step_ms_after_skip": round(float(np.percentile(run["per"][skip:], 50)), 4),
                       "p99_step_ms_after_skip": round(float(np.percentile(run["per"][skip:], 99)), 4)}
    ratios = {}
    for lv in levels:
        a = ttm["cv_ps"]["by_level"][f"{lv:g}"]
        b = ttm["baseline"]["by_level"][f"{lv:g}"]
        ratios[f"{lv:g}"] = round(b["seconds"] / a["seconds"], 3) if (a and b and a["seconds"]) else None
    ttm["speedup_cv_vs_plain"] = ratios
    cpu = cpu_baseline(args, states["cpu"]) if args.cpu_seconds > 0 else None
    log(f"[bench] cpu baseline {cpu}")
    extras = {}
    for name in [e for e in args.extras.split(",") if e]:
        t0 = time.perf_counter()
        if name == "surface":
            extras["surface_texture"] = bench_surface(args)
        elif name == "large_grid":
            extras["large_grid"] = bench_large_grid(args)
        elif name == "batch":
            extras["batched_64_jobs"] = bench_batch(args, short=True)
        elif name == "serve":
            extras["served"] = bench_serve(args, seconds=10.0)
        else:
            raise SystemExit(f"unknown extra section {name!r}")
        torch.cuda.empty_cache()
        log(f"[bench] extra {name}: {time.perf_counter() - t0:.1f} s")
    out = {
        "metric": "Msamples/s (two-state primal + adjoint paths) of a cv_ps volume optimisation step; "
                  "also time-to-param-MSE 1e-3 (CV vs plain) and p50/p99 step ms",
        "value": round(msps, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (procedural blob-grid target, seed 1)",
        "config": {"workload": f"heterogeneous volume {args.grid}^3, 4 views {args.res}x{args.res}, "
                               f"spp {args.spp}/{args.spp}, max_depth {args.max_depth}, cv_ps, Adam lr 0.01, L1, "
                               f"adjoint {args.adjoint}, {n_iter} iterations; timed: steps {first + 1}..{n_iter}",
                   "global_batch": n_pix * args.spp, "parallelism": "single GPU"},
        "p50_step_ms": round(float(np.percentile(per, 50)), 4),
        "p99_step_ms": round(float(np.percentile(per, 99)), 4),
        "time_to_param_mse": ttm,
        "roofline": {k: v for k, v in roof["volume_adjoint"].items()
                     if k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_detail": roof,
        "majorant_supergrid": sg,
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
