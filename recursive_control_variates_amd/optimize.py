"""Optimisation driver with the loop order of the reference's run_opt
(optimize.py:10-154) and its helpers (utils.py:62-87,98-127,129-192).

Per iteration i: parameters_changed (scene update from the optimizer), render
with the meta integrator (seed i, spp, spp_grad), loss against the reference
image, checkpoint every ``recomp_freq`` steps (plain render at ``spp_inf``,
seed i+1), backward through the path-replay adjoint, Adam step with fused
clamping, learning-rate halving at the ``schedule`` fractions.

Extras for measurement: if ``target`` ({key: tensor}) is given, the
parameter MSE against it is recorded every step; ``benchmark`` records the
wall-clock time of every step (GPU synchronised), excluding nothing;
statistics skip the first 17 steps as the reference does (optimize.py:105).
"""
from __future__ import annotations

import time
from typing import Dict

import numpy as np
import torch

from .cv_integrator import MetaIntegrator
from .optim import Adam, l1_loss, l2_loss
from .rendering import render
from .scene import traverse


def _loss_func(name):
    if name == "L1":
        return l1_loss
    if name == "L2":
        return l2_loss
    raise ValueError(f"Unknown loss function {name!r} (L1, L2; VGG is not part of this build)")


def render_reference(cfg, scene, integrator, params=None, target=None):
    """Average of ``ref_passes`` renders at ``spp_ref`` with seeds 17843 + j
    (utils.py:98-111), of the target parameter values if given."""
    ref_passes = int(cfg.get("ref_passes", 1))
    spp_ref = int(cfg.get("spp_ref", cfg.get("spp", 1)))
    saved = {}
    if target:
        params = params or traverse(scene)
        for k, v in target.items():
            saved[k] = params[k].detach().clone()
            params[k] = v
        params.update()
    img = None
    with torch.no_grad():
        for j in range(ref_passes):
            r = integrator.render(scene, 0, 17843 + j, spp_ref) / ref_passes
            img = r if img is None else img + r
    if target:
        for k, v in saved.items():
            params[k] = v
        params.update()
        params.update()
    return img


def initialize_parameters(cfg, opt, scene_params):
    """utils.py:62-87: run twice for cv_ps so that old = new = init."""
    for _ in range(2 if cfg.get("method", "baseline") == "cv_ps" else 1):
        for key, p in cfg["variables"].items():
            init = p["init"]
            if not isinstance(init, torch.Tensor):
                init = torch.as_tensor(np.asarray(init, np.float32))
            opt[key] = init.to(scene_params.scene.device)
            if "lr_factor" in p:
                opt.set_learning_rate({key: cfg.get("lr", 1e-2) * p["lr_factor"]})
            if "clamp" in p:
                opt.set_clamp(key, *p["clamp"])
        scene_params.update(opt)


def run_opt(cfg: Dict) -> Dict:
    scene = cfg["scene"]
    loss_func = _loss_func(cfg.get("loss", "L1"))
    method = cfg.get("method", "baseline")
    n_steps = int(cfg.get("n_steps", 0))
    lr = float(cfg.get("lr", 1e-2))
    spp = int(cfg.get("spp", 1))
    spp_grad = int(cfg.get("spp_grad", spp))
    recomp_freq = int(cfg.get("recomp_freq", 10))
    spp_inf = int(cfg.get("spp_inf", 64))
    benchmark = bool(cfg.get("benchmark", False))
    target = cfg.get("target")
    mse_threshold = cfg.get("mse_threshold")
    if "integrator" not in cfg:
        raise ValueError("No integrator specified")
    if cfg.get("upsample"):
        raise NotImplementedError("resolution upsampling during optimisation is not part of this build")
    base = cfg["integrator"]
    integrator = MetaIntegrator({
        "method": method, "integrator": base,
        "adjoint_integrator": cfg.get("adjoint_integrator", base),
        "beta1": cfg.get("beta1", 0.9), "beta2": cfg.get("beta2", 0.999),
        "force_baseline": cfg.get("force_baseline", False), "pre_update": cfg.get("pre_update", False),
        "warmup": cfg.get("warmup", 1)})
    params = traverse(scene)
    img_ref = cfg.get("img_ref")
    if img_ref is None:
        img_ref = render_reference(cfg, scene, base, params, target)

    schedule_steps = [int(n_steps * w) for w in cfg.get("schedule", [])]
    result = {"loss": np.zeros(1 + n_steps // max(1, recomp_freq), np.float32),
              "var": np.zeros(1 + n_steps // max(1, recomp_freq), np.float32),
              "step_ms": np.zeros(n_steps, np.float64),
              "param_mse": np.full(n_steps, np.nan),
              "wall_s": np.zeros(n_steps, np.float64),
              "time_to_threshold_s": None, "steps_to_threshold": None}

    opt = Adam(lr=lr)
    initialize_parameters(cfg, opt, params)
    torch.cuda.synchronize(scene.device)
    t_start = time.perf_counter()
    for i in range(n_steps):
        t0 = time.perf_counter()
        params.update(opt)
        img = render(scene, params, integrator=integrator, seed=i, spp=spp, spp_grad=spp_grad)
        loss = loss_func(img, img_ref)
        if recomp_freq > 0 and i % recomp_freq == 0 and not benchmark:
            with torch.no_grad():
                img_inf = base.render(scene, 0, i + 1, spp_inf)
                result["loss"][i // recomp_freq] = float(loss_func(img_ref, img_inf))
                result["var"][i // recomp_freq] = float(((img_inf - img.detach()) ** 2).mean())
        loss.backward()
        opt.step()
        if i in schedule_steps:
            lr *= 0.5
            opt.set_learning_rate(lr)
        if benchmark or target is not None:
            torch.cuda.synchronize(scene.device)
        t1 = time.perf_counter()
        result["step_ms"][i] = (t1 - t0) * 1e3
        result["wall_s"][i] = t1 - t_start
        if target is not None:
            mse = 0.0
            cnt = 0
            with torch.no_grad():
                for k, v in target.items():
                    d = opt[k].detach() - v.to(opt[k].device)
                    mse += float((d * d).sum())
                    cnt += d.numel()
            result["param_mse"][i] = mse / cnt
            if mse_threshold is not None and result["time_to_threshold_s"] is None and mse / cnt <= mse_threshold:
                result["time_to_threshold_s"] = t1 - t_start
                result["steps_to_threshold"] = i + 1
    params.update(opt)
    result["final"] = {k: opt[k].detach() for k in opt.keys()}
    result["integrator"] = integrator
    return result
