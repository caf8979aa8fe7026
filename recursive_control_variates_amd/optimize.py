"""Optimisation driver with the loop order of the reference's run_opt
(optimize.py:10-154) and its helpers (utils.py:62-87,98-127,129-192).

Per iteration i: parameters_changed (scene update from the optimizer), render
with the meta integrator (seed i, spp, spp_grad), loss against the reference
image, checkpoint every ``recomp_freq`` steps (plain render at ``spp_inf``,
seed i+1), backward through the path-replay adjoint, Adam step with fused
clamping, resolution doubling at the ``upsample`` fractions (with the CV
reset and the supergrid factor update), learning-rate halving at the
``schedule`` fractions.

With ``output_dir`` the run reads and writes the reference's files: the
reference image ``img_ref.exr`` and start image ``img_start.exr`` next to
``output_dir`` (reused when present, utils.py:98-111, optimize.py:91-97),
checkpoint images with ``save`` (``img_inf``, ``img``, ``weights``,
``img_H``, ``img_F`` as EXR + PNG, utils.py:129-156), ``img_final.exr`` and
the final parameters (``.vol`` grids, ``.exr`` textures, ``final_state.json``
for other values, utils.py:45-60).

``result["time_to_threshold_s"]`` is the wall clock from the start of the
loop to the end of the step (render, loss, checkpoint, backward, Adam) whose
updated parameters first reach ``mse_threshold``; it includes checkpoint
renders and the per-step MSE synchronisation.

``result["runtime"]`` holds, for steps i > 16 with ``benchmark``, the GPU-
synchronised [primal + loss, backward + step] milliseconds
(optimize.py:47,105-132); ``step_ms`` the same two phases summed for every
step. Extras for measurement: if ``target`` ({key: tensor}) is given, the
parameter MSE against it is recorded every step (the check synchronises).
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict

import numpy as np
import torch

from . import fileio
from .cv_integrator import MetaIntegrator
from .optim import Adam, l1_loss, l2_loss
from .rendering import render
from .scene import resample_nearest, traverse


def _loss_func(name):
    if name == "L1":
        return l1_loss
    if name == "L2":
        return l2_loss
    raise ValueError(f"Unknown loss function {name!r} (L1, L2; VGG is not part of this build)")


def render_reference(cfg, scene, integrator, params=None, target=None):
    """Average of ``ref_passes`` renders at ``spp_ref`` with seeds 17843 + j
    (utils.py:98-111), of the target parameter values if given. With
    ``output_dir``, ``img_ref.exr`` beside it is read if present, else
    written."""
    save_path = None
    if cfg.get("output_dir"):
        save_path = os.path.join(os.path.dirname(os.path.abspath(cfg["output_dir"])), "img_ref.exr")
        if os.path.isfile(save_path):
            return torch.from_numpy(fileio.read_exr_rgb(save_path)).to(scene.device)
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
    if save_path:
        fileio.write_exr(save_path, img)
    return img


def save_final_state(cfg, opt, output_dir):
    """utils.py:45-60: grids as .vol, textures as .exr, anything else in
    final_state.json."""
    final_state = {}
    for key in cfg["variables"].keys():
        x = opt[key].detach()
        name = key.replace(".", "_") + "_final"
        if x.dim() == 4:
            fileio.write_vol(os.path.join(output_dir, name + ".vol"), x)
        elif x.dim() == 3:
            fileio.write_exr(os.path.join(output_dir, name + ".exr"), x)
        else:
            final_state[key] = x.cpu().numpy().tolist()
    if final_state:
        with open(os.path.join(output_dir, "final_state.json"), "w") as f:
            json.dump(final_state, f)


def _checkpoint(cfg, i, integrator, base, scene, loss_func, img, img_ref, result):
    """utils.py:129-156: every recomp_freq steps, the loss of a fresh plain
    render at spp_inf (seed i + 1) and the variance of the step's image
    against it; with ``save``, the images."""
    recomp_freq = int(cfg.get("recomp_freq", 10))
    if recomp_freq <= 0 or i % recomp_freq != 0:
        return
    out_dir = cfg.get("output_dir")
    save = bool(cfg.get("save", False)) and out_dir
    with torch.no_grad():
        img_inf = base.render(scene, 0, i + 1, int(cfg.get("spp_inf", 64)))
        result["loss"][i // recomp_freq] = float(loss_func(img_ref, img_inf))
        result["var"][i // recomp_freq] = float(((img_inf - img.detach()) ** 2).mean())
        if save:
            k = i // recomp_freq
            fileio.save_img(img_inf, out_dir, "img_inf", k, save_png=True)
            if integrator.img is not None:
                fileio.save_img(integrator.img, out_dir, "img", k, save_png=True)
            method = cfg.get("method", "baseline")
            stats = getattr(integrator, "stats", None)
            if method != "baseline" and stats is not None and int(stats.n.item()) > integrator.warmup:
                fileio.save_img(integrator.w_s, out_dir, "weights", k)
                fileio.save_img(integrator.H, out_dir, "img_H", k)
                fileio.save_img(integrator.F, out_dir, "img_F", k)


def _single_medium(scene) -> int:
    """Index of the scene's only medium (utils.py get_single_medium)."""
    n = len(scene.desc.media)
    if n != 1:
        raise ValueError(f"expected exactly one medium in the scene, found {n}")
    return 0


def adjust_majorant_res_factor(scene, density_res):
    """utils.py:27-43: the largest supergrid factor <= 8 that leaves at least
    4 cells along the grid's shortest side; 0 (one global majorant) when
    none does."""
    res_factor = 8
    min_side = min(int(x) for x in density_res[:3])
    while res_factor > 1 and (min_side // res_factor) < 4:
        res_factor -= 1
    if res_factor <= 1:
        res_factor = 0
    medium = _single_medium(scene)
    if scene.majorant_resolution_factor(medium) != res_factor:
        scene.set_majorant_factor(medium, res_factor)


def upsample(scene, cfg, opt, integrator):
    """utils.py:177-192: every optimised tensor doubles its resolution by
    voxel repetition (Adam's state for it restarts, as for any shape change),
    the supergrid factor follows a density's new resolution, and the CV state
    is reset."""
    use_sg = bool(cfg.get("use_majorant_supergrid", False))
    for key in cfg["variables"].keys():
        x = opt[key]
        if not isinstance(x, torch.Tensor) or x.dim() not in (3, 4):
            raise ValueError(f"Upsampling expects a 3 or 4D tensor, got {getattr(x, 'shape', type(x))}")
        new_res = (*[2 * s for s in x.shape[:-1]], x.shape[-1])
        opt[key] = resample_nearest(x.detach(), new_res)
        if ".sigma_t." in key and use_sg:
            adjust_majorant_res_factor(scene, new_res)
    integrator.reset()


def initialize_parameters(cfg, opt, scene_params):
    """utils.py:62-87: run twice for cv_ps so that old = new = init; a grid
    initial density sets the supergrid factor when use_majorant_supergrid."""
    for _ in range(2 if cfg.get("method", "baseline") == "cv_ps" else 1):
        for key, p in cfg["variables"].items():
            init = p["init"]
            if not isinstance(init, torch.Tensor):
                init = torch.as_tensor(np.asarray(init, np.float32))
            if "sigma_t" in key and init.dim() == 4 and cfg.get("use_majorant_supergrid", False):
                adjust_majorant_res_factor(scene_params.scene, tuple(init.shape))
            opt[key] = init.to(scene_params.scene.device)
            if "lr_factor" in p:
                opt.set_learning_rate({key: cfg.get("lr", 1e-2) * p["lr_factor"]})
            if "clamp" in p:
                opt.set_clamp(key, *p["clamp"])
        scene_params.update(opt)


def _check_supported(cfg):
    """Keys of the reference's config that change the optimisation and that
    this driver does not implement raise instead of being ignored: the
    large-steps preconditioner (per-variable ``largesteps``, utils.py:167-170,
    largesteps.py) and the display-resolution render ``final_res``
    (optimize.py:84-87, utils.py:113-127). ``denoise`` reaches
    MetaIntegrator, which raises for it."""
    if "final_res" in cfg:
        raise ValueError("run_opt: 'final_res' (display-resolution renders) is not supported by this build")
    for key, var in cfg.get("variables", {}).items():
        if isinstance(var, dict) and var.get("largesteps"):
            raise ValueError(f"run_opt: variable {key!r} asks for 'largesteps' preconditioning, "
                             "which this build does not implement")
    if cfg.get("loss", "L1") == "VGG":
        raise ValueError("run_opt: the VGG loss is not part of this build (L1, L2)")


def run_opt(cfg: Dict) -> Dict:
    scene = cfg["scene"]
    loss_func = _loss_func(cfg.get("loss", "L1"))
    method = cfg.get("method", "baseline")
    n_steps = int(cfg.get("n_steps", 0))
    lr = float(cfg.get("lr", 1e-2))
    spp = int(cfg.get("spp", 1))
    spp_grad = int(cfg.get("spp_grad", spp))
    recomp_freq = int(cfg.get("recomp_freq", 10))
    benchmark = bool(cfg.get("benchmark", False))
    target = cfg.get("target")
    mse_threshold = cfg.get("mse_threshold")
    if "integrator" not in cfg:
        raise ValueError("No integrator specified")
    _check_supported(cfg)
    base = cfg["integrator"]
    integrator = MetaIntegrator({
        "method": method, "denoise": cfg.get("denoise", False), "integrator": base,
        "adjoint_integrator": cfg.get("adjoint_integrator", base),
        "beta1": cfg.get("beta1", 0.9), "beta2": cfg.get("beta2", 0.999),
        "force_baseline": cfg.get("force_baseline", False), "pre_update": cfg.get("pre_update", False),
        "warmup": cfg.get("warmup", 1)})
    params = traverse(scene)
    img_ref = cfg.get("img_ref")
    if img_ref is None:
        img_ref = render_reference(cfg, scene, base, params, target)

    upsample_steps = []
    for w in cfg.get("upsample", []):
        if not 0 < w < 1:
            raise ValueError("upsample fractions must lie in (0, 1)")
        upsample_steps.append(int(n_steps * w))
    schedule_steps = []
    for w in cfg.get("schedule", []):
        if not 0 < w < 1:
            raise ValueError("schedule fractions must lie in (0, 1)")
        schedule_steps.append(int(n_steps * w))
    result = {"loss": np.zeros(1 + n_steps // max(1, recomp_freq), np.float32),
              "var": np.zeros(1 + n_steps // max(1, recomp_freq), np.float32),
              "runtime": np.zeros((n_steps, 2), np.float32),
              "step_ms": np.zeros(n_steps, np.float64),
              "param_mse": np.full(n_steps, np.nan),
              "wall_s": np.zeros(n_steps, np.float64),
              "time_to_threshold_s": None, "steps_to_threshold": None}

    out_dir = cfg.get("output_dir")
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
    opt = Adam(lr=lr)
    initialize_parameters(cfg, opt, params)
    if out_dir:
        start_path = os.path.join(os.path.dirname(os.path.abspath(out_dir)), "img_start.exr")
        if not os.path.isfile(start_path):
            with torch.no_grad():
                fileio.write_exr(start_path, base.render(scene, 0, 2048, int(cfg.get("spp_ref", spp))))
    dev = scene.device
    torch.cuda.synchronize(dev)
    t_start = time.perf_counter()
    for i in range(n_steps):
        t0 = time.perf_counter()
        params.update(opt)
        img = render(scene, params, integrator=integrator, seed=i, spp=spp, spp_grad=spp_grad)
        loss = loss_func(img, img_ref)
        if benchmark:
            torch.cuda.synchronize(dev)
        t1 = time.perf_counter()
        _checkpoint(cfg, i, integrator, base, scene, loss_func, img, img_ref, result)
        t2 = time.perf_counter()
        loss.backward()
        opt.step()
        if benchmark or target is not None:
            torch.cuda.synchronize(dev)
        t3 = time.perf_counter()
        if benchmark and i > 16:
            result["runtime"][i] = ((t1 - t0) * 1e3, (t3 - t2) * 1e3)
        if i in upsample_steps:
            upsample(scene, cfg, opt, integrator)
        if i in schedule_steps:
            lr *= 0.5
            opt.set_learning_rate(lr)
        result["step_ms"][i] = (t1 - t0 + t3 - t2) * 1e3
        result["wall_s"][i] = t3 - t_start
        if target is not None:
            mse = 0.0
            cnt = 0
            with torch.no_grad():
                for k, v in target.items():
                    est = opt[k].detach()
                    if est.shape != v.shape:       # coarse phase of an upsampled run
                        est = resample_nearest(est, tuple(v.shape))
                    d = est - v.to(est.device)
                    mse += float((d * d).sum())
                    cnt += d.numel()
            result["param_mse"][i] = mse / cnt
            if mse_threshold is not None and result["time_to_threshold_s"] is None and mse / cnt <= mse_threshold:
                result["time_to_threshold_s"] = t3 - t_start
                result["steps_to_threshold"] = i + 1
    params.update(opt)
    if out_dir:
        fileio.write_exr(os.path.join(out_dir, "img_final.exr"),
                         render_reference(dict(cfg, output_dir=None), scene, base))
        save_final_state(cfg, opt, out_dir)
    result["final"] = {k: opt[k].detach() for k in opt.keys()}
    result["integrator"] = integrator
    return result
