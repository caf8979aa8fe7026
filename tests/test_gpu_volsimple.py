"""volpathsimple on the GPU (plugins/volpathsimple.py): primal and DRT adjoint
against the CPU oracle with the same random numbers, the albedo-grid
gradient against finite differences, the scene restrictions, and its use as
the adjoint integrator of a cv_ps optimisation."""
import numpy as np
import pytest
import torch

import render_ref as R
from conftest import announce
from helpers import furnace_box_scene, volume_box_scene
import recursive_control_variates_amd as rcv
from recursive_control_variates_amd import _native as N
from recursive_control_variates_amd.transform import T

pytestmark = pytest.mark.gpu
SIMPLE = 2
KEY = "box.interior_medium.sigma_t.data"
AKEY = "box.interior_medium.albedo.data"


def _close_fraction(a, b, rtol=1e-3, atol=1e-4):
    return np.mean(np.abs(a - b) <= atol + rtol * np.abs(b))


def _grid(seed, n=5):
    rng = np.random.default_rng(seed)
    return (0.2 + rng.uniform(0, 0.8, (n, n, n, 1))).astype(np.float32)


def test_primal_matches_oracle_sample_by_sample(cuda):
    announce("volpathsimple primal vs oracle fp32")
    g = _grid(1)
    sc = rcv.load_dict(volume_box_scene(g, scale=6.0, w=16, h=12, spp=8))
    for flags, props in ((0, {}), (N.SIMPLE_NO_NEE, {"use_nee": False})):
        integ = rcv.load_dict(dict({"type": "volpathsimple", "max_depth": 12, "rr_depth": 4}, **props))
        img = integ.render(sc, 0, 55, 8).cpu().numpy()
        ref = R.render(sc.desc, SIMPLE, 12, 4, 55, 8, precision=32, flags=flags)[1]
        assert _close_fraction(img, ref) > 0.97
        assert abs(img.mean() - ref.mean()) < 2e-3 * abs(ref.mean()) + 1e-5


@pytest.mark.parametrize("props,flags", [({}, 0), ({"use_drt": False}, N.SIMPLE_NO_DRT),
                                         ({"use_drt_subsampling": False}, N.SIMPLE_NO_DRT_SUBSAMPLING)])
def test_adjoint_matches_oracle_same_samples(cuda, props, flags):
    announce(f"volpathsimple adjoint vs oracle fp32 (flags {flags})")
    rng = np.random.default_rng(2)
    g = _grid(2)
    sc = rcv.load_dict(volume_box_scene(g, albedo=0.8, scale=5.0, w=8, h=8, spp=16))
    gimg = rng.uniform(0.5, 1.5, (8, 8, 3))
    integ = rcv.load_dict(dict({"type": "volpathsimple", "max_depth": 8, "rr_depth": 8}, **props))
    p = rcv.traverse(sc)
    x = p[KEY].detach().clone().requires_grad_(True)
    p[KEY] = x
    p.update()
    p.update()
    g_gpu = integ.adjoint(sc, p, torch.from_numpy(gimg.astype(np.float32)).cuda(), 0, 31, 32)[KEY].cpu().numpy()
    g_cpu = R.adjoint(sc.desc, 8, 8, 31, 32, gimg, precision=32, integrator_type=SIMPLE, flags=flags)
    scale = np.abs(g_cpu).max()
    close = np.abs(g_gpu - g_cpu) <= 2e-3 * scale + 2e-3 * np.abs(g_cpu)
    assert close.mean() > 0.9, close.mean()
    assert abs(g_gpu.sum() - g_cpu.sum()) < 0.05 * np.abs(g_cpu).sum()


def test_albedo_and_density_gradients_match_finite_differences(cuda):
    # DRT + collision + transmittance + NEE terms, for both parameters, against
    # central differences of two-state volume renders (shared paths)
    announce("volpathsimple adjoint vs FD")
    rng = np.random.default_rng(3)
    n = 4
    dens = (0.4 + rng.uniform(0, 0.6, (n, n, n, 1))).astype(np.float32)
    alb = rng.uniform(0.4, 0.9, (n, n, n, 3)).astype(np.float32)
    alb_dict = {"type": "gridvolume", "data": alb, "to_world": T.scale(1.0)}
    sc = rcv.load_dict(volume_box_scene(dens, albedo=alb_dict, scale=4.0, w=8, h=8, spp=16))
    gimg = rng.uniform(0.5, 1.5, (8, 8, 3)).astype(np.float32)
    ts = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": -1, "rr_depth": 4})
    simple = rcv.load_dict({"type": "volpathsimple", "max_depth": -1, "rr_depth": 4})
    p = rcv.traverse(sc)
    for key, base in ((KEY, dens), (AKEY, alb)):
        delta = rng.uniform(-1, 1, base.shape).astype(np.float32)
        eps = 2e-2
        fds, adj = [], []
        for s in range(12):
            p[key] = torch.from_numpy(base + eps * delta).cuda()
            p.update()
            p[key] = torch.from_numpy(base - eps * delta).cuda()
            p.update()
            H, F = ts.render_twostates(sc, 0, 900 + s, 2048)
            fds.append(float(np.sum(gimg * (H - F).cpu().numpy()) / (2 * eps)))
            # old = new = theta for every key: update() shifts only the keys
            # assigned since the last update, so each restore assigns twice
            x = torch.from_numpy(base).cuda().requires_grad_(True)
            for _ in range(2):
                p[key] = x
                p.update()
            gr = simple.adjoint(sc, p, torch.from_numpy(gimg).cuda(), 0, 1900 + s, 2048)[key].cpu().numpy()
            adj.append(float(np.sum(gr * delta)))
            for _ in range(2):
                p[key] = torch.from_numpy(base).cuda()
                p.update()
        fds, adj = np.array(fds), np.array(adj)
        se = np.sqrt(fds.var(ddof=1) / len(fds) + adj.var(ddof=1) / len(adj))
        assert abs(fds.mean() - adj.mean()) < 4 * se + 0.03 * abs(fds.mean()), (key, fds.mean(), adj.mean(), se)


def test_scene_restrictions(cuda):
    announce("volpathsimple scene checks")
    integ = rcv.load_dict({"type": "volpathsimple", "max_depth": 4})
    surf = rcv.load_dict(furnace_box_scene(albedo=0.5, w=4, h=4, spp=4))
    with pytest.raises(N.SceneCheckError):
        integ.render(surf, 0, 1, 4)
    d = volume_box_scene(_grid(4))
    d["light"] = {"type": "rectangle", "to_world": T.translate([0.5, 2.5, 0.5]).rotate([1, 0, 0], 90).scale(0.2),
                  "bsdf": {"type": "null"}, "emitter": {"type": "area", "radiance": {"type": "rgb", "value": 1.0}}}
    with pytest.raises(N.SceneCheckError):
        integ.render(rcv.load_dict(d), 0, 1, 4)


def test_adjoint_integrator_of_a_cv_optimisation(cuda):
    # the reference's volume setup: twostateprbvolpath primal, volpathsimple
    # adjoint (experiments.py:130-131)
    announce("cv_ps with volpathsimple adjoint")
    from recursive_control_variates_amd.optimize import run_opt
    rng = np.random.default_rng(5)
    target = rng.uniform(0.2, 1.0, (6, 6, 6, 1)).astype(np.float32)
    init = np.full_like(target, 0.5)
    sc = rcv.load_dict(volume_box_scene(init, scale=5.0, w=12, h=10, spp=8))
    cfg = {"scene": sc, "method": "cv_ps", "n_steps": 8, "lr": 0.02, "spp": 8, "spp_ref": 32, "recomp_freq": 4,
           "spp_inf": 8,
           "integrator": rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 16, "rr_depth": 16}),
           "adjoint_integrator": rcv.load_dict({"type": "volpathsimple", "max_depth": 16, "rr_depth": 16}),
           "variables": {KEY: {"init": init, "clamp": (0.0, 250.0)}},
           "target": {KEY: torch.from_numpy(target).cuda()}}
    res = run_opt(cfg)
    assert np.isfinite(res["param_mse"]).all()
    assert res["param_mse"][-1] < res["param_mse"][0]
    assert bool(torch.isfinite(res["final"][KEY]).all())
