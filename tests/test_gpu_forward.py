"""Forward mode (render_forward, plugins/cv_integrator.py:79-85 and
plugins/metaintegrator.py:132-138 delegating to the upstream PRB /
PRBVolpath forward mode): the forward replay gathers the tangents where the
path-replay adjoint scatters, so on the same samples <g, J t> equals
<J^T g, t> up to float rounding (the adjoint is pinned against the oracle and
finite differences); per pixel it equals the central difference of the
two-state render with shared paths; and the plugins delegate to it."""
import numpy as np
import pytest
import torch

from conftest import announce
from helpers import furnace_box_scene, homogeneous_box_scene, volume_box_scene
import recursive_control_variates_amd as rcv
from recursive_control_variates_amd.transform import T

pytestmark = pytest.mark.gpu


def _bind(sc, key, value):
    p = rcv.traverse(sc)
    x = torch.as_tensor(value).cuda().requires_grad_(True)
    for _ in range(2):
        p[key] = x
        p.update()
    return p, x


def _volume(seed=21, n=6):
    rng = np.random.default_rng(seed)
    dens = (0.3 + rng.uniform(0, 0.7, (n, n, n, 1))).astype(np.float32)
    alb = rng.uniform(0.5, 0.95, (n, n, n, 3)).astype(np.float32)
    alb_dict = {"type": "gridvolume", "data": alb, "to_world": T.scale(1.0)}
    return rng, dens, alb, volume_box_scene(dens, albedo=alb_dict, scale=12.0, w=16, h=12, spp=8)


def _transpose_check(sc, integ, key, base, gimg, tangent, seed, spp, rtol=2e-4):
    p, _ = _bind(sc, key, base)
    dimg = integ.render_forward(sc, p, 0, seed, spp, tangents={key: tangent})
    g = integ.adjoint(sc, p, gimg, 0, seed, spp)[key]
    lhs = float((gimg.double() * dimg.double()).sum())
    rhs = float((g.double() * tangent.double()).sum())
    print(f"[forward] {key}: <g, J t> {lhs:.6e}  <J^T g, t> {rhs:.6e}", flush=True)
    assert abs(lhs) > 0
    assert abs(lhs - rhs) <= rtol * abs(lhs) + 1e-9, (lhs, rhs)
    return dimg


@pytest.mark.parametrize("which", ["sigma_t", "albedo", "sigma_t_absorber"])
def test_forward_is_transpose_of_adjoint_volume(cuda, which):
    # The density's derivative sums null-collision terms (-scale / sigma_n)
    # and real-collision terms (+scale / sigma_t) of opposite signs per voxel,
    # which the adjoint adds with float32 atomics: the two sides then agree to
    # the rounding of those cancelling sums (2e-3). Without scattering (albedo
    # 0: null terms only, no NEE) and for the albedo (one sign per channel)
    # they agree to 2e-4.
    announce(f"forward vs adjoint transpose (volume {which})")
    rng, dens, alb, d = _volume()
    if which == "sigma_t_absorber":
        alb = np.zeros_like(alb)
        d["box"]["interior"]["incoming"]["albedo"]["data"] = alb
    sc = rcv.load_dict(d)
    base = alb if which == "albedo" else dens
    key = "box.interior_medium." + ("albedo" if which == "albedo" else "sigma_t") + ".data"
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (12, 16, 3)).astype(np.float32)).cuda()
    tangent = torch.from_numpy(rng.uniform(-1, 1, base.shape).astype(np.float32)).cuda()
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": 12, "rr_depth": 12})
    for s in range(2):
        _transpose_check(sc, integ, key, base, gimg, tangent, 30 + s, 32, rtol=2e-3 if which == "sigma_t" else 2e-4)


@pytest.mark.parametrize("key", ["box.interior_medium.sigma_t.value.value", "box.interior_medium.albedo.value.value"])
def test_forward_is_transpose_of_adjoint_homogeneous(cuda, key):
    announce(f"forward vs adjoint transpose (homogeneous {key})")
    sc = rcv.load_dict(homogeneous_box_scene(sigma_t=0.9, albedo=(0.3, 0.6, 0.9), scale=2.0, w=16, h=12))
    base = rcv.traverse(sc)[key].detach().cpu().numpy()
    rng = np.random.default_rng(3)
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (12, 16, 3)).astype(np.float32)).cuda()
    tangent = torch.from_numpy(rng.uniform(0.5, 1.0, base.shape).astype(np.float32)).cuda()
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8})
    _transpose_check(sc, integ, key, base, gimg, tangent, 7, 16, rtol=1e-3)


def test_forward_is_transpose_of_adjoint_texture(cuda):
    announce("forward vs adjoint transpose (texture)")
    rng = np.random.default_rng(14)
    tex = rng.uniform(0.2, 0.8, (4, 4, 3)).astype(np.float32)
    sc = rcv.load_dict(furnace_box_scene(albedo=0.5, w=16, h=12, textured=True, tex=tex))
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (12, 16, 3)).astype(np.float32)).cuda()
    tangent = torch.from_numpy(rng.uniform(-1, 1, tex.shape).astype(np.float32)).cuda()
    integ = rcv.load_dict({"type": "prb", "max_depth": 4, "rr_depth": 100})
    _transpose_check(sc, integ, "back.bsdf.reflectance.data", tex, gimg, tangent, 5, 16)


def test_forward_equals_per_pixel_shared_path_difference(cuda):
    # the two-state render of (theta - eps t, theta + eps t) shares its paths
    # with the forward replay at theta for the same seed (the averaged
    # decisions equal theta's; the grid maximum, hence the majorant, is kept)
    announce("forward vs per-pixel shared-path central difference")
    rng, dens, alb, d = _volume(22)
    sc = rcv.load_dict(d)
    key = "box.interior_medium.sigma_t.data"
    t = rng.uniform(-1, 1, dens.shape).astype(np.float32)
    t[dens > 0.9 * dens.max()] = 0.0
    eps = 1e-3 * float(dens.max())
    ts = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 12, "rr_depth": 12})
    plain = rcv.load_dict({"type": "prbvolpath", "max_depth": 12, "rr_depth": 12})
    p = rcv.traverse(sc)
    for v in (dens - eps * t, dens + eps * t):
        p[key] = torch.from_numpy(v).cuda()
        p.update()
    H, F = ts.render_twostates(sc, 0, 51, 64)
    fd = ((F.double() - H.double()) / (2 * eps)).cpu().numpy()
    p, _ = _bind(sc, key, dens)
    dimg = plain.render_forward(sc, p, 0, 51, 64, tangents={key: torch.from_numpy(t).cuda()}).double().cpu().numpy()
    scale = np.abs(fd).max()
    close = np.abs(dimg - fd) <= 0.01 * np.abs(fd) + 1e-3 * scale
    print(f"[forward] per-pixel agreement {close.mean():.4f}", flush=True)
    assert close.mean() > 0.97
    assert abs(dimg.sum() - fd.sum()) <= 0.01 * np.abs(fd).sum()


def test_forward_plugin_delegation_and_grad_tangent(cuda):
    # cv / meta delegate to the adjoint integrator (plugins/cv_integrator.py:
    # 79-85); without explicit tangents the parameters' .grad is the tangent
    announce("forward plugin delegation")
    rng, dens, alb, d = _volume(23)
    sc = rcv.load_dict(d)
    key = "box.interior_medium.sigma_t.data"
    p, x = _bind(sc, key, dens)
    t = torch.from_numpy(rng.uniform(-1, 1, dens.shape).astype(np.float32)).cuda()
    x.grad = t.clone()
    cv = rcv.load_dict({"type": "cv", "integrator": {"type": "twostateprbvolpath", "max_depth": 8, "rr_depth": 8}})
    meta = rcv.load_dict({"type": "meta", "method": "cv_ps",
                          "integrator": {"type": "twostateprbvolpath", "max_depth": 8, "rr_depth": 8}})
    direct = cv.integrator.render_forward(sc, p, 0, 3, 8, tangents={key: t})
    assert torch.equal(cv.render_forward(sc, p, 0, 3, 8), direct)
    assert torch.equal(meta.render_forward(sc, p, 0, 3, 8), direct)
    assert float(direct.abs().sum()) > 0
    simple = rcv.load_dict({"type": "volpathsimple"})
    with pytest.raises(NotImplementedError):
        simple.render_forward(sc, p, 0, 3, 8)
