"""Pins of the oracle's volpathsimple (plugins/volpathsimple.py): the primal
against closed forms and against the volume path tracer (both estimate the
same radiance with unbounded paths), and every adjoint variant (DRT with and
without subsampling and MIS, the plain collision estimator, with and without
NEE) against central finite differences. CPU only."""
import numpy as np
import pytest

import render_ref as R
from helpers import box_chord, camera_rays, volume_box_scene
from recursive_control_variates_amd import scene_desc

VOL, SIMPLE = 1, 2
NO_NEE, NO_DRT, NO_SUB, NO_MIS = 1, 2, 4, 8


def _desc(d):
    return scene_desc.describe(d)


def test_white_furnace():
    g = np.full((4, 4, 4, 1), 0.7, np.float32)
    g[1:3, 1:3, 1:3] = 1.5
    d = _desc(volume_box_scene(g, albedo=1.0, scale=3.0, w=6, h=6))
    img, _, var, _ = R.render(d, SIMPLE, -1, 5, 11, 1024, precision=64, with_variance=True)
    se = np.sqrt(var.mean() / (img.size / 3 * 1024))
    assert abs(img.mean() - 1.0) < 5 * se + 1e-12, (img.mean(), se)


def test_absorber_transmittance_closed_form():
    rng = np.random.default_rng(1)
    g = rng.uniform(0.2, 1.0, (3, 3, 3, 1)).astype(np.float32)
    scale, w, h = 2.0, 5, 5
    d = _desc(volume_box_scene(g, albedo=0.0, scale=scale, w=w, h=h, fov=50))
    spp = 2048
    img, _, var, _ = R.render(d, SIMPLE, 16, 16, 3, spp, precision=64, with_variance=True)
    from test_oracle_supergrid import trilinear
    view = d.sensors[0].views[0]
    nj, nq = 10, 400
    jj = (np.arange(nj) + 0.5) / nj
    jx, jy = [a.ravel() for a in np.meshgrid(jj, jj)]
    expect = np.zeros((h, w))
    for py in range(h):
        for px in range(w):
            o, dd = camera_rays(view.to_world, view.fov, w, h, np.full(jx.shape, px), np.full(jx.shape, py), jx, jy)
            with np.errstate(divide="ignore", invalid="ignore"):
                inv = 1.0 / dd
                t0 = np.maximum(np.max(np.minimum((0 - o) * inv, (1 - o) * inv), axis=1), 0)
            L = box_chord(o, dd)
            ts = t0[:, None] + L[:, None] * (np.arange(nq) + 0.5)[None] / nq
            pts = o[:, None, :] + dd[:, None, :] * ts[..., None]
            tau = scale * trilinear(g[..., 0], np.clip(pts, 0, 1)).mean(axis=1) * L
            expect[py, px] = np.exp(-tau).mean()
    z = (img[..., 0] - expect) / np.sqrt(var[..., 0] / spp + 1e-12)
    assert np.abs(z).max() < 5, np.abs(z).max()


def test_primal_matches_volume_path_tracer():
    rng = np.random.default_rng(2)
    g = (0.2 + rng.uniform(0, 0.8, (5, 5, 5, 1))).astype(np.float32)
    alb = rng.uniform(0.5, 0.9, (5, 5, 5, 3)).astype(np.float32)
    from recursive_control_variates_amd.transform import T
    d = _desc(volume_box_scene(g, albedo={"type": "gridvolume", "data": alb, "to_world": T.scale(1.0)},
                               scale=5.0, w=6, h=6))
    spp = 2048
    a, _, va, _ = R.render(d, SIMPLE, -1, 4, 21, spp, precision=64, with_variance=True)
    b, _, vb, _ = R.render(d, VOL, -1, 4, 22, spp, precision=64, with_variance=True)
    c, _, vc, _ = R.render(d, SIMPLE, -1, 4, 23, spp, precision=64, with_variance=True, flags=NO_NEE)
    for x, vx in ((a, va), (c, vc)):
        z = (x - b) / np.sqrt((vx + vb) / spp + 1e-12)
        assert np.abs(z).max() < 5.5
        assert abs(z.mean()) < 0.5


@pytest.fixture(scope="module")
def fd_case():
    # d<g, E[img]>/d(density) along delta by a central difference of two-state
    # volume renders (theta +/- eps delta share paths)
    rng = np.random.default_rng(3)
    grid = (0.3 + rng.uniform(0, 0.7, (4, 4, 4, 1))).astype(np.float32)
    d = _desc(volume_box_scene(grid, albedo=0.8, scale=4.0, w=6, h=6))
    gimg = rng.uniform(0.5, 1.5, (6, 6, 3))
    delta = rng.uniform(-1, 1, grid.shape).astype(np.float32)
    eps = 2e-2
    fds = []
    for s in range(24):
        Hp, Fm = R.render(d, VOL, -1, 4, 400 + s, 8192, params_old=[grid + eps * delta],
                          params_new=[grid - eps * delta], precision=64)
        fds.append(np.sum(gimg * (Hp - Fm)) / (2 * eps))
    return d, gimg, delta, np.array(fds)


@pytest.mark.parametrize("flags", [0, NO_SUB, NO_MIS, NO_DRT, NO_NEE, NO_SUB | NO_MIS])
def test_adjoint_matches_finite_differences(fd_case, flags):
    # every estimator variant of volpathsimple's adjoint has the same expectation
    d, gimg, delta, fds = fd_case
    adj = np.array([np.sum(R.adjoint(d, -1, 4, 600 + s, 8192, gimg, precision=64, integrator_type=SIMPLE,
                                     flags=flags) * delta) for s in range(8)])
    se = np.sqrt(fds.var(ddof=1) / len(fds) + adj.var(ddof=1) / len(adj))
    assert abs(fds.mean() - adj.mean()) < 4 * se + 0.02 * abs(fds.mean()), (fds.mean(), adj.mean(), se)
