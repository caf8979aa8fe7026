"""Pins of the oracle's majorant supergrid and DDA free flight
(plugins/twostatemedium.py:25-70, 82-85): the cell maxima against the
textbook one-voxel dilation and against the trilinear field they must bound,
a one-cell supergrid reducing to the global majorant, a closed-form
transmittance through a heterogeneous absorber, unbiasedness against the
global-majorant tracker (plain and two-state), and the adjoint against
finite differences. CPU only."""
import numpy as np
import pytest

import render_ref as R
from helpers import box_chord, camera_rays, volume_box_scene
from recursive_control_variates_amd import scene_desc

VOL = 1


def _desc(d):
    return scene_desc.describe(d)


def trilinear(grid, p):
    """Cell-centred, edge-clamped trilinear lookup of a (z, y, x) grid at
    points p (..., 3) in the unit cube, float64."""
    n = np.array(grid.shape[::-1], np.float64)          # (nx, ny, nz)
    u = p * n - 0.5
    f = np.floor(u)
    t = u - f
    i0 = np.clip(f.astype(np.int64), 0, n.astype(np.int64) - 1)
    i1 = np.clip(f.astype(np.int64) + 1, 0, n.astype(np.int64) - 1)
    out = 0.0
    for c in range(8):
        ix = i1[..., 0] if c & 1 else i0[..., 0]
        iy = i1[..., 1] if c & 2 else i0[..., 1]
        iz = i1[..., 2] if c & 4 else i0[..., 2]
        w = ((t[..., 0] if c & 1 else 1 - t[..., 0]) * (t[..., 1] if c & 2 else 1 - t[..., 1])
             * (t[..., 2] if c & 4 else 1 - t[..., 2]))
        out = out + w * grid[iz, iy, ix].astype(np.float64)
    return out


def test_supergrid_textbook_dilation():
    # n divisible by k: cell c covers voxels [c k, (c+1) k) and a trilinear
    # lookup inside reads one more voxel on each side: max over
    # [c k - 1, (c+1) k] clamped to the grid
    rng = np.random.default_rng(0)
    g = rng.uniform(0, 1, (8, 8, 8, 1)).astype(np.float32)
    d = _desc(volume_box_scene(g, majorant_factor=2))
    sg = R.majorant_grid(d, [g])
    assert sg.shape == (4, 4, 4)
    k = 2
    expect = np.zeros((4, 4, 4), np.float32)
    for cz in range(4):
        for cy in range(4):
            for cx in range(4):
                sl = [slice(max(c * k - 1, 0), min((c + 1) * k, 7) + 1) for c in (cz, cy, cx)]
                expect[cz, cy, cx] = g[sl[0], sl[1], sl[2], 0].max()
    assert np.array_equal(sg, expect)
    # without a factor there is no supergrid
    assert R.majorant_grid(_desc(volume_box_scene(g)), [g]) is None


def test_supergrid_bounds_trilinear_field_ragged():
    # n not divisible by k (10 x 7 x 5 voxels, k = 3 -> 4 x 3 x 2 cells):
    # every cell value bounds the interpolated density anywhere in the cell,
    # corners and faces included
    rng = np.random.default_rng(1)
    g = rng.uniform(0, 1, (5, 7, 10, 1)).astype(np.float32)
    d = _desc(volume_box_scene(g, majorant_factor=3))
    sg = R.majorant_grid(d, [g])
    assert sg.shape == (2, 3, 4)
    res = np.array([4, 3, 2])
    m = 9
    s = np.linspace(0, 1, m)
    for cz in range(2):
        for cy in range(3):
            for cx in range(4):
                c = np.array([cx, cy, cz])
                lo, hi = c / res, (c + 1) / res
                pts = np.stack(np.meshgrid(*[lo[a] + (hi[a] - lo[a]) * s for a in range(3)], indexing="ij"), -1)
                pts = np.concatenate([pts.reshape(-1, 3), lo + (hi - lo) * rng.uniform(0, 1, (500, 3))])
                vals = trilinear(g[..., 0], pts)
                assert vals.max() <= sg[cz, cy, cx] * (1 + 1e-6), (c, vals.max(), sg[cz, cy, cx])
                # and the bound is not vacuous: it is a voxel within one voxel of the cell
                assert sg[cz, cy, cx] <= g.max()
    assert np.isclose(sg.max(), g.max())


def test_single_cell_supergrid_is_the_global_majorant():
    # one cell holding the grid maximum: the DDA's first segment samples
    # t = mint + tau / m, the global-majorant formula
    rng = np.random.default_rng(2)
    g = rng.uniform(0, 1, (6, 6, 6, 1)).astype(np.float32)
    base = R.render(_desc(volume_box_scene(g, scale=6.0, w=8, h=8)), VOL, 8, 8, 5, 32, precision=64)[1]
    d1 = _desc(volume_box_scene(g, scale=6.0, w=8, h=8, majorant_factor=6))
    assert R.majorant_grid(d1, [g]).shape == (1, 1, 1)
    one = R.render(d1, VOL, 8, 8, 5, 32, precision=64)[1]
    same = np.all(base == one, axis=-1)
    assert same.mean() >= 0.95, same.mean()
    assert abs(one.mean() - base.mean()) < 0.02 * base.mean()


def test_supergrid_transmittance_closed_form():
    # albedo 0: E[L] = exp(-scale * integral of the trilinear density along
    # the camera ray), integrated by the midpoint rule in float64
    rng = np.random.default_rng(3)
    g = rng.uniform(0, 1, (8, 8, 8, 1)).astype(np.float32)
    g[2:5, 3:6, 1:4] *= 4.0                       # a dense core the coarse cells separate
    scale, w, h = 3.0, 6, 6
    d = _desc(volume_box_scene(g, albedo=0.0, scale=scale, w=w, h=h, fov=50, majorant_factor=2))
    spp = 2048
    img, _, var, _ = R.render(d, VOL, 16, 16, 7, spp, precision=64, with_variance=True)
    view = d.sensors[0].views[0]
    nj, nq = 10, 600
    jj = (np.arange(nj) + 0.5) / nj
    jx, jy = [a.ravel() for a in np.meshgrid(jj, jj)]
    expect = np.zeros((h, w))
    for py in range(h):
        for px in range(w):
            o, dd = camera_rays(view.to_world, view.fov, w, h, np.full(jx.shape, px), np.full(jx.shape, py), jx, jy)
            with np.errstate(divide="ignore", invalid="ignore"):
                inv = 1.0 / dd
                t0 = np.max(np.minimum((0 - o) * inv, (1 - o) * inv), axis=1)
            t0 = np.maximum(t0, 0)
            L = box_chord(o, dd)
            ts = t0[:, None] + L[:, None] * (np.arange(nq) + 0.5)[None] / nq
            pts = o[:, None, :] + dd[:, None, :] * ts[..., None]
            tau = scale * trilinear(g[..., 0], np.clip(pts, 0, 1)).mean(axis=1) * L
            expect[py, px] = np.exp(-tau).mean()
    z = (img[..., 0] - expect) / np.sqrt(var[..., 0] / spp + 1e-12)
    assert np.abs(z).max() < 5, np.abs(z).max()
    assert expect.min() < 0.5                      # the medium is optically thick somewhere


def test_supergrid_unbiased_against_global_majorant():
    # plain and two-state renders with the supergrid agree in expectation
    # with the global-majorant tracker (E[H] = plain(old), E[F] = plain(new))
    rng = np.random.default_rng(4)
    ga = rng.uniform(0, 1, (8, 8, 8, 1)).astype(np.float32)
    ga[:3] *= 0.1                                  # sparse region: cells with small majorants
    gb = (ga * rng.uniform(0.5, 1.6, ga.shape)).astype(np.float32)
    spp = 1024
    sg = _desc(volume_box_scene(ga, scale=6.0, w=6, h=6, majorant_factor=2))
    gl = _desc(volume_box_scene(ga, scale=6.0, w=6, h=6))
    H, F, vH, vF = R.render(sg, VOL, 8, 8, 31, spp, params_old=[ga], params_new=[gb], precision=64,
                            with_variance=True)
    Pa, _, vA, _ = R.render(gl, VOL, 8, 8, 32, spp, params_old=[ga], params_new=[ga], precision=64,
                            with_variance=True)
    Pb, _, vB, _ = R.render(gl, VOL, 8, 8, 33, spp, params_old=[gb], params_new=[gb], precision=64,
                            with_variance=True)
    for z in ((H - Pa) / np.sqrt((vH + vA) / spp + 1e-12), (F - Pb) / np.sqrt((vF + vB) / spp + 1e-12)):
        assert np.abs(z).max() < 5.5
        assert abs(z.mean()) < 0.5


def test_supergrid_adjoint_matches_finite_differences():
    rng = np.random.default_rng(5)
    grid = (0.1 + rng.uniform(0, 0.9, (6, 6, 6, 1))).astype(np.float32)
    grid[:2] *= 0.2
    d = _desc(volume_box_scene(grid, albedo=0.8, scale=4.0, w=6, h=6, majorant_factor=2))
    gimg = rng.uniform(0.5, 1.5, (6, 6, 3))
    delta = rng.uniform(-1, 1, grid.shape).astype(np.float32)
    eps = 2e-2
    fds, adj = [], []
    for s in range(4):
        Hp, Fm = R.render(d, VOL, 8, 8, 140 + s, 2048, params_old=[grid + eps * delta],
                          params_new=[grid - eps * delta], precision=64)
        fds.append(np.sum(gimg * (Hp - Fm)) / (2 * eps))
        gr = R.adjoint(d, 8, 8, 160 + s, 2048, gimg, precision=64)
        adj.append(np.sum(gr * delta))
    fds, adj = np.array(fds), np.array(adj)
    se = np.sqrt(fds.var(ddof=1) / len(fds) + adj.var(ddof=1) / len(adj))
    assert abs(fds.mean() - adj.mean()) < 4 * se + 0.02 * abs(fds.mean()), (fds, adj)
