"""GPU path-replay adjoint (K5): against the oracle adjoint with the same
random numbers, against central finite differences (two-state renders of
theta +/- eps delta share their paths), and through the autograd surface."""
import numpy as np
import pytest
import torch

import render_ref as R
from conftest import announce
from helpers import furnace_box_scene, volume_box_scene
import recursive_control_variates_amd as rcv
from recursive_control_variates_amd.optim import l2_loss

pytestmark = pytest.mark.gpu
KEY = "box.interior_medium.sigma_t.data"


def _setup(seed=0, n=5, w=8, h=8, scale=5.0, majorant_factor=0):
    rng = np.random.default_rng(seed)
    grid = (0.2 + rng.uniform(0, 0.8, (n, n, n, 1))).astype(np.float32)
    sc = rcv.load_dict(volume_box_scene(grid, albedo=0.8, scale=scale, w=w, h=h, spp=16,
                                        majorant_factor=majorant_factor))
    return rng, grid, sc


def _adjoint(sc, integ, gimg, seed, spp):
    p = rcv.traverse(sc)
    x = p[KEY].detach().clone().requires_grad_(True)
    p[KEY] = x
    p.update()
    p.update()
    grads = integ.adjoint(sc, p, torch.from_numpy(gimg.astype(np.float32)).cuda(), 0, seed, spp)
    return grads[KEY].cpu().numpy()


@pytest.mark.parametrize("factor", [0, 2])
def test_adjoint_matches_oracle_same_samples(cuda, factor):
    # factor 2: free flights and NEE ratio tracking through a 3^3 supergrid
    announce(f"adjoint vs oracle fp32 (majorant factor {factor})")
    rng, grid, sc = _setup(1, majorant_factor=factor)
    gimg = rng.uniform(0.5, 1.5, (8, 8, 3))
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8})
    g_gpu = _adjoint(sc, integ, gimg, 31, 32)
    g_cpu = R.adjoint(sc.desc, 8, 8, 31, 32, gimg, precision=32)
    scale = np.abs(g_cpu).max()
    close = np.abs(g_gpu - g_cpu) <= 1e-3 * scale + 1e-3 * np.abs(g_cpu)
    assert close.mean() > 0.97, close.mean()
    assert abs(g_gpu.sum() - g_cpu.sum()) < 0.03 * np.abs(g_cpu).sum()


def test_adjoint_with_tape_overflows_matches_oracle_same_samples(cuda):
    # With a supergrid the adjoint records each path's collisions in its
    # primal pass and scatters them once the radiance is known; a path with
    # more collisions than the record holds is replayed (collision terms only,
    # its NEE terms were scattered in the primal pass). A thick, bright medium
    # with deep paths makes both kinds frequent.
    announce("adjoint vs oracle fp32 (supergrid, thick, tape overflows)")
    rng = np.random.default_rng(4)
    grid = (0.2 + rng.uniform(0, 0.8, (5, 5, 5, 1))).astype(np.float32)
    w, h, spp, depth = 8, 8, 16, 40
    sc = rcv.load_dict(volume_box_scene(grid, albedo=0.95, scale=40.0, w=w, h=h, spp=spp, majorant_factor=2))
    gimg = rng.uniform(0.5, 1.5, (h, w, 3))
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": depth, "rr_depth": depth})
    sc.set_counting(True)
    g_gpu = _adjoint(sc, integ, gimg, 13, spp)
    cnt = sc.read_counters()
    sc.set_counting(False)
    replays = cnt["paths"] - w * h * spp
    assert 0 < replays < w * h * spp, cnt
    g_cpu = R.adjoint(sc.desc, depth, depth, 13, spp, gimg, precision=32)
    scale = np.abs(g_cpu).max()
    close = np.abs(g_gpu - g_cpu) <= 1e-3 * scale + 1e-3 * np.abs(g_cpu)
    assert close.mean() > 0.97, close.mean()
    assert abs(g_gpu.sum() - g_cpu.sum()) < 0.03 * np.abs(g_cpu).sum()


def test_adjoint_with_more_nees_than_slots_matches_oracle_same_samples(cuda):
    # Without a supergrid the primal pass keeps dLoss/dL . c of each NEE in
    # the lane's LDS slots (at most 32) and the replay scatters the shadow
    # ray's gradient while tracking it; NEEs past the slots are tracked a
    # second time in the replay. A thick, bright medium with paths of up to
    # 64 scattering events makes most paths go past the slots.
    announce("adjoint vs oracle fp32 (thick, more NEEs per path than LDS slots)")
    rng = np.random.default_rng(6)
    grid = (0.2 + rng.uniform(0, 0.8, (5, 5, 5, 1))).astype(np.float32)
    w, h, spp, depth = 8, 8, 16, 64
    sc = rcv.load_dict(volume_box_scene(grid, albedo=0.95, scale=40.0, w=w, h=h, spp=spp))
    gimg = rng.uniform(0.5, 1.5, (h, w, 3))
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": depth, "rr_depth": depth})
    sc.set_counting(True)
    integ.render(sc, 0, 17, spp)
    primal = sc.read_counters()
    sc.set_counting(False)
    sc.set_counting(True)
    g_gpu = _adjoint(sc, integ, gimg, 17, spp)
    cnt = sc.read_counters()
    sc.set_counting(False)
    # the adjoint's primal and replay passes track every shadow ray once each
    # (the render of the same samples, twice); past the slots, once more
    assert cnt["nee_steps"] > 2.1 * primal["nee_steps"], (cnt, primal)
    g_cpu = R.adjoint(sc.desc, depth, depth, 17, spp, gimg, precision=32)
    scale = np.abs(g_cpu).max()
    close = np.abs(g_gpu - g_cpu) <= 1e-3 * scale + 1e-3 * np.abs(g_cpu)
    assert close.mean() > 0.97, close.mean()
    assert abs(g_gpu.sum() - g_cpu.sum()) < 0.03 * np.abs(g_cpu).sum()


def test_adjoint_matches_finite_differences(cuda):
    announce("adjoint vs FD")
    rng, grid, sc = _setup(2)
    gimg = rng.uniform(0.5, 1.5, (8, 8, 3))
    delta = rng.uniform(-1, 1, grid.shape).astype(np.float32)
    eps = 2e-2
    ts = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 8, "rr_depth": 8})
    plain = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8})
    p = rcv.traverse(sc)
    fds, adj = [], []
    for s in range(8):
        p[KEY] = torch.from_numpy(grid + eps * delta).cuda()
        p.update()
        p[KEY] = torch.from_numpy(grid - eps * delta).cuda()
        p.update()                                  # old = +eps, new = -eps
        H, F = ts.render_twostates(sc, 0, 500 + s, 1024)
        fds.append(float(np.sum(gimg * (H - F).cpu().numpy()) / (2 * eps)))
        p[KEY] = torch.from_numpy(grid).cuda()
        p.update()
        p.update()
        adj.append(float(np.sum(_adjoint(sc, plain, gimg, 700 + s, 1024) * delta)))
    fds, adj = np.array(fds), np.array(adj)
    se = np.sqrt(fds.var(ddof=1) / len(fds) + adj.var(ddof=1) / len(adj))
    assert abs(fds.mean() - adj.mean()) < 4 * se + 0.02 * abs(fds.mean()), (fds.mean(), adj.mean(), se)


def test_autograd_render_and_loss(cuda):
    announce("adjoint autograd")
    rng, grid, sc = _setup(3)
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8})
    ref = torch.rand(8, 8, 3, device="cuda")
    p = rcv.traverse(sc)
    x = p[KEY].detach().clone().requires_grad_(True)
    p[KEY] = x
    p.update()
    p.update()
    img = rcv.render(sc, p, integrator=integ, seed=4, spp=16)
    loss = l2_loss(img, ref)
    np.testing.assert_allclose(float(loss), float(((img.detach() - ref) ** 2).mean()), rtol=1e-5)
    loss.backward()
    from recursive_control_variates_amd.rendering import sample_tea_32
    gimg = (2 * (img.detach() - ref) / img.numel()).cpu().numpy()
    direct = _adjoint(sc, integ, gimg, sample_tea_32(4, 1)[0], 16)
    np.testing.assert_allclose(x.grad.cpu().numpy(), direct, rtol=1e-4, atol=1e-7)


def test_texture_adjoint_matches_finite_differences(cuda):
    # surfaces: dL/d(texture) by path replay vs two-state central differences
    announce("texture adjoint vs FD")
    rng = np.random.default_rng(9)
    tex = rng.uniform(0.3, 0.7, (4, 4, 3)).astype(np.float32)
    sc = rcv.load_dict(furnace_box_scene(albedo=0.5, w=8, h=8, textured=True, tex=tex))
    key = "back.bsdf.reflectance.data"
    gimg = rng.uniform(0.5, 1.5, (8, 8, 3))
    delta = rng.uniform(-1, 1, tex.shape).astype(np.float32)
    eps = 2e-2
    ts = rcv.load_dict({"type": "twostateprb", "max_depth": 4, "rr_depth": 10})
    plain = rcv.load_dict({"type": "prb", "max_depth": 4, "rr_depth": 10})
    p = rcv.traverse(sc)
    fds, adj = [], []
    for s in range(6):
        p[key] = torch.from_numpy(tex + eps * delta).cuda()
        p.update()
        p[key] = torch.from_numpy(tex - eps * delta).cuda()
        p.update()
        H, F = ts.render_twostates(sc, 0, 300 + s, 512)
        fds.append(float(np.sum(gimg * (H - F).cpu().numpy()) / (2 * eps)))
        x = torch.from_numpy(tex).cuda().requires_grad_(True)
        p[key] = x
        p.update()
        p.update()
        g = plain.adjoint(sc, p, torch.from_numpy(gimg.astype(np.float32)).cuda(), 0, 400 + s, 512)[key]
        adj.append(float(np.sum(g.cpu().numpy() * delta)))
    fds, adj = np.array(fds), np.array(adj)
    se = np.sqrt(fds.var(ddof=1) / len(fds) + adj.var(ddof=1) / len(adj))
    assert abs(fds.mean() - adj.mean()) < 4 * se + 0.02 * abs(fds.mean()), (fds.mean(), adj.mean(), se)


def test_albedo_grid_adjoint_matches_finite_differences(cuda):
    # dL/d(albedo grid (n, n, n, 3)) by path replay vs two-state central differences
    announce("albedo adjoint vs FD")
    from recursive_control_variates_amd.transform import T
    rng = np.random.default_rng(12)
    n = 4
    dens = (0.4 + rng.uniform(0, 0.6, (n, n, n, 1))).astype(np.float32)
    alb = rng.uniform(0.4, 0.9, (n, n, n, 3)).astype(np.float32)
    alb_dict = {"type": "gridvolume", "data": alb, "to_world": T.scale(1.0)}
    sc = rcv.load_dict(volume_box_scene(dens, albedo=alb_dict, scale=5.0, w=8, h=8, spp=16))
    key = "box.interior_medium.albedo.data"
    gimg = rng.uniform(0.5, 1.5, (8, 8, 3))
    delta = rng.uniform(-1, 1, alb.shape).astype(np.float32)
    eps = 2e-2
    ts = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 8, "rr_depth": 8})
    plain = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8})
    p = rcv.traverse(sc)
    fds, adj = [], []
    for s in range(6):
        p[key] = torch.from_numpy(alb + eps * delta).cuda()
        p.update()
        p[key] = torch.from_numpy(alb - eps * delta).cuda()
        p.update()
        H, F = ts.render_twostates(sc, 0, 800 + s, 1024)
        fds.append(float(np.sum(gimg * (H - F).cpu().numpy()) / (2 * eps)))
        x = torch.from_numpy(alb).cuda().requires_grad_(True)
        p[key] = x
        p.update()
        p.update()
        g = plain.adjoint(sc, p, torch.from_numpy(gimg.astype(np.float32)).cuda(), 0, 900 + s, 1024)[key]
        adj.append(float(np.sum(g.cpu().numpy() * delta)))
    fds, adj = np.array(fds), np.array(adj)
    se = np.sqrt(fds.var(ddof=1) / len(fds) + adj.var(ddof=1) / len(adj))
    assert abs(fds.mean() - adj.mean()) < 4 * se + 0.02 * abs(fds.mean()), (fds.mean(), adj.mean(), se)


@pytest.mark.parametrize("which", ["sigma_t", "albedo"])
def test_adjoint_equals_finite_difference_with_common_random_numbers(cuda, which):
    # The two-state render of (theta + eps delta, theta - eps delta) and the
    # path-replay adjoint at theta trace the same paths for the same seed: the
    # central difference of the shared-path estimator is then the per-path
    # derivative the adjoint scatters, up to O(eps^2) and float rounding --
    # seed by seed, with no Monte Carlo tolerance (dense medium, paths long
    # enough to reach max_depth).
    announce(f"adjoint == FD with common random numbers ({which})")
    from recursive_control_variates_amd.transform import T
    rng = np.random.default_rng(21)
    n = 6
    dens = (0.3 + rng.uniform(0, 0.7, (n, n, n, 1))).astype(np.float32)
    alb = rng.uniform(0.5, 0.95, (n, n, n, 3)).astype(np.float32)
    alb_dict = {"type": "gridvolume", "data": alb, "to_world": T.scale(1.0)}
    sc = rcv.load_dict(volume_box_scene(dens, albedo=alb_dict, scale=12.0, w=16, h=12, spp=8))
    key = f"box.interior_medium.{which}.data"
    base = dens if which == "sigma_t" else alb
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (12, 16, 3)).astype(np.float32)).cuda()
    delta = rng.uniform(-1, 1, base.shape).astype(np.float32)
    # the majorant (grid maximum) must not move, or free flights would differ
    delta[base > 0.9 * base.max()] = 0.0
    eps = 1e-3 * float(base.max())
    ts = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 12, "rr_depth": 12})
    plain = rcv.load_dict({"type": "prbvolpath", "max_depth": 12, "rr_depth": 12})
    p = rcv.traverse(sc)
    for s in range(3):
        p[key] = torch.from_numpy(base + eps * delta).cuda()
        p.update()
        p[key] = torch.from_numpy(base - eps * delta).cuda()
        p.update()
        H, F = ts.render_twostates(sc, 0, 40 + s, 64)
        fd = float(((H.double() - F.double()) * gimg.double()).sum()) / (2 * eps)
        p[key] = torch.from_numpy(base).cuda().requires_grad_(True)
        p.update()
        p.update()
        g = plain.adjoint(sc, p, gimg, 0, 40 + s, 64)[key]
        ad = float((g.double() * torch.from_numpy(delta).cuda().double()).sum())
        assert abs(fd - ad) <= 0.01 * abs(fd), (s, fd, ad)


@pytest.mark.parametrize("kind", ["prb", "volpathsimple"])
def test_persistent_adjoint_equals_static_launch(cuda, kind):
    # The per-sample adjoints (surface path tracer, volpathsimple) run as a
    # persistent launch by default and as a static one-thread-per-chunk launch
    # with samples_per_thread > 0; both trace the same samples, so the
    # gradients differ only by the order of the float scatter-adds.
    announce(f"persistent vs static adjoint launch ({kind})")
    from recursive_control_variates_amd import scenes
    rng = np.random.default_rng(9)
    if kind == "prb":
        sc = rcv.load_dict(scenes.cornell_scene_dict(scenes.texture_target(16, 1), twostate=False, film_res=40,
                                                     spp=32))
        key = scenes.TEXTURE_KEY
        integ = rcv.load_dict({"type": "prb", "max_depth": 5, "rr_depth": 3})
    else:
        rng2, grid, sc = _setup(9, w=40, h=30)
        key = KEY
        integ = rcv.load_dict({"type": "volpathsimple", "max_depth": 8, "rr_depth": 4})
    h, w = sc.film_shape(0)[:2]
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (h, w, 3)).astype(np.float32)).cuda()
    p = rcv.traverse(sc)
    p[key] = p[key].detach().clone().requires_grad_(True)
    p.update()
    p.update()
    grads = []
    for spt in (0, 4):
        integ.samples_per_thread = spt
        grads.append(integ.adjoint(sc, p, gimg, 0, 17, 32)[key].double().cpu().numpy())
    a, b = grads
    scale = np.abs(b).max()
    assert scale > 0
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5 * scale), np.abs(a - b).max() / scale


def test_scatter_log_across_tiles_equals_atomics(cuda):
    # The persistent volume adjoint logs the density scatter-adds and adds
    # them up per 16384-voxel tile (in 64-bit fixed point): a 48 x 48 x 40
    # grid spans six tiles whose boundaries cut through a slice, so records
    # go to tile bins and to the seam bins read by two tiles. The static
    # launch (samples_per_thread > 0) traces the same samples and adds the
    # same products with float atomics.
    announce("scatter log over six tiles vs atomics (prbvolpath)")
    rng = np.random.default_rng(21)
    grid = (0.2 + rng.uniform(0, 0.8, (40, 48, 48, 1))).astype(np.float32)
    sc = rcv.load_dict(volume_box_scene(grid, albedo=0.8, scale=5.0, w=64, h=48, spp=32))
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (48, 64, 3)).astype(np.float32)).cuda()
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8})
    p = rcv.traverse(sc)
    p[KEY] = p[KEY].detach().clone().requires_grad_(True)
    p.update()
    p.update()
    grads = []
    for spt in (0, 4):
        integ.samples_per_thread = spt
        grads.append(integ.adjoint(sc, p, gimg, 0, 23, 32)[KEY].double().cpu().numpy())
    a, b = grads
    scale = np.abs(b).max()
    assert scale > 0 and (b != 0).mean() > 0.5, (b != 0).mean()
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5 * scale), np.abs(a - b).max() / scale


@pytest.mark.parametrize("chunks", ["1", "40"])
def test_scatter_log_past_the_pool_equals_atomics(cuda, monkeypatch, chunks):
    # With the chunk pool capped (RCV_SCATTER_LOG_CHUNKS) the wavefronts draw
    # it empty early in the launch; from then on their scatter-adds fall back
    # to atomics, and the chunks they filled are still added up.
    announce(f"scatter log with a pool of {chunks} chunks vs atomics (prbvolpath)")
    monkeypatch.setenv("RCV_SCATTER_LOG_CHUNKS", chunks)
    rng = np.random.default_rng(23)
    grid = (0.2 + rng.uniform(0, 0.8, (40, 48, 48, 1))).astype(np.float32)
    sc = rcv.load_dict(volume_box_scene(grid, albedo=0.8, scale=5.0, w=48, h=32, spp=32))
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (32, 48, 3)).astype(np.float32)).cuda()
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8})
    p = rcv.traverse(sc)
    p[KEY] = p[KEY].detach().clone().requires_grad_(True)
    p.update()
    p.update()
    grads = []
    for spt in (0, 4):
        integ.samples_per_thread = spt
        grads.append(integ.adjoint(sc, p, gimg, 0, 37, 32)[KEY].double().cpu().numpy())
    a, b = grads
    scale = np.abs(b).max()
    assert scale > 0 and (b != 0).mean() > 0.5, (b != 0).mean()
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5 * scale), np.abs(a - b).max() / scale


def test_scatter_log_non_finite_values_match_atomics(cuda):
    # A tile whose records hold a NaN is summed with float atomics, as the
    # static launch does: the same voxels come out NaN, the rest as before.
    announce("scatter log with a NaN image gradient vs atomics (prbvolpath)")
    rng = np.random.default_rng(22)
    grid = (0.2 + rng.uniform(0, 0.8, (24, 24, 24, 1))).astype(np.float32)
    sc = rcv.load_dict(volume_box_scene(grid, albedo=0.8, scale=5.0, w=32, h=24, spp=32))
    gimg_np = rng.uniform(0.5, 1.5, (24, 32, 3)).astype(np.float32)
    gimg_np[12, 16, 1] = np.nan
    gimg = torch.from_numpy(gimg_np).cuda()
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8})
    p = rcv.traverse(sc)
    p[KEY] = p[KEY].detach().clone().requires_grad_(True)
    p.update()
    p.update()
    grads = []
    for spt in (0, 4):
        integ.samples_per_thread = spt
        grads.append(integ.adjoint(sc, p, gimg, 0, 29, 32)[KEY].double().cpu().numpy())
    a, b = grads
    assert np.isnan(b).any() and not np.isnan(b).all()
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b)
    scale = np.abs(b[ok]).max()
    assert np.allclose(a[ok], b[ok], rtol=1e-4, atol=1e-5 * scale), np.abs(a[ok] - b[ok]).max() / scale


def _param_index(desc, key):
    return [p.key for p in desc.params].index(key)


def _close_to_oracle(g_gpu, g_cpu, frac=0.97):
    scale = np.abs(g_cpu).max()
    assert scale > 0
    close = np.abs(g_gpu - g_cpu) <= 1e-3 * scale + 1e-3 * np.abs(g_cpu)
    assert close.mean() > frac, close.mean()
    assert abs(g_gpu.sum() - g_cpu.sum()) < 0.03 * np.abs(g_cpu).sum()
    # per channel as well: a mixed-up channel shows here
    for c in range(g_cpu.shape[-1]):
        a, b = g_gpu[..., c].sum(), g_cpu[..., c].sum()
        assert abs(a - b) < 0.03 * np.abs(g_cpu[..., c]).sum(), (c, a, b)


def test_albedo_adjoint_matches_oracle_same_samples(cuda):
    # dL/d(albedo grid) of the HIP path-replay adjoint against the oracle's
    # recorded-primal-pass adjoint, same random numbers (float32 both)
    announce("albedo adjoint vs oracle fp32")
    from recursive_control_variates_amd.transform import T
    rng = np.random.default_rng(31)
    n = 5
    dens = (0.3 + rng.uniform(0, 0.7, (n, n, n, 1))).astype(np.float32)
    alb = (np.array([0.25, 0.55, 0.9], np.float32) + rng.uniform(-0.05, 0.05, (n, n, n, 3))).astype(np.float32)
    sc = rcv.load_dict(volume_box_scene(dens, albedo={"type": "gridvolume", "data": alb, "to_world": T.scale(1.0)},
                                        scale=5.0, w=8, h=8, spp=32))
    key = "box.interior_medium.albedo.data"
    gimg = rng.uniform(0.5, 1.5, (8, 8, 3))
    integ = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8})
    p = rcv.traverse(sc)
    p[key] = p[key].detach().clone().requires_grad_(True)
    p.update()
    p.update()
    g_gpu = integ.adjoint(sc, p, torch.from_numpy(gimg.astype(np.float32)).cuda(), 0, 37, 32)[key].cpu().numpy()
    g_cpu = R.adjoint(sc.desc, 8, 8, 37, 32, gimg, param=_param_index(sc.desc, key), precision=32)
    assert g_gpu.shape == g_cpu.shape == alb.shape
    _close_to_oracle(g_gpu, g_cpu)


def test_texture_adjoint_matches_oracle_same_samples(cuda):
    # dL/d(back-wall texture) of the HIP surface adjoint (Cornell box of the
    # surface workload) against the oracle's, same random numbers
    announce("texture adjoint vs oracle fp32")
    from recursive_control_variates_amd import scenes
    rng = np.random.default_rng(32)
    sc = rcv.load_dict(scenes.cornell_scene_dict(scenes.texture_target(16, 0), twostate=False, film_res=32, spp=16))
    key = scenes.TEXTURE_KEY
    gimg = rng.uniform(0.5, 1.5, (32, 32, 3))
    integ = rcv.load_dict({"type": "prb", "max_depth": 4})
    p = rcv.traverse(sc)
    p[key] = p[key].detach().clone().requires_grad_(True)
    p.update()
    p.update()
    g_gpu = integ.adjoint(sc, p, torch.from_numpy(gimg.astype(np.float32)).cuda(), 0, 41, 16)[key].cpu().numpy()
    g_cpu = R.adjoint(sc.desc, 4, 5, 41, 16, gimg, param=_param_index(sc.desc, key), precision=32,
                      integrator_type=0)
    assert g_gpu.shape == g_cpu.shape == (16, 16, 3)
    _close_to_oracle(g_gpu, g_cpu)
