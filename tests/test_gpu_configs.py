"""The two largest workloads at their stated sizes:

* 128^3 density + 128^3 x 3 albedo grid, 4 views of 512^2, 32 spp: the
  path-replay adjoint against the shared-path central difference of the
  two-state render, per seed, along directions supported on 32 seed-chosen
  voxels of each grid;
* the 64-job batch (seeds 100..163, image sides U{64..512}, grids
  {32, 64, 128}^3, spp U{4..64}, max_depth 8-32): one graph-captured batched
  iteration of every job, a subset checked against the single-job driver.
"""
import time

import numpy as np
import pytest
import torch

from conftest import announce
import recursive_control_variates_amd as rcv

pytestmark = pytest.mark.gpu


def test_large_grid_adjoint_equals_shared_path_difference(cuda):
    announce("128^3 density + albedo, 4 x 512^2, 32 spp: adjoint vs shared-path central difference")
    from recursive_control_variates_amd import scenes
    res, grid, spp = 512, 128, 32
    dens = scenes.blob_grid(grid, 1)
    alb = scenes.albedo_grid(grid, 1)
    sc = rcv.load_dict(scenes.volume_scene_dict(dens, film_res=res, spp=spp, albedo_grid_data=alb))
    assert sc.film_shape(0) == (res, 4 * res, 3)
    p = rcv.traverse(sc)
    ts = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 16, "rr_depth": 16})
    prb = rcv.load_dict({"type": "prbvolpath", "max_depth": 16, "rr_depth": 16})
    n_pix = res * res * 4
    rng = np.random.default_rng(7)
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (res, 4 * res, 3)).astype(np.float32) / (n_pix * 3)).cuda()
    for key, base in ((scenes.VOLUME_KEY, dens), (scenes.ALBEDO_KEY, alb)):
        # 32 voxels inside the medium, away from the grid maximum (moving the
        # maximum would move the majorant and with it every free flight)
        ok = (dens.reshape(-1) > 0.3 * dens.max()) & (base.reshape(-1, base.shape[-1]).max(1) < 0.9 * base.max())
        vox = np.random.default_rng(11).choice(np.flatnonzero(ok), 32, replace=False)
        delta = np.zeros((dens.size, base.shape[-1]), np.float32)
        delta[vox] = np.random.default_rng(100).uniform(-1, 1, (32, base.shape[-1]))
        delta = delta.reshape(base.shape)
        eps = 1e-3 * float(np.abs(base).max())
        for s in range(2):
            t0 = time.perf_counter()
            p[key] = torch.from_numpy(base + eps * delta).cuda()
            p.update()
            p[key] = torch.from_numpy(base - eps * delta).cuda()
            p.update()
            H, F = ts.render_twostates(sc, 0, 500 + s, spp)
            fd = float(((H.double() - F.double()) * gimg.double()).sum()) / (2 * eps)
            x = torch.from_numpy(base).cuda().requires_grad_(True)
            for _ in range(2):
                p[key] = x
                p.update()
            g = prb.adjoint(sc, p, gimg, 0, 500 + s, spp)[key]
            ad = float((g.double() * torch.from_numpy(delta).cuda().double()).sum())
            torch.cuda.synchronize()
            print(f"[config3] {key} seed {s}: fd {fd:.6e} adjoint {ad:.6e} "
                  f"({time.perf_counter() - t0:.2f}s)", flush=True)
            assert abs(fd) > 0
            assert abs(fd - ad) <= 0.03 * abs(fd), (key, s, fd, ad)
        # restore old = new = the base value for the next key
        for _ in range(2):
            p[key] = torch.from_numpy(base).cuda()
            p.update()


def test_sixty_four_job_batch_follows_the_single_job_driver(cuda):
    announce("64-job batch (seeds 100..163), graph-captured, subset vs single-job driver")
    from recursive_control_variates_amd.batch import BatchOptimizer
    from recursive_control_variates_amd.cv_integrator import MetaIntegrator
    from recursive_control_variates_amd.optim import Adam, l1_loss
    from recursive_control_variates_amd.rendering import render
    from recursive_control_variates_amd.serve import make_job
    seeds = list(range(100, 164))
    t0 = time.perf_counter()
    made = [make_job(s, "volume", "cv_ps") for s in seeds]
    sides = [m["side"] for _, m in made]
    grids = [m["grid"] for _, m in made]
    spps = [m["spp"] for _, m in made]
    depths = [m["max_depth"] for _, m in made]
    print(f"[config4] {len(made)} jobs built in {time.perf_counter() - t0:.1f}s; sides {min(sides)}..{max(sides)}, "
          f"grids {sorted(set(grids))}, spp {min(spps)}..{max(spps)}, depth {min(depths)}..{max(depths)}",
          flush=True)
    assert all(64 <= s <= 512 and s % 32 == 0 for s in sides)
    assert set(grids) <= {32, 64, 128} and len(set(grids)) >= 2
    assert all(4 <= s <= 64 for s in spps) and all(8 <= d <= 32 for d in depths)
    batch = BatchOptimizer([sp for sp, _ in made], use_graph=True)
    t0 = time.perf_counter()
    batch.step(1)
    batch.synchronize()
    print(f"[config4] first (captured) iteration {time.perf_counter() - t0:.2f}s", flush=True)
    first = {j: batch.image(j).detach().clone() for j in range(len(seeds))}
    batch.step(1)
    batch.synchronize()
    for j in range(len(seeds)):
        assert batch.job_state(j)[0] == 2
        assert torch.isfinite(batch.param(j)).all()
    # the three smallest jobs and the largest one through the single-job driver
    order = sorted(range(len(seeds)), key=lambda j: sides[j] ** 2 * spps[j])
    for j in order[:3] + order[-1:]:
        spec, _ = make_job(seeds[j], "volume", "cv_ps")
        base = rcv.load_dict(dict(spec.integrator, type="twostateprbvolpath"))
        meta = MetaIntegrator({"method": "cv_ps", "integrator": base})
        p = rcv.traverse(spec.scene)
        opt = Adam(lr=spec.lr)
        opt.set_clamp(spec.key, *spec.clamp)
        for _ in range(2):
            opt[spec.key] = torch.as_tensor(spec.init).cuda()
            p.update(opt)
        for i in range(2):
            p.update(opt)
            img = render(spec.scene, p, integrator=meta, seed=spec.seed + i, spp=spec.spp, spp_grad=spec.spp_grad)
            if i == 0:
                np.testing.assert_allclose(first[j].cpu().numpy(), img.detach().cpu().numpy(), rtol=1e-5,
                                           atol=1e-6)
            l1_loss(img, spec.img_ref).backward()
            opt.step()
        d = np.abs(batch.param(j).cpu().numpy() - opt[spec.key].detach().cpu().numpy())
        print(f"[config4] job {j} (side {sides[j]}, grid {grids[j]}, spp {spps[j]}): "
              f"param max diff {d.max():.3g}, median {np.median(d):.3g}", flush=True)
        assert d.max() <= 2 * spec.lr * 2 + 1e-6 and np.median(d) <= 1e-4
    batch.close()
