"""Batched optimisations (rcv_batch_*): a batch of one job follows the
single-job driver; a job's results are the same alone and in a mix of
uneven jobs; hipGraph replay matches eager launches; the autotuner runs."""
import numpy as np
import pytest
import torch

from conftest import announce
from helpers import furnace_box_scene, volume_box_scene
import recursive_control_variates_amd as rcv
from recursive_control_variates_amd.batch import BatchOptimizer, JobSpec, autotune
from recursive_control_variates_amd.cv_integrator import MetaIntegrator
from recursive_control_variates_amd.optim import Adam, l1_loss
from recursive_control_variates_amd.rendering import render

pytestmark = pytest.mark.gpu
VKEY = "box.interior_medium.sigma_t.data"
TKEY = "back.bsdf.reflectance.data"


def vol_spec(seed, n=6, w=12, h=10, spp=8, method="cv_ps", max_depth=8, twostate=True, majorant_factor=0):
    rng = np.random.default_rng(seed)
    target = rng.uniform(0.2, 1.0, (n, n, n, 1)).astype(np.float32)
    init = np.full_like(target, 0.5)
    d = volume_box_scene(init, scale=5.0, w=w, h=h, spp=spp, twostate=twostate, majorant_factor=majorant_factor)
    sc = rcv.load_dict(d)
    p = rcv.traverse(sc)
    p[VKEY] = torch.from_numpy(target).cuda()
    p.update()
    p.update()
    ref = rcv.load_dict({"type": "prbvolpath", "max_depth": max_depth, "rr_depth": max_depth}).render(sc, 0, 999, 64)
    p[VKEY] = torch.from_numpy(init).cuda()
    p.update()
    p.update()
    return JobSpec(scene=sc, key=VKEY, init=init, img_ref=ref, method=method, spp=spp, seed=0,
                   integrator={"type": "prbvolpath", "max_depth": max_depth, "rr_depth": max_depth},
                   lr=0.02, clamp=(0.0, 1e9))


def surf_spec(seed, w=10, h=10, spp=8, method="cv_ps"):
    rng = np.random.default_rng(seed)
    tex = rng.uniform(0.2, 0.8, (4, 4, 3)).astype(np.float32)
    sc = rcv.load_dict(furnace_box_scene(albedo=0.5, w=w, h=h, spp=spp, textured=True, tex=tex))
    ref = torch.from_numpy(rng.uniform(1.0, 2.0, (h, w, 3)).astype(np.float32)).cuda()
    init = np.full((4, 4, 3), 0.5, np.float32)
    return JobSpec(scene=sc, key=TKEY, init=init, img_ref=ref, method=method, spp=spp, seed=0,
                   integrator={"type": "prb", "max_depth": 3, "rr_depth": 10}, lr=0.02, clamp=(0.0, 1.0))


def driver_run(spec, steps):
    """The single-job path: meta integrator + autograd render + Adam."""
    sc = spec.scene
    base = rcv.load_dict(dict(spec.integrator, type="twostateprbvolpath" if spec.method == "cv_ps"
                              else "prbvolpath"))
    meta = MetaIntegrator({"method": spec.method, "integrator": base})
    p = rcv.traverse(sc)
    opt = Adam(lr=spec.lr)
    opt.set_clamp(spec.key, *spec.clamp)
    for _ in range(2 if spec.method == "cv_ps" else 1):
        opt[spec.key] = torch.as_tensor(spec.init).cuda()
        p.update(opt)
    imgs = []
    for i in range(steps):
        p.update(opt)
        img = render(sc, p, integrator=meta, seed=i, spp=spec.spp, spp_grad=spec.spp)
        imgs.append(img.detach().clone())
        l1_loss(img, spec.img_ref).backward()
        opt.step()
    return opt[spec.key].detach(), imgs


def test_batch_of_one_follows_driver(cuda):
    announce("batch of one vs driver")
    spec = vol_spec(1)
    ref_param, ref_imgs = driver_run(vol_spec(1), 3)
    b = BatchOptimizer([spec], use_graph=False)
    b.step(1)
    np.testing.assert_allclose(b.image(0).cpu().numpy(), ref_imgs[0].cpu().numpy(), rtol=1e-5, atol=1e-6)
    b.step(2)
    # gradients are float-atomic sums in a different order (the batch splits
    # a pixel's samples over more threads); Adam's first steps move by about
    # lr * sign(g), so near-zero gradients may differ by up to 2 lr per step
    _params_close(b.param(0), ref_param, 0.02, 3)
    it, _ = b.job_state(0)
    assert it == 3


def test_supergrid_job_follows_driver(cuda):
    # the batch rebuilds the optimised grid's supergrids after every rotation,
    # as the single-job path does when the parameter is rebound
    announce("batch supergrid job vs driver")
    mk = lambda: vol_spec(11, n=8, w=12, h=10, spp=8, majorant_factor=2)
    ref_param, ref_imgs = driver_run(mk(), 3)
    b = BatchOptimizer([mk()], use_graph=True)
    b.step(1)
    b.synchronize()
    np.testing.assert_allclose(b.image(0).cpu().numpy(), ref_imgs[0].cpu().numpy(), rtol=1e-5, atol=1e-6)
    b.step(2)
    b.synchronize()
    _params_close(b.param(0), ref_param, 0.02, 3)
    assert b.job_state(0)[0] == 3


def _grads_close(a, b):
    a = a.cpu().numpy()
    b = b.cpu().numpy()
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-5 * max(np.abs(b).max(), 1e-12))


def _params_close(a, b, lr, steps):
    # float atomics make gradient sums order-dependent in the last bits; Adam's
    # early steps are ~lr * sign(g), so an element whose gradient is near 0
    # may move by up to 2 lr per step differently; the bulk stays identical
    d = np.abs(a.cpu().numpy() - b.cpu().numpy())
    assert d.max() <= 2 * lr * steps + 1e-6, d.max()
    assert np.median(d) <= 1e-4, np.median(d)


def test_same_alone_and_in_the_mix(cuda):
    announce("alone vs mix")
    mk = [lambda: vol_spec(2, n=6, w=12, h=10, spp=8),
          lambda: vol_spec(3, n=8, w=20, h=8, spp=4, method="baseline", twostate=False),
          lambda: surf_spec(4, w=16, h=12, spp=8),
          lambda: vol_spec(5, n=4, w=8, h=8, spp=16, max_depth=12)]
    for j in (0, 2):
        alone = BatchOptimizer([mk[j]()], use_graph=False)
        mix = BatchOptimizer([f() for f in mk], use_graph=False)
        alone.step(1)
        mix.step(1)
        # first iteration: identical images (same random streams, same kernels)
        assert torch.equal(alone.image(0), mix.image(j))
        _grads_close(mix.jobs[j].grad, alone.jobs[0].grad)
        alone.step(3)
        mix.step(3)
        _params_close(alone.param(0), mix.param(j), 0.02, 4)
    for j in range(4):
        assert torch.isfinite(mix.param(j)).all()


def test_graph_replay_matches_eager(cuda):
    announce("graph vs eager")
    specs = lambda: [vol_spec(6, w=12, h=12, spp=8), surf_spec(7, spp=4), vol_spec(8, n=5, spp=4)]
    eager = BatchOptimizer(specs(), use_graph=False)
    graph = BatchOptimizer(specs(), use_graph=True)
    eager.step(1)
    graph.step(1)
    graph.synchronize()
    eager.synchronize()
    for j in range(3):
        assert torch.equal(graph.jobs[j].img, eager.jobs[j].img)
        _grads_close(graph.jobs[j].grad, eager.jobs[j].grad)
    eager.step(3)
    graph.step(3)
    for j in range(3):
        _params_close(graph.param(j), eager.param(j), 0.02, 4)
        assert graph.job_state(j)[0] == 4


def test_autotune(cuda):
    announce("autotune")
    best, table = autotune(lambda: [vol_spec(9, spp=4), surf_spec(10, spp=4)], blocks=(64, 256),
                           grabs=(64, 512), steps=2)
    assert best in table and len(table) == 4
    assert all(v > 0 for v in table.values())
