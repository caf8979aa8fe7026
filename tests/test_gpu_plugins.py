"""The reference's plugin surface on the device: the tutorial's drop-in
``cv`` integrator (tutorial.ipynb:312-321, plugins/cv_integrator.py:47-94)
and the ``meta`` integrator's ``cv_pss`` and ``force_baseline`` branches
(plugins/metaintegrator.py:87-130), each checked bit for bit against the
NumPy combine oracle fed with the same images."""
import numpy as np
import pytest
import torch

import welford_ref as W
from conftest import announce
from helpers import volume_box_scene
import recursive_control_variates_amd as rcv
from recursive_control_variates_amd.optim import Adam, l1_loss

pytestmark = pytest.mark.gpu
KEY = "box.interior_medium.sigma_t.data"


def _scene(seed=0, n=6, w=12, h=10):
    rng = np.random.default_rng(seed)
    grid = (0.2 + rng.uniform(0, 0.8, (n, n, n, 1))).astype(np.float32)
    return rng, grid, rcv.load_dict(volume_box_scene(grid, albedo=0.8, scale=5.0, w=w, h=h, spp=8))


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _states(rng, grid, n):
    """A sequence of parameter states, as an optimisation produces them."""
    out = [grid]
    for _ in range(n - 1):
        out.append(np.clip(out[-1] * rng.uniform(0.9, 1.1, grid.shape), 0, None).astype(np.float32))
    return out


def test_tutorial_cv_dict_render_matches_combine_oracle(cuda):
    # {'type': 'cv', 'integrator': {'type': 'twostateprbvolpath', ...}}: each
    # render() draws (H, F) from the two-state integrator and blends; the
    # state must equal the NumPy combine fed with the same H, F
    announce("cv plugin render vs combine oracle")
    rng, grid, sc = _scene(1)
    cv = rcv.load_dict({"type": "cv", "integrator": {"type": "twostateprbvolpath", "max_depth": 8,
                                                     "rr_depth": 8}})
    inner = cv.integrator
    p = rcv.traverse(sc)
    ref = None
    for it, g in enumerate(_states(rng, grid, 7)):
        p[KEY] = torch.from_numpy(g).cuda()
        p.update()
        img = cv.render(sc, 0, it, 8)
        # the inner render is deterministic: the same call reproduces H, F
        H, F = inner.render_twostates(sc, 0, it, 8)
        assert torch.equal(H, cv.H) and torch.equal(F, cv.F)
        if ref is None:
            ref = W.CVStateRef(tuple(H.shape), warmup=cv.warmup, beta1=cv.beta1, beta2=cv.beta2)
        expect = ref.step(H.cpu().numpy(), F.cpu().numpy())
        assert _bits_equal(img.cpu().numpy(), expect), it
        got = cv.state.buf.cpu().numpy()
        for k, a in enumerate(ref.arrays()):
            assert _bits_equal(got[k], a), (it, k)
        assert cv.it == ref.it
    assert np.any(ref.w_s > 0)          # the weights did engage
    # reset() (utils.py:192) starts afresh: the next image is F
    cv.reset()
    img = cv.render(sc, 0, 99, 8)
    assert torch.equal(img, cv.F)


def test_tutorial_cv_loop_trains(cuda):
    # the tutorial's loop body, unchanged (tutorial.ipynb:363 / cell 24):
    # render with the cv integrator, L1 loss, backward (delegated to the
    # adjoint integrator), Adam, clamp
    announce("cv plugin tutorial loop")
    rng, grid, sc = _scene(2)
    target = rcv.load_dict({"type": "prbvolpath", "max_depth": 8, "rr_depth": 8}).render(sc, 0, 12345, 256)
    cv = rcv.load_dict({"type": "cv", "integrator": {"type": "twostateprbvolpath", "max_depth": 8,
                                                     "rr_depth": 8}})
    p = rcv.traverse(sc)
    opt = Adam(lr=0.05)
    opt[KEY] = torch.full(grid.shape, 0.3, device="cuda")
    p.update(opt)
    losses = []
    for it in range(12):
        p.update(opt)
        img = rcv.render(sc, p, integrator=cv, spp=4, spp_grad=8, seed=it)
        loss = l1_loss(img, target)
        loss.backward()
        assert torch.isfinite(opt[KEY].grad).all()
        opt.step()
        opt[KEY] = torch.clamp(opt[KEY].detach(), 1e-3, 10)
        losses.append(float(loss))
    assert np.mean(losses[-4:]) < np.mean(losses[:4])


def test_cv_render_backward_delegates_to_adjoint_integrator(cuda):
    # 'adjoint_integrator' (plugins/cv_integrator.py:22,87-94): the cv
    # integrator's adjoint is exactly the given integrator's
    announce("cv adjoint delegation")
    rng, grid, sc = _scene(3)
    cv = rcv.load_dict({"type": "cv", "integrator": {"type": "twostateprbvolpath", "max_depth": 8},
                        "adjoint_integrator": {"type": "volpathsimple", "max_depth": 8}})
    p = rcv.traverse(sc)
    p[KEY] = p[KEY].detach().clone().requires_grad_(True)
    p.update()
    p.update()
    gimg = torch.from_numpy(rng.uniform(0.5, 1.5, (10, 12, 3)).astype(np.float32)).cuda()
    a = cv.adjoint(sc, p, gimg, 0, 7, 8)[KEY]
    b = cv.adjoint_integrator.adjoint(sc, p, gimg, 0, 7, 8)[KEY]
    scale = float(b.abs().max())
    assert scale > 0
    torch.testing.assert_close(a, b, rtol=1e-4, atol=1e-5 * scale)
    cv.render_backward(sc, p, gimg, 0, 7, 8)
    assert p[KEY].grad is not None and torch.isfinite(p[KEY].grad).all()


def test_meta_cv_pss_matches_combine_oracle(cuda):
    # cv_pss (plugins/metaintegrator.py:87-91,117-119): F = render(seed) of the
    # current state, H = the previous iteration's render(seed + 1) of the
    # previous state (zero at the first iteration); combined as cv_ps
    announce("meta cv_pss vs combine oracle")
    rng, grid, sc = _scene(4)
    base = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 8, "rr_depth": 8})
    meta = rcv.load_dict({"type": "meta", "method": "cv_pss", "integrator": base})
    p = rcv.traverse(sc)
    ref = None
    H_next = None
    for it, g in enumerate(_states(rng, grid, 6)):
        p[KEY] = torch.from_numpy(g).cuda()
        p.update()
        img = meta.render(sc, 0, it, 8)
        F = base.render(sc, 0, it, 8)
        H = torch.zeros_like(F) if H_next is None else H_next
        assert torch.equal(meta.F, F)
        if ref is None:
            ref = W.CVStateRef(tuple(F.shape), warmup=meta.warmup, beta1=meta.beta1, beta2=meta.beta2)
        expect = ref.step(H.cpu().numpy(), F.cpu().numpy())
        assert _bits_equal(img.cpu().numpy(), expect), it
        H_next = base.render(sc, 0, it + 1, 8)
        assert torch.equal(meta.H, H_next)
    assert np.any(ref.w_s > 0)


@pytest.mark.parametrize("method", ["cv_ps", "cv_pss", "baseline"])
def test_meta_force_baseline_returns_the_noisy_render(cuda, method):
    # force_baseline (plugins/metaintegrator.py:124-128): the returned image
    # is the plain render of this seed while the CV state advances as without
    # the flag
    announce(f"meta force_baseline ({method})")
    rng, grid, sc = _scene(5)
    base = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 8, "rr_depth": 8})
    forced = rcv.load_dict({"type": "meta", "method": method, "force_baseline": True, "integrator": base})
    normal = rcv.load_dict({"type": "meta", "method": method, "integrator": base})
    p = rcv.traverse(sc)
    for it, g in enumerate(_states(rng, grid, 5)):
        p[KEY] = torch.from_numpy(g).cuda()
        p.update()
        out_f = forced.render(sc, 0, it, 8)
        out_n = normal.render(sc, 0, it, 8)
        plain = base.render(sc, 0, it, 8)
        assert torch.equal(out_f, plain), it
        if method == "baseline":
            assert torch.equal(out_n, plain)
        else:
            assert torch.equal(forced.img, normal.img), it
            assert torch.equal(out_n, normal.img)
