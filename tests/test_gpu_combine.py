"""GPU parity of the fused CV combine (rcv_cv_step) against the NumPy oracle:
bit-exact on random, adversarial and full-size streams; graph replay."""
import numpy as np
import pytest
import torch

import welford_ref as W
from conftest import announce

pytestmark = pytest.mark.gpu


def _run(shape, steps, warmup=1, pre=False, b1=0.9, b2=0.999, seed=0, inject=False):
    from recursive_control_variates_amd.welford import CVState
    rng = np.random.default_rng(seed)
    st = CVState(shape, "cuda", b1, b2)
    ref = W.CVStateRef(shape, warmup=warmup, beta1=b1, beta2=b2, pre_update=pre)
    base = rng.uniform(0.1, 2.0, shape).astype(np.float32)
    for it in range(steps):
        H = (base * rng.uniform(0.8, 1.2, shape)).astype(np.float32)
        F = (H + rng.normal(0, 0.05, shape)).astype(np.float32)
        if inject and it in (2, 5):
            flat = H.reshape(-1)
            flat[:3] = [np.nan, np.inf, -np.inf]
            F.reshape(-1)[3:5] = [np.nan, 0.0]
        st.step(torch.from_numpy(H).cuda(), torch.from_numpy(F).cuda(), warmup, pre)
        ref.step(H, F)
    torch.cuda.synchronize()
    return st, ref


def _assert_same(st, ref):
    got = st.buf.cpu().numpy()
    for k, a in enumerate(ref.arrays()):
        g = got[k]
        same = (g.view(np.uint32) == a.view(np.uint32)) | (np.isnan(g) & np.isnan(a))
        assert same.all(), f"array {k}: {np.count_nonzero(~same)} mismatches"
    c = st.counters.cpu().numpy()
    assert c[0] == ref.it and c[1] == ref.stats.n and c[2] == 0


@pytest.mark.parametrize("shape", [(8, 16, 3), (7, 5, 3), (1, 1, 1), (33, 17, 3)])
@pytest.mark.parametrize("pre", [False, True])
def test_combine_bitwise_random(shape, pre):
    announce("combine", shape, pre)
    st, ref = _run(shape, 9, pre=pre)
    _assert_same(st, ref)


@pytest.mark.parametrize("warmup", [0, 1, 4])
def test_combine_warmup(warmup):
    announce("combine warmup", warmup)
    st, ref = _run((16, 16, 3), 8, warmup=warmup, seed=3)
    _assert_same(st, ref)


def test_combine_nan_inf_guard():
    announce("combine nan/inf")
    st, ref = _run((4, 8, 3), 8, inject=True, seed=5)
    _assert_same(st, ref)
    w = st.w_s.cpu().numpy()
    assert np.all(np.isfinite(w)) and np.all((w >= 0) & (w <= 1))


def test_combine_full_size_volume_film():
    # (256, 4*256, 3): the 4-view 256^2 volume workload's CV state
    announce("combine full size")
    st, ref = _run((256, 1024, 3), 6, seed=7)
    _assert_same(st, ref)


def test_combine_betas():
    announce("combine betas")
    st, ref = _run((10, 10, 3), 7, b1=0.5, b2=0.75, seed=8)
    _assert_same(st, ref)


def test_combine_graph_replay():
    # counters live on the device, so a captured step replays correctly
    from recursive_control_variates_amd.welford import CVState
    announce("combine graph replay")
    shape = (32, 64, 3)
    rng = np.random.default_rng(11)
    st = CVState(shape, "cuda")
    ref = W.CVStateRef(shape)
    Hd = torch.zeros(shape, device="cuda")
    Fd = torch.zeros(shape, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        st.step(Hd, Fd, 1)        # warm-up launch outside the graph, it = 0
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    ref.step(np.zeros(shape, np.float32), np.zeros(shape, np.float32))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        st.step(Hd, Fd, 1)
    for _ in range(6):
        H = rng.uniform(0.5, 1.5, shape).astype(np.float32)
        F = (H + rng.normal(0, 0.1, shape)).astype(np.float32)
        Hd.copy_(torch.from_numpy(H))
        Fd.copy_(torch.from_numpy(F))
        g.replay()
        torch.cuda.synchronize()
        ref.step(H, F)
        _assert_same(st, ref)


def test_welford_single_variable():
    from recursive_control_variates_amd.welford import WelfordVarianceEstimator
    announce("welford single")
    rng = np.random.default_rng(2)
    est = WelfordVarianceEstimator()
    ref = W.WelfordVarianceEstimatorRef((5, 7))
    for _ in range(6):
        x = rng.normal(1, 0.5, (5, 7)).astype(np.float32)
        est.update(torch.from_numpy(x).cuda())
        ref.update(x)
    assert np.array_equal(est.var.cpu().numpy().view(np.uint32), ref.var.view(np.uint32))
    assert np.array_equal(est.mean.cpu().numpy().view(np.uint32), ref.mean.view(np.uint32))
    # get() is a host-side helper (torch division): equal to 1e-6 relative
    np.testing.assert_allclose(est.get().cpu().numpy(), ref.get(), rtol=1e-6)
