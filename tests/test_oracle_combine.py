"""Pins of the combine oracle (oracle/welford_ref.py) by closed forms and
invariants, so that a dropped term, a wrong index or a swapped operand in
the restatement of plugins/welford.py / plugins/cv_integrator.py fails."""
import numpy as np
import pytest

import welford_ref as W

f32 = np.float32


def run_stats(xs, ys, b1=0.9, b2=0.999):
    st = W.StatisticsEstimatorRef(np.shape(xs[0]), b1, b2)
    for x, y in zip(xs, ys):
        st.update(np.asarray(x, f32), np.asarray(y, f32))
    return st


@pytest.mark.parametrize("b1", [0.5, 0.9, 0.99])
def test_ema_mean_closed_form(b1):
    # bias-corrected EMA mean == sum_i b1^(n-i) (1-b1) x_i / (1 - b1^n)
    rng = np.random.default_rng(1)
    xs = rng.normal(2.0, 1.0, size=(25, 7)).astype(f32)
    st = run_stats(xs, xs, b1=b1)
    n = len(xs)
    w = np.array([b1 ** (n - 1 - i) * (1 - b1) for i in range(n)])
    expect = (w[:, None] * xs.astype(np.float64)).sum(0) / (1 - b1 ** n)
    got = st.mean_x.astype(np.float64) / (1 - b1 ** n)
    np.testing.assert_allclose(got, expect, rtol=2e-5)
    assert st.n == n


@pytest.mark.parametrize("b1,b2", [(0.9, 0.999), (0.5, 0.5), (0.3, 0.8)])
def test_two_sample_moments_closed_form(b1, b2):
    # hand derivation: after updates x1, x2 the corrected variance is
    # (x2-x1)^2 b1 / ((1+b1)(1+b2)); covariance (x2-x1)(y2-y1) b1/((1+b1)(1+b2))
    x = [np.array([1.0, -2.0, 0.25], f32), np.array([3.0, 5.0, 0.75], f32)]
    y = [np.array([0.5, 1.0, 2.0], f32), np.array([-1.5, 4.0, 2.5], f32)]
    st = run_stats(x, y, b1, b2)
    vx, vy, cv = st.get()
    k = b1 / ((1 + b1) * (1 + b2))
    dx = (x[1] - x[0]).astype(np.float64)
    dy = (y[1] - y[0]).astype(np.float64)
    np.testing.assert_allclose(vx, dx * dx * k, rtol=1e-5)
    np.testing.assert_allclose(vy, dy * dy * k, rtol=1e-5)
    np.testing.assert_allclose(cv, dx * dy * k, rtol=1e-5)


def test_first_update_has_zero_variance():
    st = run_stats([np.array([4.0, -1.0], f32)], [np.array([2.0, 3.0], f32)])
    assert np.all(st.var_x == 0) and np.all(st.var_y == 0) and np.all(st.cov == 0)
    np.testing.assert_allclose(st.mean_x / W.bias_correction(0.9, 1), [4.0, -1.0], rtol=1e-6)


def test_shift_invariance_and_symmetry():
    rng = np.random.default_rng(3)
    xs = rng.normal(0, 1, size=(30, 64)).astype(f32)
    ys = (xs * 0.5 + rng.normal(0, 0.2, size=xs.shape)).astype(f32)
    a = run_stats(xs, ys)
    b = run_stats(xs + f32(8.0), ys - f32(2.0))
    for u, v in zip(a.get(), b.get()):
        np.testing.assert_allclose(u, v, rtol=2e-3, atol=1e-4)
    # cov(x, x) is the variance bit for bit; swapping roles swaps var_x / var_y
    c = run_stats(xs, xs)
    assert np.array_equal(c.cov, c.var_x)
    d = run_stats(ys, xs)
    assert np.array_equal(d.var_x, a.var_y) and np.array_equal(d.var_y, a.var_x)


def test_covariance_is_bilinear():
    rng = np.random.default_rng(4)
    xs = rng.normal(0, 1, size=(20, 16)).astype(f32)
    y1 = rng.normal(0, 1, size=xs.shape).astype(f32)
    y2 = rng.normal(0, 1, size=xs.shape).astype(f32)
    c12 = run_stats(xs, y1 + y2).get()[2]
    c1 = run_stats(xs, y1).get()[2]
    c2 = run_stats(xs, y2).get()[2]
    np.testing.assert_allclose(c12, c1 + c2, rtol=1e-3, atol=1e-5)


def test_weight_recursion_reproduces_running_mean_weights():
    # With a constant variance estimate s (v0 = v1 = cov = s, i.e. H == F),
    # the recursion gives w = 1/2, 2/3, 3/4, ...: the running mean of the
    # analytic weighting j/(j+1) (figures/steady_state/generate_data.py:58-59).
    s = np.full(5, 0.37, f32)
    w = np.zeros(5, f32)
    vn = np.zeros(5, f32)
    for j in range(1, 40):
        w, vn = W.cv_weights(w, vn, s, s, s)
        np.testing.assert_allclose(w, j / (j + 1), rtol=2e-5)


def test_weight_guards_and_clamp():
    w0 = np.zeros(4, f32)
    vn0 = np.zeros(4, f32)
    v0 = np.array([0.0, 1.0, 1.0, 1.0], f32)
    v1 = np.array([0.0, 1.0, 1.0, 1.0], f32)
    cov = np.array([0.0, -0.5, 5.0, 0.5], f32)
    w, vn = W.cv_weights(w0, vn0, v0, v1, cov)
    # 0/0 -> NaN -> 0 ; negative -> 0 ; > 1 -> 1 ; 0.5/(1+1) = 0.25
    np.testing.assert_array_equal(w, np.array([0.0, 0.0, 1.0, 0.25], f32))
    # v0 + v_n == 0 with cov != 0 -> inf -> 0   (v_n = 0*(..) + v1 - 0 = -1)
    w, vn = W.cv_weights(np.zeros(1, f32), np.zeros(1, f32), np.array([1.0], f32),
                         np.array([-1.0], f32), np.array([1.0], f32))
    assert vn[0] == -1.0 and w[0] == 0.0


def test_blend():
    img = np.array([1.0, 2.0], f32)
    H = np.array([0.5, 0.25], f32)
    F = np.array([3.0, -1.0], f32)
    np.testing.assert_array_equal(W.blend(np.zeros(2, f32), img, H, F), F)
    np.testing.assert_array_equal(W.blend(np.ones(2, f32), img, H, F), img - H + F)


def test_cv_state_sequence_constant_stream():
    # constant H == F == c with beta1 = 0.5 and powers of two: every bias-
    # corrected mean is exactly c, the variances are exactly 0, the weights
    # 0/0 -> 0 and the image is F at every step
    st = W.CVStateRef((3,), warmup=1, beta1=0.5, beta2=0.5)
    c = np.array([0.25, 0.5, 2.0], f32)
    for _ in range(6):
        img = st.step(c, c)
        np.testing.assert_array_equal(img, c)
        assert np.all(st.w_s == 0) and np.all(st.stats.var_x == 0)
    assert st.it == 6 and st.stats.n == 5
    # general constants: rounding noise may make the variance non-zero, but
    # img = w (img - H) + F stays exactly F because img == H == F
    st = W.CVStateRef((3,), warmup=1)
    c = np.array([0.2, 0.4, 0.8], f32)
    for _ in range(8):
        np.testing.assert_array_equal(st.step(c, c), c)
        assert np.all((st.w_s >= 0) & (st.w_s <= 1))


def test_cv_state_warmup_and_order():
    rng = np.random.default_rng(9)
    H = [rng.normal(1, 0.2, 8).astype(f32) for _ in range(5)]
    F = [(h + rng.normal(0, 0.05, 8)).astype(f32) for h in H]
    st = W.CVStateRef((8,), warmup=1)
    out0 = st.step(H[0], F[0])
    np.testing.assert_array_equal(out0, F[0])           # it = 0: img = F, no update
    assert st.stats.n == 0
    out1 = st.step(H[1], F[1])
    np.testing.assert_array_equal(out1, F[1])           # it = 1 <= warmup: w = 0
    assert st.stats.n == 1
    st.step(H[2], F[2])                                 # it = 2: weights from n = 1 stats
    assert st.stats.n == 2
    # with n = 1 the corrected variances are zero -> weights 0/0 -> 0
    assert np.all(st.w_s == 0)
    pre = W.CVStateRef((8,), warmup=1, pre_update=True)
    for h, f in zip(H[:3], F[:3]):
        pre.step(h, f)
    # pre-update: weights at it = 2 use n = 2 statistics (non-zero variance)
    assert pre.stats.n == 2
    assert np.any(pre.w_s != 0)


def run_var(xs, b1=0.9, b2=0.999):
    est = W.WelfordVarianceEstimatorRef(np.shape(xs[0]), b1, b2)
    for x in xs:
        est.update(np.asarray(x, f32))
    return est


@pytest.mark.parametrize("b1,b2", [(0.9, 0.999), (0.5, 0.5), (0.3, 0.8)])
def test_single_variable_two_sample_closed_form(b1, b2):
    # WelfordVarianceEstimator (plugins/welford.py:4-36) after x1, x2:
    # dx1 = x2 - x1, dx2 = b1 (x2 - x1) / (1 + b1), so the corrected variance
    # is (x2 - x1)^2 b1 / ((1 + b1)(1 + b2)); the first update adds nothing
    x = [np.array([1.0, -2.0, 0.25, 7.0], f32), np.array([3.0, 5.0, 0.75, 7.0], f32)]
    est = run_var(x[:1], b1, b2)
    # zero up to the rounding of mean / (1 - b1) against x
    assert np.all(np.abs(est.var) <= 1e-6 * x[0] * x[0]) and est.n == 1
    est = run_var(x, b1, b2)
    d = (x[1] - x[0]).astype(np.float64)
    # (the constant column is zero up to rounding of the bias-corrected mean)
    np.testing.assert_allclose(est.get(), d * d * b1 / ((1 + b1) * (1 + b2)), rtol=1e-5, atol=5e-6)
    # bias-corrected mean: (b1 x1 + x2) / (1 + b1)
    np.testing.assert_allclose(est.mean / W.bias_correction(b1, 2), (b1 * x[0] + x[1]) / (1 + b1), rtol=1e-6)
    assert est.n == 2


@pytest.mark.parametrize("b1,b2", [(0.9, 0.999), (0.5, 0.5)])
def test_single_variable_estimator_equals_the_pair_estimator(b1, b2):
    # the single-variable estimator is the x-half of StatisticsEstimator
    # (welford.py:27-33 vs :59-73): on one stream both give the same mean and
    # variance bit for bit, and get() the same corrected variance
    rng = np.random.default_rng(11)
    xs = rng.normal(0.3, 2.0, size=(23, 40)).astype(f32)
    ys = rng.normal(-1.0, 0.5, size=xs.shape).astype(f32)
    one = run_var(xs, b1, b2)
    pair = run_stats(xs, ys, b1, b2)
    assert one.n == pair.n == len(xs)
    assert np.array_equal(one.mean.view(np.uint32), pair.mean_x.view(np.uint32))
    assert np.array_equal(one.var.view(np.uint32), pair.var_x.view(np.uint32))
    assert np.array_equal(one.get().view(np.uint32), pair.get()[0].view(np.uint32))


def test_single_variable_constant_and_shift():
    # constant input: zero variance at every step; a shifted stream has the
    # same variance (to rounding) as the original
    c = np.array([0.25, 0.5, 2.0], f32)
    est = W.WelfordVarianceEstimatorRef((3,), 0.5, 0.5)
    for _ in range(7):
        est.update(c)
        assert np.all(est.var == 0)
    rng = np.random.default_rng(12)
    xs = rng.normal(0, 1, size=(30, 32)).astype(f32)
    np.testing.assert_allclose(run_var(xs).get(), run_var(xs + f32(5.0)).get(), rtol=3e-3, atol=1e-4)
