"""CPU oracle for the CV / EMA-Welford combine step.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this file;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker.

It restates, in NumPy float32 with the reference's operation order (no
fused multiply-add, one rounding per operation):

* ``StatisticsEstimator.update/get``     plugins/welford.py:38-78
* ``WelfordVarianceEstimator.update/get`` plugins/welford.py:4-36
* ``CVIntegrator.render`` combine        plugins/cv_integrator.py:57-77
* ``MetaIntegrator.render`` cv branch    plugins/metaintegrator.py:93-113
  (``pre_update`` ordering at :95-96 versus :110-111)

Bias corrections ``1 - beta**n`` are evaluated as ``1 - exp(n*log(beta))`` in
float64 and rounded once to float32. The reference evaluates ``beta**n`` with
Dr.Jit's float32 ``pow`` whose last-ulp behaviour cannot be reproduced here;
the double evaluation is the build's definition and the HIP kernel uses the
identical expression, so kernel and oracle agree bit for bit.

Pins (tests/test_oracle_combine.py): closed-form EMA mean, a hand-derived
two-sample variance, shift invariance, cov(x, x) == var(x), bilinearity of
the covariance, the analytic j/(j+1) weight sequence of the recursion under a
constant variance estimate, the NaN/Inf guard and the [0, 1] clamp.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def bias_correction(beta: float, n: int) -> np.float32:
    """``1 - beta**n`` evaluated in float64, rounded once to float32."""
    return f32(1.0 - math.exp(n * math.log(float(f32(beta))))) if n > 0 else f32(0.0)


class StatisticsEstimatorRef:
    """plugins/welford.py:38-78 -- EMA mean/var/cov of two image streams."""

    def __init__(self, shape, beta1=0.9, beta2=0.999):
        self.n = 0
        self.mean_x = np.zeros(shape, f32)
        self.mean_y = np.zeros(shape, f32)
        self.var_x = np.zeros(shape, f32)
        self.var_y = np.zeros(shape, f32)
        self.cov = np.zeros(shape, f32)
        self.beta1 = f32(beta1)
        self.beta2 = f32(beta2)

    def update(self, x, y):
        """plugins/welford.py:59-75, same order of operations."""
        x = np.asarray(x, f32)
        y = np.asarray(y, f32)
        b1, b2 = self.beta1, self.beta2
        one = f32(1.0)
        with np.errstate(all="ignore"):
            a = bias_correction(b1, self.n)
            if self.n > 0:                                   # welford.py:61-62
                dx1 = x - self.mean_x / a
                dy1 = y - self.mean_y / a
            else:
                dx1 = x.copy()
                dy1 = y.copy()
            self.mean_y = b1 * self.mean_y + (one - b1) * y  # welford.py:64
            self.mean_x = b1 * self.mean_x + (one - b1) * x  # welford.py:65
            self.n += 1                                      # welford.py:67
            a2 = bias_correction(b1, self.n)
            dx2 = x - self.mean_x / a2                       # welford.py:68
            dy2 = y - self.mean_y / a2                       # welford.py:69
            c = one - b2
            self.var_x = b2 * self.var_x + (c * dx1) * dx2   # welford.py:71
            self.var_y = b2 * self.var_y + (c * dy1) * dy2   # welford.py:72
            self.cov = b2 * self.cov + (c * dx1) * dy2       # welford.py:73
        for name in ("mean_x", "mean_y", "var_x", "var_y", "cov"):
            setattr(self, name, np.asarray(getattr(self, name), f32))

    def get(self):
        """plugins/welford.py:77-78."""
        b = bias_correction(self.beta2, self.n)
        with np.errstate(all="ignore"):
            return self.var_x / b, self.var_y / b, self.cov / b


class WelfordVarianceEstimatorRef:
    """plugins/welford.py:4-36 -- single-variable EMA variance."""

    def __init__(self, shape, beta1=0.9, beta2=0.999):
        self.n = 0
        self.mean = np.zeros(shape, f32)
        self.var = np.zeros(shape, f32)
        self.beta1 = f32(beta1)
        self.beta2 = f32(beta2)

    def update(self, x):
        """plugins/welford.py:27-33."""
        x = np.asarray(x, f32)
        b1, b2 = self.beta1, self.beta2
        one = f32(1.0)
        with np.errstate(all="ignore"):
            a = bias_correction(b1, self.n)
            dx1 = x - self.mean / a if self.n > 0 else x.copy()
            self.mean = np.asarray(b1 * self.mean + (one - b1) * x, f32)
            self.n += 1
            dx2 = x - self.mean / bias_correction(b1, self.n)
            self.var = np.asarray(b2 * self.var + ((one - b2) * dx1) * dx2, f32)

    def get(self):
        """plugins/welford.py:35-36."""
        with np.errstate(all="ignore"):
            return self.var / bias_correction(self.beta2, self.n)


def cv_weights(w_s, v_n, v0, v1, cov):
    """Weight recursion, plugins/cv_integrator.py:63-68.

    Returns the new ``(w_s, v_n)``.
    """
    w_s = np.asarray(w_s, f32)
    two = f32(2.0)
    with np.errstate(all="ignore"):
        v_n = (w_s * w_s) * (v_n + v0) + v1 - (two * w_s) * cov   # :63
        w = cov / (v0 + v_n)                                        # :65
    w = np.where(np.isnan(w) | np.isinf(w), f32(0.0), w)           # :67
    w = np.minimum(np.maximum(w, f32(0.0)), f32(1.0))               # :68
    return np.asarray(w, f32), np.asarray(v_n, f32)


def blend(w_s, img, H, F):
    """Recursive blend, plugins/cv_integrator.py:70."""
    with np.errstate(all="ignore"):
        return np.asarray(w_s * (img - H) + F, f32)


class CVStateRef:
    """Per-pixel CV state of ``CVIntegrator`` / ``MetaIntegrator`` (cv_ps).

    ``init_buffers`` semantics (plugins/cv_integrator.py:32-45): every array
    starts at zero and ``it`` at 0.
    """

    def __init__(self, shape, warmup=1, beta1=0.9, beta2=0.999, pre_update=False):
        self.shape = tuple(shape)
        self.img = np.zeros(shape, f32)
        self.v_n = np.zeros(shape, f32)
        self.w_s = np.zeros(shape, f32)
        self.H = np.zeros(shape, f32)
        self.F = np.zeros(shape, f32)
        self.stats = StatisticsEstimatorRef(shape, beta1, beta2)
        self.it = 0
        self.warmup = warmup
        self.pre_update = pre_update

    def step(self, H, F):
        """One combine: plugins/metaintegrator.py:93-113 (cv_integrator.py:60-76
        is the ``pre_update=False`` case)."""
        self.H = np.asarray(H, f32)
        self.F = np.asarray(F, f32)
        if self.it > 0 and self.pre_update:                 # metaintegrator.py:95-96
            self.stats.update(self.H, self.F)
        if self.it > self.warmup:                           # cv_integrator.py:62
            v0, v1, cov = self.stats.get()
            self.w_s, self.v_n = cv_weights(self.w_s, self.v_n, v0, v1, cov)
        self.img = blend(self.w_s, self.img, self.H, self.F)
        if self.it > 0 and not self.pre_update:             # cv_integrator.py:74-75
            self.stats.update(self.H, self.F)
        self.it += 1
        return self.img

    def arrays(self):
        """The eight device arrays in the order the C-ABI state uses."""
        s = self.stats
        return [self.img, self.v_n, self.w_s, s.mean_x, s.mean_y, s.var_x, s.var_y, s.cov]
