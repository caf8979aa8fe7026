"""Device-resident EMA-Welford statistics (plugins/welford.py).

``CVState`` holds the per-pixel control-variate state of CVIntegrator /
MetaIntegrator (plugins/cv_integrator.py:32-45): img, v_n, w_s and the
StatisticsEstimator moments mean_x, mean_y, var_x, var_y, cov, as one
(8, H, W, 3) float32 block, plus device int32 counters (it, n). One
``step`` is one fused kernel (librcv ``rcv_cv_step``).
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _native as N

_FIELDS = ("img", "v_n", "w_s", "mean_x", "mean_y", "var_x", "var_y", "cov")


def _bias(beta: float, n: int) -> float:
    """1 - beta**n as the kernels evaluate it (double, rounded to float)."""
    b = float(torch.tensor(beta, dtype=torch.float32))
    return float(torch.tensor(1.0 - math.exp(n * math.log(b)), dtype=torch.float32)) if n > 0 else 0.0


class CVState:
    def __init__(self, shape, device, beta1=0.9, beta2=0.999):
        self.shape = tuple(shape)
        self.device = torch.device(device)
        self.beta1 = float(beta1)
        self.beta2 = float(beta2)
        self.buf = torch.zeros((8,) + self.shape, dtype=torch.float32, device=self.device)
        self.counters = torch.zeros(4, dtype=torch.int32, device=self.device)
        self._c = N.CVState(*[C.c_void_p(self.buf[i].data_ptr()) for i in range(8)],
                            C.c_void_p(self.counters.data_ptr()))
        self.n_elem = self.buf[0].numel()

    def __getattr__(self, name):
        if name in _FIELDS:
            return self.buf[_FIELDS.index(name)]
        raise AttributeError(name)

    def step(self, H: torch.Tensor, F: torch.Tensor, warmup: int, pre_update: bool = False):
        if tuple(H.shape) != self.shape or tuple(F.shape) != self.shape:
            raise ValueError(f"CV images {tuple(H.shape)}/{tuple(F.shape)} do not match state {self.shape}")
        H = H.contiguous()
        F = F.contiguous()
        N.call("rcv_cv_step", C.byref(self._c), C.c_void_p(H.data_ptr()), C.c_void_p(F.data_ptr()),
               self.n_elem, int(warmup), C.c_float(self.beta1), C.c_float(self.beta2), int(bool(pre_update)),
               C.c_void_p(N.stream_handle(self.device)))


class StatisticsEstimator:
    """View of the moments inside a CVState (plugins/welford.py:38-78)."""

    def __init__(self, state: CVState):
        self._s = state
        self.beta1 = state.beta1
        self.beta2 = state.beta2

    @property
    def n(self) -> torch.Tensor:
        return self._s.counters[1:2]

    mean_x = property(lambda self: self._s.mean_x)
    mean_y = property(lambda self: self._s.mean_y)
    var_x = property(lambda self: self._s.var_x)
    var_y = property(lambda self: self._s.var_y)
    cov = property(lambda self: self._s.cov)

    def get(self):
        """(var_x, var_y, cov) / (1 - beta2^n) (plugins/welford.py:77-78)."""
        b = torch.tensor(_bias(self.beta2, int(self.n.item())), dtype=torch.float32, device=self.var_x.device)
        return self.var_x / b, self.var_y / b, self.cov / b


class WelfordVarianceEstimator:
    """Single-variable EMA variance (plugins/welford.py:4-36), device resident."""

    def __init__(self, beta1=0.9, beta2=0.999):
        self.beta1 = float(beta1)
        self.beta2 = float(beta2)
        self.mean = None
        self.var = None
        self.counter = None

    @property
    def n(self):
        return 0 if self.counter is None else self.counter[0:1]

    def update(self, x: torch.Tensor):
        x = x.detach().contiguous().float()
        if self.mean is None:
            self.mean = torch.zeros_like(x)
            self.var = torch.zeros_like(x)
            self.counter = torch.zeros(2, dtype=torch.int32, device=x.device)
        N.call("rcv_welford_update", C.c_void_p(self.mean.data_ptr()), C.c_void_p(self.var.data_ptr()),
               C.c_void_p(self.counter.data_ptr()), C.c_void_p(x.data_ptr()), x.numel(),
               C.c_float(self.beta1), C.c_float(self.beta2), C.c_void_p(N.stream_handle(x.device)))

    def get(self):
        n = int(self.counter[0].item()) if self.counter is not None else 0
        return self.var / torch.tensor(_bias(self.beta2, n), dtype=torch.float32, device=self.var.device)
