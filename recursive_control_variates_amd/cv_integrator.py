"""'cv' and 'meta' integrators: recursive control variates around an inner
integrator.

CVIntegrator (plugins/cv_integrator.py:9-94) wraps a two-state integrator
('path space' variant). MetaIntegrator (plugins/metaintegrator.py:7-147) is
what the optimisation driver uses: ``method`` in {'baseline', 'cv_ps',
'cv_pss'}, ``pre_update``, ``force_baseline``. Denoising (OptiX) is not part
of this build and raises.

State lives on the device (welford.CVState); ``it`` is mirrored on the host
because the reference keeps it as a Python int (cv_integrator.py:43,61,76)
and checkpoint code reads it.
"""
from __future__ import annotations

from typing import Optional

import torch

from .integrators import TwoStatePRBIntegrator, TwoStatePRBVolpathIntegrator
from .welford import CVState, StatisticsEstimator


class _CVBase:
    def _init_props(self, props):
        self.warmup = int(props.get("warmup", 1))
        self.beta1 = float(props.get("beta1", 0.9))
        self.beta2 = float(props.get("beta2", 0.999))

    def reset(self):
        """Drops the CV state; the next render starts afresh
        (plugins/cv_integrator.py:29-30, called by upsampling utils.py:192)."""
        self.init = False

    def init_buffers(self, shape=None, device=None):
        if self.init:
            return
        self._shape = shape
        self._device = device
        self.state: Optional[CVState] = None
        self.stats: Optional[StatisticsEstimator] = None
        self.it = 0
        self.H = None
        self.F = None
        self.init = True

    def _ensure_state(self, like: torch.Tensor):
        if self.state is None or tuple(self.state.shape) != tuple(like.shape):
            if self.state is not None:
                raise ValueError("image shape changed without reset()")
            self.state = CVState(like.shape, like.device, self.beta1, self.beta2)
            self.stats = StatisticsEstimator(self.state)

    @property
    def img(self):
        return None if not self.init or self.state is None else self.state.img

    @property
    def w_s(self):
        return None if not self.init or self.state is None else self.state.w_s

    @property
    def v_n(self):
        return None if not self.init or self.state is None else self.state.v_n

    def _combine(self, pre_update: bool) -> torch.Tensor:
        self._ensure_state(self.F)
        self.state.step(self.H, self.F, self.warmup, pre_update)
        self.it += 1
        return self.state.img.clone()

    # adjoint delegation (plugins/cv_integrator.py:79-94)
    def render_forward(self, scene, params, sensor=0, seed=0, spp=0, tangents=None):
        return self.adjoint_integrator.render_forward(scene, params, sensor, seed, spp, tangents)

    def render_backward(self, scene, params, grad_in, sensor=0, seed=0, spp=0):
        return self.adjoint_integrator.render_backward(scene, params, grad_in, sensor, seed, spp)

    def adjoint(self, scene, params, grad_in, sensor=0, seed=0, spp=0):
        return self.adjoint_integrator.adjoint(scene, params, grad_in, sensor, seed, spp)


class CVIntegrator(_CVBase):
    """'cv' (plugins/cv_integrator.py:9)."""

    def __init__(self, props):
        props = dict(props)
        self.integrator = props.get("integrator")
        if type(self.integrator) not in (TwoStatePRBIntegrator, TwoStatePRBVolpathIntegrator):
            raise ValueError("CV integrator expects a 'twostate' integrator as nested integrator !")
        self.adjoint_integrator = props.get("adjoint_integrator", self.integrator)
        self._init_props(props)
        self.reset()

    def render(self, scene, sensor: int = 0, seed: int = 0, spp: int = 0, develop: bool = True,
               evaluate: bool = True) -> torch.Tensor:
        assert isinstance(sensor, int)
        self.init_buffers()
        self.H, self.F = self.integrator.render_twostates(scene, sensor, seed, spp)
        return self._combine(pre_update=False)


class MetaIntegrator(_CVBase):
    """'meta' (plugins/metaintegrator.py:7)."""

    METHODS = ("baseline", "cv_ps", "cv_pss")

    def __init__(self, props):
        props = dict(props)
        self.denoise = bool(props.get("denoise", False))
        if self.denoise:
            raise NotImplementedError("denoising (OptiX) is not part of this build")
        self.force_baseline = bool(props.get("force_baseline", False))
        self.integrator = props.get("integrator")
        self.adjoint_integrator = props.get("adjoint_integrator", self.integrator)
        self.method = props.get("method", "baseline")
        if self.method not in self.METHODS:
            raise ValueError(f"unknown method {self.method!r}")
        self.pre_update = bool(props.get("pre_update", False))
        self._init_props(props)
        self.reset()

    def init_buffers(self, shape=None, device=None):
        if self.init:
            return
        super().init_buffers(shape, device)
        if "cv" not in self.method:
            self._baseline_img = None

    @property
    def img(self):
        if "cv" not in self.method:
            return getattr(self, "_baseline_img", None)
        return super().img

    def render(self, scene, sensor: int = 0, seed: int = 0, spp: int = 0, develop: bool = True,
               evaluate: bool = True) -> torch.Tensor:
        assert isinstance(sensor, int)
        self.init_buffers()
        img_noisy = None
        if self.method == "cv_ps":
            self.H, self.F = self.integrator.render_twostates(scene, sensor, seed, spp)
        else:
            img_noisy = self.integrator.render(scene, sensor, seed, spp, develop, evaluate)
        if self.method == "cv_pss":
            self.F = img_noisy
            if self.H is None:
                self.H = torch.zeros_like(self.F)  # TensorXf(0.0) broadcast (metaintegrator.py:52)
        if "cv" in self.method:
            img = self._combine(pre_update=self.pre_update)
        else:
            self._baseline_img = img_noisy
            img = img_noisy
        if self.method == "cv_pss":
            # re-render the current state with the next iteration's seed
            # (plugins/metaintegrator.py:117-119)
            self.H = self.integrator.render(scene, sensor, seed + 1, spp, develop, evaluate)
        if self.force_baseline:
            if self.method != "cv_ps":
                return img_noisy
            return self.integrator.render(scene, sensor, seed, spp, develop, evaluate)
        return img
