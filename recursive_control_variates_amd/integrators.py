"""Integrator plugins backed by librcv.

* ``prb`` / ``prbvolpath``: plain single-state renders of the NEW state (the
  upstream PRBIntegrator / PRBVolpathIntegrator the reference subclasses and
  the ``baseline`` method renders with, plugins/metaintegrator.py:87-88).
* ``twostateprb`` (plugins/twostatepath.py:97-364) and ``twostateprbvolpath``
  (plugins/twostatevolpath.py:11-482): add ``check_scene`` and
  ``render_twostates`` -> (img_old, img_new); their ``render_backward`` is
  the inherited adjoint on the NEW state.

Props and defaults follow the reference plugin docstrings: ``max_depth`` 6
(-1 = unbounded), ``rr_depth`` 5, ``hide_emitters`` False, ``xyz`` False
(plugins/twostatepath.py:106-118,149). ``samples_per_thread`` and
``block_size`` are this build's launch-shape knobs (0 = automatic); they
change results only through the float summation order of a pixel.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict

import torch

from . import _native as N

# linear sRGB -> CIE XYZ (D65), used when xyz=True (plugins/twostatepath.py:361-364)
_SRGB_TO_XYZ = torch.tensor([[0.412453, 0.357580, 0.180423],
                             [0.212671, 0.715160, 0.072169],
                             [0.019334, 0.119193, 0.950227]], dtype=torch.float32)


def _props(props) -> dict:
    return dict(props or {})


class _LaunchEvents:
    """Records a (start, end) pair of timing events on the launch stream
    around each native launch of one kind, for benchmarks that time kernels
    inside a running optimisation."""

    def __init__(self):
        self.pairs = []

    def around(self, fn):
        import torch
        s = torch.cuda.current_stream()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn()
        b.record(s)
        self.pairs.append((a, b))

    def elapsed_ms(self):
        return [a.elapsed_time(b) for a, b in self.pairs]


def _launch(events, fn):
    if events is None:
        fn()
    else:
        events.around(fn)


class RenderIntegrator:
    """Common native render / adjoint surface."""

    INTEGRATOR_TYPE = N.INTEGRATOR_PATH
    # Optional timing hooks: a _LaunchEvents per launch kind (None: off).
    adjoint_events = None
    twostate_events = None

    def time_launches(self, on: bool = True):
        """Start (or stop) recording timing events around every adjoint and
        two-state launch of this integrator; returns (adjoint, two-state)
        recorders."""
        self.adjoint_events = _LaunchEvents() if on else None
        self.twostate_events = _LaunchEvents() if on else None
        return self.adjoint_events, self.twostate_events

    def __init__(self, props=None):
        p = _props(props)
        self.max_depth = int(p.get("max_depth", 6))
        self.rr_depth = int(p.get("rr_depth", 5))
        self.hide_emitters = bool(p.get("hide_emitters", False))
        self.samples_per_thread = int(p.get("samples_per_thread", 0))
        self.block_size = int(p.get("block_size", 0))
        self.xyz = bool(p.get("xyz", False))
        if self.max_depth < -1 or self.max_depth == 0:
            raise ValueError("max_depth must be -1 or positive")

    def desc(self) -> N.IntegratorDesc:
        d = N.IntegratorDesc()
        d.type = self.INTEGRATOR_TYPE
        d.max_depth = self.max_depth
        d.rr_depth = self.rr_depth
        d.hide_emitters = int(self.hide_emitters)
        d.samples_per_thread = self.samples_per_thread
        d.block_size = self.block_size
        return d

    @staticmethod
    def _spp(scene, sensor, spp):
        return int(spp) if spp else scene.sensors()[sensor].spp

    def _post(self, img):
        if not self.xyz:
            return img
        return img @ _SRGB_TO_XYZ.to(img.device).T

    def _pre_grad(self, grad):
        if not self.xyz:
            return grad
        return grad @ _SRGB_TO_XYZ.to(grad.device)

    def render(self, scene, sensor: int = 0, seed: int = 0, spp: int = 0, develop: bool = True,
               evaluate: bool = True) -> torch.Tensor:
        if not isinstance(sensor, int):
            raise TypeError("sensor must be an int index")
        scene.check_params()
        spp = self._spp(scene, sensor, spp)
        img = torch.empty(scene.film_shape(sensor), dtype=torch.float32, device=scene.device)
        d = self.desc()
        N.call("rcv_render", scene.handle, C.byref(d), sensor, C.c_uint32(seed & 0xFFFFFFFF), spp,
               C.c_void_p(img.data_ptr()), C.c_void_p(N.stream_handle(scene.device)))
        return self._post(img)

    def adjoint(self, scene, params, grad_in: torch.Tensor, sensor: int = 0, seed: int = 0,
                spp: int = 0) -> Dict[str, torch.Tensor]:
        """Path-replay adjoint: {key: dLoss/dparam} for the differentiable
        parameters of ``params`` (NEW state)."""
        scene.check_params()
        spp = self._spp(scene, sensor, spp)
        grad_in = self._pre_grad(grad_in.detach()).to(scene.device, torch.float32).contiguous()
        if tuple(grad_in.shape) != tuple(scene.film_shape(sensor)):
            raise ValueError(f"grad_in shape {tuple(grad_in.shape)} != film {scene.film_shape(sensor)}")
        keys = [k for k, _ in params.differentiable()] if params is not None else []
        grads = {k: torch.zeros(scene.params[k].shape, dtype=torch.float32, device=scene.device) for k in keys}
        ids = N.i32_array([scene.params[k].index for k in keys] or [0])
        ptrs = (C.c_void_p * max(1, len(keys)))(*[grads[k].data_ptr() for k in keys])
        d = self.desc()
        _launch(self.adjoint_events, lambda: N.call(
            "rcv_render_backward", scene.handle, C.byref(d), sensor, C.c_uint32(seed & 0xFFFFFFFF), spp,
            C.c_void_p(grad_in.data_ptr()), len(keys), ids, ptrs, C.c_void_p(N.stream_handle(scene.device))))
        return grads

    def render_backward(self, scene, params, grad_in, sensor: int = 0, seed: int = 0, spp: int = 0) -> None:
        """Accumulates dLoss/dparam into ``.grad`` of the differentiable
        parameters (side effect, like the reference)."""
        grads = self.adjoint(scene, params, grad_in, sensor, seed, spp)
        for k, src in params.differentiable():
            g = grads[k].to(src.device)
            src.grad = g if src.grad is None else src.grad + g

    def render_forward(self, scene, params, sensor: int = 0, seed: int = 0, spp: int = 0, tangents=None):
        """Forward-mode derivative of the NEW-state image: d(image)/d(theta) .
        tangent. The tangent of each differentiable parameter of ``params``
        is its ``.grad`` (where Dr.Jit keeps forward-mode gradients as well),
        or ``tangents[key]`` when given; parameters without one are held
        fixed. Paths are replayed with the primal's random numbers, so for one
        seed the result is the transpose of ``adjoint`` on the same samples."""
        scene.check_params()
        spp = self._spp(scene, sensor, spp)
        if tangents is None:
            tangents = {k: t.grad for k, t in params.differentiable() if t.grad is not None}
        keys, ptrs, keep = [], [], []
        for k, t in tangents.items():
            if k not in scene.params:
                raise KeyError(f"unknown parameter {k!r}")
            t = t.detach().to(scene.device, torch.float32).contiguous()
            if tuple(t.shape) != tuple(scene.params[k].shape):
                raise ValueError(f"{k}: tangent shape {tuple(t.shape)} != {tuple(scene.params[k].shape)}")
            keys.append(k)
            keep.append(t)
            ptrs.append(t.data_ptr())
        dimg = torch.empty(scene.film_shape(sensor), dtype=torch.float32, device=scene.device)
        ids = N.i32_array([scene.params[k].index for k in keys] or [0])
        arr = (C.c_void_p * max(1, len(keys)))(*ptrs)
        d = self.desc()
        N.call("rcv_render_forward", scene.handle, C.byref(d), sensor, C.c_uint32(seed & 0xFFFFFFFF), spp,
               len(keys), ids, arr, C.c_void_p(dimg.data_ptr()), C.c_void_p(N.stream_handle(scene.device)))
        del keep
        return self._post(dimg)


class PRBIntegrator(RenderIntegrator):
    INTEGRATOR_TYPE = N.INTEGRATOR_PATH


class PRBVolpathIntegrator(RenderIntegrator):
    INTEGRATOR_TYPE = N.INTEGRATOR_VOLPATH


class VolpathSimpleIntegrator(RenderIntegrator):
    """'volpathsimple' (plugins/volpathsimple.py:23-48): one medium in a
    null-BSDF boundary lit by an environment; its adjoint uses differential
    ratio tracking. Props ``use_nee``, ``use_drt``, ``use_drt_subsampling``,
    ``use_drt_mis`` (all True by default), ``max_depth``, ``rr_depth``,
    ``hide_emitters``. The reference uses it as the ``adjoint_integrator`` of
    volume optimisations (experiments.py:131)."""

    INTEGRATOR_TYPE = N.INTEGRATOR_VOLPATH_SIMPLE

    def __init__(self, props=None):
        super().__init__(props)
        p = _props(props)
        self.use_nee = bool(p.get("use_nee", True))
        self.use_drt = bool(p.get("use_drt", True))
        self.use_drt_subsampling = bool(p.get("use_drt_subsampling", True))
        self.use_drt_mis = bool(p.get("use_drt_mis", True))

    def desc(self) -> N.IntegratorDesc:
        d = super().desc()
        d.flags = ((0 if self.use_nee else N.SIMPLE_NO_NEE) | (0 if self.use_drt else N.SIMPLE_NO_DRT)
                   | (0 if self.use_drt_subsampling else N.SIMPLE_NO_DRT_SUBSAMPLING)
                   | (0 if self.use_drt_mis else N.SIMPLE_NO_DRT_MIS))
        return d

    def render_forward(self, scene, params, sensor: int = 0, seed: int = 0, spp: int = 0, tangents=None):
        raise NotImplementedError("volpathsimple has no forward mode: its adjoint is written with backward-mode "
                                  "calls only (plugins/volpathsimple.py:183,502,593,620)")

    def __repr__(self):
        return f"VolpathSimpleIntegrator[max_depth = {self.max_depth}, use_drt = {self.use_drt}]"


class _TwoStateMixin:
    def _check_scene(self, scene):
        if getattr(self, "_checked_scene", None) is scene:
            return
        bs, ms = scene.twostate_flags()
        N.call("rcv_scene_check_twostate", scene.handle, self.INTEGRATOR_TYPE, N.i32_array(bs or [0]), len(bs),
               N.i32_array(ms or [0]), len(ms))
        self._checked_scene = scene

    def check_scene(self, scene):
        self._check_scene(scene)

    def render_twostates(self, scene, sensor: int = 0, seed: int = 0, spp: int = 1):
        """(img_old, img_new) = (H, F) from one set of paths carrying both
        states' throughputs (plugins/twostatepath.py:179-223)."""
        self._check_scene(scene)
        if not isinstance(sensor, int):
            raise TypeError("sensor must be an int index")
        scene.check_params()
        spp = self._spp(scene, sensor, spp)
        shape = scene.film_shape(sensor)
        img_old = torch.empty(shape, dtype=torch.float32, device=scene.device)
        img_new = torch.empty(shape, dtype=torch.float32, device=scene.device)
        d = self.desc()
        _launch(self.twostate_events, lambda: N.call(
            "rcv_render_twostates", scene.handle, C.byref(d), sensor, C.c_uint32(seed & 0xFFFFFFFF), spp,
            C.c_void_p(img_old.data_ptr()), C.c_void_p(img_new.data_ptr()),
            C.c_void_p(N.stream_handle(scene.device))))
        return self._post(img_old), self._post(img_new)


class TwoStatePRBIntegrator(_TwoStateMixin, PRBIntegrator):
    """'twostateprb' (plugins/twostatepath.py:97)."""


class TwoStatePRBVolpathIntegrator(_TwoStateMixin, PRBVolpathIntegrator):
    """'twostateprbvolpath' (plugins/twostatevolpath.py:11)."""

    def __init__(self, props=None):
        super().__init__(props)

    def __repr__(self):
        return f"TwoStatePRBVolpathIntegrator[max_depth = {self.max_depth}]"
