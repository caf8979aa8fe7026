"""Many concurrent optimisations advanced by one device-resident iteration
(librcv ``rcv_batch_*``): the multi-job form of the reference driver's loop
body (optimize.py:104-129) for jobs of uneven image, grid and sample counts.

Each job owns a scene, one optimised parameter, a method ('cv_ps' two-state
+ CV combine, or 'baseline' plain render), its CV state and Adam state. One
``step()`` advances every job by one iteration with a fixed sequence of
persistent / elementwise kernels; with ``use_graph`` the sequence is captured
into a hipGraph on the first step and replayed afterwards. ``autotune``
picks the persistent kernels' block size and queue grab by timing a
throw-away copy of the batch.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional

import numpy as np
import torch

from . import _native as N
from .scene import Scene
from .scene_desc import describe


class BatchJob(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p), ("integrator", N.IntegratorDesc), ("sensor", C.c_int32), ("seed", C.c_uint32),
        ("spp", C.c_int32), ("spp_grad", C.c_int32), ("method", C.c_int32), ("loss", C.c_int32),
        ("param", C.c_int32), ("param_ring", C.c_void_p * 3), ("adam_m", C.c_void_p), ("adam_v", C.c_void_p),
        ("grad", C.c_void_p), ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
        ("clamp_lo", C.c_float), ("clamp_hi", C.c_float), ("img_ref", C.c_void_p), ("H", C.c_void_p),
        ("F", C.c_void_p), ("img", C.c_void_p), ("grad_img", C.c_void_p), ("cv_state", C.c_void_p * 8),
        ("warmup", C.c_int32), ("cv_beta1", C.c_float), ("cv_beta2", C.c_float), ("pre_update", C.c_int32),
        ("resume", C.c_int32), ("iteration", C.c_int32), ("ring_old", C.c_int32), ("ring_new", C.c_int32),
        ("ring_next", C.c_int32), ("next_key", C.c_uint32),
    ]


_BATCH_SIGS = {
    "rcv_batch_create": [C.POINTER(C.c_void_p)],
    "rcv_batch_destroy": [C.c_void_p],
    "rcv_batch_add_job": [C.c_void_p, C.POINTER(BatchJob), C.POINTER(C.c_int32)],
    "rcv_batch_commit": [C.c_void_p, C.c_void_p],
    "rcv_batch_set_launch": [C.c_void_p, C.c_int32, C.c_int32],
    "rcv_batch_step": [C.c_void_p, C.c_int32, C.c_void_p],
    "rcv_batch_job_state": [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_void_p],
    "rcv_batch_set_active": [C.c_void_p, C.POINTER(C.c_uint32), C.c_int32, C.c_void_p],
    "rcv_batch_read_job": [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                           C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.c_void_p],
}
N.SIGNATURES.update(_BATCH_SIGS)


def _lib():
    lib = N.lib()
    for name, args in _BATCH_SIGS.items():
        fn = getattr(lib, name)
        if fn.argtypes != args:
            fn.argtypes = args
            fn.restype = C.c_int
    return lib


@dataclass
class JobSpec:
    """One optimisation: scene (dict or Scene), optimised key, its initial
    value, the reference image and the run_opt-style hyper-parameters."""
    scene: Any
    key: str
    init: Any
    img_ref: Any
    integrator: Dict[str, Any] = field(default_factory=lambda: {"type": "prbvolpath", "max_depth": 16})
    method: str = "cv_ps"
    loss: str = "L1"
    spp: int = 16
    spp_grad: int = 0
    seed: int = 0
    lr: float = 1e-2
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    clamp: tuple = (-float("inf"), float("inf"))
    warmup: int = 1
    cv_beta1: float = 0.9
    cv_beta2: float = 0.999
    pre_update: bool = False
    sensor: int = 0


_ITYPE = {"prb": N.INTEGRATOR_PATH, "path": N.INTEGRATOR_PATH, "twostateprb": N.INTEGRATOR_PATH,
          "prbvolpath": N.INTEGRATOR_VOLPATH, "volpath": N.INTEGRATOR_VOLPATH,
          "twostateprbvolpath": N.INTEGRATOR_VOLPATH}


class _JobBuffers:
    def __init__(self, spec: JobSpec, device):
        self.spec = spec
        sc = spec.scene
        if isinstance(sc, dict):
            sc = Scene(describe(sc), device)
        self.scene = sc
        sp = sc.params[spec.key]
        self.param_index = sp.index
        init = spec.init if isinstance(spec.init, torch.Tensor) else torch.as_tensor(np.asarray(spec.init, np.float32))
        init = init.to(sc.device, torch.float32).reshape(sp.shape).contiguous()
        self.ring = [init.clone(), init.clone(), torch.empty_like(init)]
        self.m = torch.zeros_like(init)
        self.v = torch.zeros_like(init)
        self.grad = torch.zeros_like(init)
        shape = sc.film_shape(spec.sensor)
        ref = spec.img_ref if isinstance(spec.img_ref, torch.Tensor) else torch.as_tensor(spec.img_ref)
        self.img_ref = ref.to(sc.device, torch.float32).reshape(shape).contiguous()
        self.H = torch.zeros(shape, device=sc.device)
        self.F = torch.zeros(shape, device=sc.device)
        self.img = torch.zeros(shape, device=sc.device)
        self.grad_img = torch.zeros(shape, device=sc.device)
        self.cv = torch.zeros((8,) + tuple(shape), device=sc.device)
        # bind the two states of the optimised parameter to ring[0] / ring[1]
        sp.states = [self.ring[0], self.ring[1]]
        sp.incoming = self.ring[1]
        sc._bind(sp)

    def footprint_bytes(self) -> int:
        """Device bytes this job holds (its buffers; the scene block is small)."""
        ts = self.ring + [self.m, self.v, self.grad, self.img_ref, self.H, self.F, self.img, self.grad_img, self.cv]
        return int(sum(t.numel() * t.element_size() for t in ts))

    def native(self, resume=None) -> BatchJob:
        s = self.spec
        itype = _ITYPE[s.integrator["type"]]
        d = N.IntegratorDesc()
        d.type = itype
        d.max_depth = int(s.integrator.get("max_depth", 6))
        d.rr_depth = int(s.integrator.get("rr_depth", 5))
        d.hide_emitters = int(bool(s.integrator.get("hide_emitters", False)))
        d.samples_per_thread = int(s.integrator.get("samples_per_thread", 0))
        j = BatchJob()
        j.scene = self.scene.handle.value
        j.integrator = d
        j.sensor = s.sensor
        j.seed = s.seed & 0xFFFFFFFF
        j.spp = s.spp
        j.spp_grad = s.spp_grad or s.spp
        j.method = 1 if s.method == "cv_ps" else 0
        if s.method not in ("cv_ps", "baseline"):
            raise ValueError(f"batched jobs support methods 'cv_ps' and 'baseline', not {s.method!r}")
        j.loss = N.LOSS_L1 if s.loss == "L1" else N.LOSS_L2
        j.param = self.param_index
        for k in range(3):
            j.param_ring[k] = self.ring[k].data_ptr()
        j.adam_m, j.adam_v, j.grad = self.m.data_ptr(), self.v.data_ptr(), self.grad.data_ptr()
        j.lr, j.beta1, j.beta2, j.eps = s.lr, s.beta1, s.beta2, s.eps
        j.clamp_lo, j.clamp_hi = float(s.clamp[0]), float(s.clamp[1])
        j.img_ref, j.H, j.F = self.img_ref.data_ptr(), self.H.data_ptr(), self.F.data_ptr()
        j.img, j.grad_img = self.img.data_ptr(), self.grad_img.data_ptr()
        for k in range(8):
            j.cv_state[k] = self.cv[k].data_ptr()
        j.warmup = s.warmup
        j.cv_beta1, j.cv_beta2 = s.cv_beta1, s.cv_beta2
        j.pre_update = int(s.pre_update)
        if resume is not None:
            j.resume = 1
            j.iteration, j.ring_old, j.ring_new, j.ring_next, j.next_key = resume
        return j


class BatchOptimizer:
    """``specs``: JobSpec (new jobs) or _JobBuffers taken from another batch;
    ``resume``: per job None or the tuple from ``read_job`` (state carried
    over when a continuous-batching engine rebuilds its batch)."""

    def __init__(self, specs: List, device="cuda", use_graph=True, block_size=0, queue_grab=0, resume=None,
                 stream=None):
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.use_graph = bool(use_graph)
        self.jobs = [s if isinstance(s, _JobBuffers) else _JobBuffers(s, self.device) for s in specs]
        resume = resume or [None] * len(self.jobs)
        self.stream = stream if stream is not None else torch.cuda.Stream(device=self.device)
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        self._h = C.c_void_p()
        lib = _lib()
        N.check(lib.rcv_batch_create(C.byref(self._h)), "rcv_batch_create")
        for jb, rs in zip(self.jobs, resume):
            out = C.c_int32()
            desc = jb.native(rs)
            N.check(lib.rcv_batch_add_job(self._h, C.byref(desc), C.byref(out)), "rcv_batch_add_job")
        N.check(lib.rcv_batch_commit(self._h, C.c_void_p(self.stream.cuda_stream)), "rcv_batch_commit")
        self.set_launch(block_size, queue_grab)
        self.iterations = 0

    def set_launch(self, block_size=0, queue_grab=0):
        N.check(_lib().rcv_batch_set_launch(self._h, int(block_size), int(queue_grab)), "rcv_batch_set_launch")
        self.launch = (int(block_size) or 256, int(queue_grab) or 256)

    def step(self, n: int = 1):
        lib = _lib()
        for _ in range(n):
            N.check(lib.rcv_batch_step(self._h, int(self.use_graph), C.c_void_p(self.stream.cuda_stream)),
                    "rcv_batch_step")
            self.iterations += 1

    def synchronize(self):
        self.stream.synchronize()

    def job_state(self, j: int):
        it = C.c_int32()
        slot = C.c_int32()
        N.check(_lib().rcv_batch_job_state(self._h, j, C.byref(it), C.byref(slot),
                                           C.c_void_p(self.stream.cuda_stream)), "rcv_batch_job_state")
        return it.value, slot.value

    def read_job(self, j: int):
        vals = [C.c_int32() for _ in range(4)] + [C.c_uint32()]
        N.check(_lib().rcv_batch_read_job(self._h, j, *[C.byref(v) for v in vals],
                                          C.c_void_p(self.stream.cuda_stream)), "rcv_batch_read_job")
        return tuple(v.value for v in vals)

    def set_active(self, mask):
        """Continuous batching: only jobs with mask[j] true advance."""
        n = len(self.jobs)
        words = (C.c_uint32 * max(1, (n + 31) // 32))()
        for j, on in enumerate(mask):
            if on:
                words[j >> 5] |= 1 << (j & 31)
        N.check(_lib().rcv_batch_set_active(self._h, words, n, C.c_void_p(self.stream.cuda_stream)),
                "rcv_batch_set_active")

    def param(self, j: int) -> torch.Tensor:
        """Latest optimiser value of job j's parameter (synchronises)."""
        _, slot = self.job_state(j)
        return self.jobs[j].ring[slot]

    def image(self, j: int) -> torch.Tensor:
        """Image of job j's last iteration (the CV image for cv_ps)."""
        self.synchronize()
        return self.jobs[j].img

    def close(self):
        if self._h.value:
            _lib().rcv_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(make_specs, blocks=(64, 128, 256, 512), grabs=(64, 128, 256, 1024), steps=3, device="cuda"):
    """Times ``steps`` batched iterations for every (block size, queue grab)
    pair on a throw-away batch built by ``make_specs()`` and returns the
    fastest pair with the full table of timings (ms per iteration). Block
    sizes the device cannot launch (LDS per block) are left out of the
    table."""
    table = {}
    opt = BatchOptimizer(make_specs(), device=device, use_graph=False)
    opt.step(1)
    opt.synchronize()
    for b in blocks:
        for g in grabs:
            try:
                opt.set_launch(b, g)
            except N.RcvError:
                continue
            opt.step(1)
            opt.synchronize()
            t0 = time.perf_counter()
            opt.step(steps)
            opt.synchronize()
            table[(b, g)] = (time.perf_counter() - t0) * 1e3 / steps
    opt.close()
    best = min(table, key=table.get)
    return best, table
