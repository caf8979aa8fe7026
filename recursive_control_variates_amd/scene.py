"""Device scene: a SceneDescription committed to librcv, with its parameter
tensors resident in device memory.

Two-state adapters (plugins/twostatebsdf.py:46-56, plugins/twostatemedium.py:
189-200) keep three copies (old, new, incoming) and copy old <- new <-
incoming on every update. Here the copies are tensor references: an update
rebinds the native OLD state to the tensor that was NEW and the NEW state to
the incoming tensor, so no parameter data moves (the optimizer writes each
step's result to a fresh buffer and never mutates a bound tensor; in-place
writes to a bound tensor are detected and rejected at render time).
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _native as N
from .scene_desc import SceneDescription, describe


class SceneParam:
    """One differentiable scene parameter (texture or grid)."""

    def __init__(self, index: int, key: str, kind: int, shape: tuple, twostate: bool, device):
        self.index = index
        self.key = key
        self.kind = kind
        self.shape = shape
        self.twostate = twostate
        self.device = device
        self.incoming: Optional[torch.Tensor] = None
        self.states: List[Optional[torch.Tensor]] = [None, None]  # [old, new]
        self.versions = [None, None]

    def check_unmodified(self):
        for s, t in enumerate(self.states):
            if t is not None and t._version != self.versions[s]:
                raise RuntimeError(
                    f"parameter {self.key!r}: a tensor bound as scene state was modified in place; "
                    "assign a new tensor through SceneParameters instead")


class Scene:
    """mi.Scene counterpart for the hot path (created by ``load_dict``)."""

    def __init__(self, desc: SceneDescription, device="cuda"):
        self.desc = desc
        dev = torch.device(device)
        if dev.type != "cuda":
            raise ValueError("Scene lives on a HIP device ('cuda' in torch)")
        self.device = dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())
        self._handle = C.c_void_p()
        N.call("rcv_scene_create", C.byref(self._handle))
        self.params: Dict[str, SceneParam] = {}
        self._param_list: List[SceneParam] = []
        try:
            self._build()
        except Exception:
            self._destroy()
            raise

    # -------------------------------------------------------------- build
    def _build(self):
        h = self._handle
        d = self.desc
        with torch.cuda.device(self.device):
            stream = N.stream_handle(self.device)
            for i, p in enumerate(d.params):
                out = C.c_int32()
                if p.kind == N.PARAM_VALUE:
                    nz = ny = nx = 1
                    ch = p.shape[0]
                elif p.kind == N.PARAM_TEXTURE:
                    ny, nx, ch = p.shape
                    nz = 1
                else:
                    nz, ny, nx, ch = p.shape
                N.call("rcv_scene_add_param", h, p.kind, nx, ny, nz, ch, C.byref(out))
                assert out.value == i
                sp = SceneParam(i, p.key, p.kind, p.shape, p.twostate, self.device)
                self.params[p.key] = sp
                self._param_list.append(sp)
            for b in d.bsdfs:
                out = C.c_int32()
                N.call("rcv_scene_add_bsdf", h, b.type, N.f32_array(b.rgb), b.texture_param, C.byref(out))
            for m in d.media:
                md = N.MediumDesc()
                md.sigma_t_param = m.sigma_param
                md.albedo_param = m.albedo_param
                md.albedo_rgb[:] = list(m.albedo_rgb)
                md.scale = m.scale
                md.bbox_min[:] = list(m.bbox_min)
                md.bbox_max[:] = list(m.bbox_max)
                md.sample_emitters = int(m.sample_emitters)
                md.majorant_factor = m.majorant_factor
                md.homogeneous = int(m.homogeneous)
                out = C.c_int32()
                N.call("rcv_scene_add_medium", h, C.byref(md), C.byref(out))
            for e in d.emitters:
                out = C.c_int32()
                N.call("rcv_scene_add_emitter", h, e.type, N.f32_array(e.rgb), C.byref(out))
            for s in d.shapes:
                out = C.c_int32()
                if s.kind == "sphere":
                    N.call("rcv_scene_add_sphere", h, N.f32_array(s.center), C.c_float(s.radius), s.bsdf,
                           s.interior, s.exterior, C.byref(out))
                elif s.box is not None:
                    N.call("rcv_scene_add_box", h, N.f32_array(np.asarray(s.box, np.float32).reshape(12)), s.bsdf,
                           s.interior, s.exterior, C.byref(out))
                else:
                    v = np.ascontiguousarray(s.vertices, np.float32)
                    f = np.ascontiguousarray(s.faces, np.int32)
                    uv = None if s.uv is None else np.ascontiguousarray(s.uv, np.float32)
                    N.call("rcv_scene_add_mesh", h, v.ctypes.data_as(C.POINTER(C.c_float)),
                           None if uv is None else uv.ctypes.data_as(C.POINTER(C.c_float)), len(v),
                           f.ctypes.data_as(C.POINTER(C.c_int32)), len(f), s.bsdf, s.emitter, s.interior,
                           s.exterior, C.byref(out))
            for sen in d.sensors:
                mats = np.concatenate([np.asarray(v.to_world, np.float32).reshape(16) for v in sen.views])
                fovs = [v.fov for v in sen.views]
                out = C.c_int32()
                N.call("rcv_scene_add_sensor", h, sen.film_width, sen.film_height, len(sen.views),
                       N.f32_array(mats), N.f32_array(fovs), C.byref(out))
            # initial data: incoming == new == old (initialize_parameters, utils.py:62-87)
            for p, sp in zip(d.params, self._param_list):
                t = torch.from_numpy(p.init).to(self.device)
                sp.incoming = t
                sp.states = [t, t]
                sp.versions = [t._version, t._version]
            N.call("rcv_scene_commit", h, C.c_void_p(stream))
            for sp in self._param_list:
                self._bind(sp)

    def _bind(self, sp: SceneParam):
        stream = C.c_void_p(N.stream_handle(self.device))
        for s in (N.STATE_OLD, N.STATE_NEW):
            t = sp.states[s]
            if t.dtype != torch.float32 or not t.is_contiguous() or t.device != self.device:
                raise ValueError(f"{sp.key}: parameters must be contiguous float32 tensors on {self.device}")
            if tuple(t.shape) != tuple(sp.shape):
                raise ValueError(f"{sp.key}: shape {tuple(t.shape)} does not match {sp.shape}")
            N.call("rcv_scene_bind_param", self._handle, sp.index, s, C.c_void_p(t.data_ptr()), stream)
            sp.versions[s] = t._version

    # ------------------------------------------------------------ updates
    def set_incoming(self, key: str, value: torch.Tensor):
        sp = self.params[key]
        if not isinstance(value, torch.Tensor):
            value = torch.as_tensor(np.asarray(value, np.float32))
        value = value.detach().to(self.device, torch.float32)
        if tuple(value.shape) != tuple(sp.shape):
            if value.dim() != len(sp.shape) or value.shape[-1] != sp.shape[-1]:
                raise ValueError(f"{key}: shape {tuple(value.shape)} is not a resolution change of {sp.shape}")
        sp.incoming = value.contiguous()

    def parameters_changed(self, keys):
        """old <- new <- incoming for two-state parameters; plain parameters
        take the incoming value for both states.

        An incoming value of another resolution (upsampling, utils.py:177-192)
        resizes the parameter. Both states of one parameter share one
        resolution here, so the state that becomes OLD is the previous NEW
        state resampled to the new resolution by voxel repetition (upstream
        keeps the coarse volume as OLD for one iteration; that iteration's
        OLD image is not used, because upsampling also resets the CV state)."""
        for k in keys:
            sp = self.params[k]
            if tuple(sp.incoming.shape) != tuple(sp.shape):
                self._resize(sp, tuple(sp.incoming.shape))
            if sp.twostate:
                sp.states = [sp.states[N.STATE_NEW], sp.incoming]
            else:
                sp.states = [sp.incoming, sp.incoming]
            self._bind(sp)

    def _resize(self, sp: SceneParam, shape: tuple):
        if sp.kind == N.PARAM_VALUE:
            raise ValueError(f"{sp.key}: a value parameter keeps its shape {sp.shape}")
        if sp.kind == N.PARAM_TEXTURE:
            ny, nx, _ = shape
            nz = 1
        else:
            nz, ny, nx, _ = shape
        N.call("rcv_scene_resize_param", self._handle, sp.index, nx, ny, nz,
               C.c_void_p(N.stream_handle(self.device)))
        sp.shape = shape
        self.desc.params[sp.index].shape = shape
        sp.states = [resample_nearest(t, shape) if t is not None else None for t in sp.states]

    def set_majorant_factor(self, medium: int, factor: int):
        """Medium.set_majorant_resolution_factor (plugins/twostatemedium.py:
        174-178): 0 = one global majorant, k > 0 = supergrid of ceil(n / k)
        cells per axis."""
        N.call("rcv_scene_set_majorant_factor", self._handle, medium, int(factor),
               C.c_void_p(N.stream_handle(self.device)))
        self.desc.media[medium].majorant_factor = int(factor)

    def majorant_resolution_factor(self, medium: int = 0) -> int:
        return int(self.desc.media[medium].majorant_factor)

    def check_params(self):
        for sp in self._param_list:
            sp.check_unmodified()

    # ------------------------------------------------------------- access
    @property
    def handle(self):
        return self._handle

    def sensors(self):
        return self.desc.sensors

    def film_shape(self, sensor: int = 0):
        s = self.desc.sensors[sensor]
        return (s.film_height, s.film_width, 3)

    def param_state(self, key: str, state: int) -> torch.Tensor:
        return self.params[key].states[state]

    def twostate_flags(self):
        return ([int(b.twostate) for b in self.desc.bsdfs], [int(m.twostate) for m in self.desc.media])

    def set_counting(self, enabled: bool):
        N.call("rcv_scene_set_counting", self._handle, int(enabled))

    def read_counters(self) -> Dict[str, int]:
        out = (C.c_int64 * 8)()
        N.call("rcv_scene_read_counters", self._handle, out, 8, C.c_void_p(N.stream_handle(self.device)))
        return dict(zip(N.COUNTER_NAMES, [int(x) for x in out]))

    def majorant_grid(self, medium: int = 0, state: int = 1):
        """Unscaled majorant supergrid of a medium for one state (0 old,
        1 new) as a float32 array of shape (z, y, x), or None when the medium
        uses one global majorant (Medium.majorant_grid(), read at
        plugins/twostatemedium.py:25-29)."""
        res = (C.c_int32 * 3)()
        stream = C.c_void_p(N.stream_handle(self.device))
        N.call("rcv_scene_majorant_grid", self._handle, medium, state, None, 0, res, stream)
        if res[0] == 0:
            return None
        out = np.empty((res[2], res[1], res[0]), np.float32)
        N.call("rcv_scene_majorant_grid", self._handle, medium, state, out.ctypes.data_as(C.c_void_p), out.size,
               res, stream)
        return out

    def _destroy(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            try:
                N.lib().rcv_scene_destroy(self._handle)
            except Exception:
                pass
            self._handle = None

    def __del__(self):
        self._destroy()


def resample_nearest(t: torch.Tensor, shape: tuple) -> torch.Tensor:
    """Nearest-neighbour resampling of a (ny, nx, c) texture or (nz, ny, nx, c)
    grid to ``shape`` (voxel repetition for integer factors, the tensor case
    of dr.upsample used at utils.py:184)."""
    x = t.movedim(-1, 0).unsqueeze(0)
    y = torch.nn.functional.interpolate(x, size=tuple(shape[:-1]), mode="nearest")
    return y.squeeze(0).movedim(0, -1).contiguous()


class SceneParameters:
    """Result of ``traverse(scene)`` (mi.traverse): key -> incoming tensor.

    ``update(values)`` assigns the given values (e.g. an optimizer) and then
    runs the two-state rotation for every changed key.
    """

    def __init__(self, scene: Scene):
        self.scene = scene
        self._dirty = set()
        # tensors assigned with requires_grad: render() returns their gradients
        self._grad_sources: Dict[str, torch.Tensor] = {}

    def keys(self):
        return list(self.scene.params.keys())

    def items(self):
        return [(k, self[k]) for k in self.keys()]

    def __contains__(self, key):
        return key in self.scene.params

    def __getitem__(self, key):
        src = self._grad_sources.get(key)
        return src if src is not None else self.scene.params[key].incoming

    def __setitem__(self, key, value):
        self.scene.set_incoming(key, value)
        if isinstance(value, torch.Tensor) and value.requires_grad:
            self._grad_sources[key] = value
        else:
            self._grad_sources.pop(key, None)
        self._dirty.add(key)

    def update(self, values=None):
        if values is not None:
            for k in values.keys():
                if k in self.scene.params:
                    self[k] = values[k]
        keys = sorted(self._dirty)
        self._dirty.clear()
        self.scene.parameters_changed(keys)
        return keys

    def differentiable(self):
        """(key, tensor) pairs assigned with requires_grad, in key order."""
        return [(k, self._grad_sources[k]) for k in sorted(self._grad_sources)]


def load_scene(scene_dict, device="cuda") -> Scene:
    return Scene(describe(scene_dict), device)


def traverse(scene: Scene) -> SceneParameters:
    return SceneParameters(scene)
