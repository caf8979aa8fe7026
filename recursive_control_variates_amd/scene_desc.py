"""Dict scene descriptions -> flat scene description (pure Python / NumPy).

Accepts the subset of Mitsuba's ``load_dict`` vocabulary that the hot path's
scenes use (experiments.py:20-583, tutorial.ipynb cells 5/21) and that this
build implements:

* shapes: ``rectangle`` (local [-1,1]^2, normal +z), ``cube`` (local [-1,1]^3,
  outward normals), ``sphere`` (``center``/``radius``), ``mesh`` (explicit
  ``vertices``/``faces``/optional ``uv``); each with ``to_world``, ``bsdf``,
  ``emitter``, ``interior``/``exterior`` media;
* BSDFs: ``diffuse`` (``reflectance`` rgb / float / ``bitmap`` with ``data``
  or an OpenEXR ``filename``), ``null``, and the ``twostate`` adapter
  (plugins/twostatebsdf.py:4-56);
* media: ``heterogeneous`` with a ``gridvolume`` ``sigma_t`` (1 channel;
  ``data`` or a Mitsuba ``.vol`` ``filename``, placed by ``to_world`` or,
  with ``use_grid_bbox``, by the file's box) and an albedo (float / rgb /
  3-channel ``gridvolume``), ``scale``,
  ``sample_emitters``, ``majorant_resolution_factor``; ``homogeneous`` with
  a grey ``sigma_t`` (float or ``spectrum``/``rgb`` of one value) and an
  albedo (float / ``rgb`` / ``spectrum``), ``scale``, ``sample_emitters``
  (tutorial.ipynb:77-88), at most one per scene; and the
  ``twostatemedium`` adapter (plugins/twostatemedium.py:5-200);
* emitters: ``area`` (on a shape), ``constant`` (environment);
* sensors: ``perspective`` (``fov`` along x, ``to_world``) and ``batch`` of
  perspective sensors side by side on one ``hdrfilm`` (experiments.py:28-50);
  box reconstruction filter; ``independent`` sampler ``sample_count``.

Differentiable parameters get the reference's key names, e.g.
``back.bsdf.reflectance.data`` or ``object.interior_medium.sigma_t.data``
(experiments.py:104,546), ``bunny.interior_medium.sigma_t.value.value``
for a homogeneous medium (tutorial.ipynb:181); the two-state adapters expose the ``incoming``
copy's parameters under the same keys (plugins/twostatebsdf.py:43-44,
plugins/twostatemedium.py:186-187).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional

import numpy as np

from .transform import Transform4f

PARAM_TEXTURE = 0
PARAM_GRID = 1
PARAM_VALUE = 2
BSDF_DIFFUSE = 0
BSDF_NULL = 1
EMITTER_AREA = 0
EMITTER_CONSTANT = 1

SHAPE_TYPES = ("rectangle", "cube", "sphere", "mesh")
SENSOR_TYPES = ("perspective", "batch")


@dataclass
class ParamDesc:
    key: str
    kind: int
    shape: tuple            # texture (H, W, C); grid (D, H, W, C); value (C,)
    init: np.ndarray        # float32 initial data (incoming == new == old)
    twostate: bool


@dataclass
class BsdfDesc:
    type: int
    rgb: tuple
    texture_param: int
    twostate: bool


@dataclass
class MediumDesc:
    sigma_param: int
    albedo_param: int
    albedo_rgb: tuple
    scale: float
    bbox_min: tuple
    bbox_max: tuple
    sample_emitters: bool
    majorant_factor: int
    twostate: bool
    homogeneous: bool = False


@dataclass
class EmitterDesc:
    type: int
    rgb: tuple


@dataclass
class ShapeDesc:
    name: str
    kind: str                       # 'mesh' | 'sphere'
    bsdf: int
    emitter: int
    interior: int
    exterior: int
    vertices: Optional[np.ndarray] = None   # (n, 3) float32
    uv: Optional[np.ndarray] = None         # (n, 2) float32
    faces: Optional[np.ndarray] = None      # (m, 3) int32
    center: Optional[tuple] = None
    radius: float = 0.0
    box: Optional[np.ndarray] = None        # 'cube': (3, 4) float32 object-to-world map of [-1,1]^3


@dataclass
class ViewDesc:
    to_world: np.ndarray   # 4x4 float64, columns (left, up, forward, origin)
    fov: float             # degrees, along x


@dataclass
class SensorDesc:
    film_width: int
    film_height: int
    views: List[ViewDesc]
    spp: int


@dataclass
class SceneDescription:
    params: List[ParamDesc] = field(default_factory=list)
    bsdfs: List[BsdfDesc] = field(default_factory=list)
    media: List[MediumDesc] = field(default_factory=list)
    emitters: List[EmitterDesc] = field(default_factory=list)
    shapes: List[ShapeDesc] = field(default_factory=list)
    sensors: List[SensorDesc] = field(default_factory=list)

    def param_index(self, key: str) -> int:
        for i, p in enumerate(self.params):
            if p.key == key:
                return i
        raise KeyError(key)

    # flat arrays (used by the native scene and by test oracles)
    def triangles(self):
        """(T, 9) vertices, (T, 6) uvs, (T,) shape ids, ordered by shape."""
        tv, tu, ts = [], [], []
        for si, s in enumerate(self.shapes):
            if s.kind != "mesh":
                continue
            v = s.vertices[s.faces].reshape(-1, 9)
            uv = (s.uv if s.uv is not None else np.zeros((len(s.vertices), 2), np.float32))[s.faces].reshape(-1, 6)
            tv.append(v)
            tu.append(uv)
            ts.append(np.full(len(s.faces), si, np.int32))
        if not tv:
            return np.zeros((0, 9), np.float32), np.zeros((0, 6), np.float32), np.zeros(0, np.int32)
        return (np.concatenate(tv).astype(np.float32), np.concatenate(tu).astype(np.float32),
                np.concatenate(ts).astype(np.int32))

    def spheres(self):
        sp, ss = [], []
        for si, s in enumerate(self.shapes):
            if s.kind == "sphere":
                sp.append([*s.center, s.radius])
                ss.append(si)
        return np.asarray(sp, np.float32).reshape(-1, 4), np.asarray(ss, np.int32)


def _grid_data(d, prefix):
    """(grid, file bbox or None) of a gridvolume: inline ``data`` or a
    ``filename`` of a float32 Mitsuba .vol file; the file's box is used only
    with ``use_grid_bbox`` (experiments.py sets it to False and places the
    grid with ``to_world``)."""
    if "data" in d:
        return d["data"], None
    if "filename" in d:
        from .fileio import read_vol
        arr, lo, hi = read_vol(d["filename"])
        return arr, ((lo, hi) if d.get("use_grid_bbox", False) else None)
    raise ValueError(f"{prefix}: gridvolume needs 'data' or 'filename'")


def _bitmap_data(d, prefix):
    """Texels of a bitmap: inline ``data`` or an OpenEXR ``filename``."""
    if "data" in d:
        return d["data"]
    if "filename" in d:
        fn = d["filename"]
        if not fn.lower().endswith(".exr"):
            raise ValueError(f"{prefix}: bitmap files must be OpenEXR (.exr)")
        from .fileio import read_exr_rgb
        return read_exr_rgb(fn)
    raise ValueError(f"{prefix}: bitmap needs 'data' or 'filename'")


def _box_matrix(lo, hi):
    m = np.eye(4)
    for k in range(3):
        m[k, k] = hi[k] - lo[k]
        m[k, 3] = lo[k]
    return m


def _as_rgb(v) -> tuple:
    if isinstance(v, dict):
        t = v.get("type")
        if t in ("rgb", "spectrum"):
            return _as_rgb(v["value"])
        raise ValueError(f"unsupported spectrum type {t!r}")
    a = np.broadcast_to(np.asarray(v, np.float64), (3,))
    return (float(a[0]), float(a[1]), float(a[2]))


def _to_numpy(data) -> np.ndarray:
    if hasattr(data, "detach"):
        data = data.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(data, dtype=np.float32))


def _transform(d) -> Transform4f:
    t = d.get("to_world")
    if t is None:
        return Transform4f()
    if isinstance(t, Transform4f):
        return t
    return Transform4f(t)


# -------------------------------------------------------------- geometry
def _rectangle():
    v = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float64)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return v, uv, f


def _cube():
    # 6 faces of [-1,1]^3, counter-clockwise seen from outside (outward normals)
    verts, uvs, faces = [], [], []
    quads = [
        ([1, -1, -1], [1, 1, -1], [1, 1, 1], [1, -1, 1]),      # +x
        ([-1, -1, -1], [-1, -1, 1], [-1, 1, 1], [-1, 1, -1]),  # -x
        ([-1, 1, -1], [-1, 1, 1], [1, 1, 1], [1, 1, -1]),      # +y
        ([-1, -1, -1], [1, -1, -1], [1, -1, 1], [-1, -1, 1]),  # -y
        ([-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]),      # +z
        ([-1, -1, -1], [-1, 1, -1], [1, 1, -1], [1, -1, -1]),  # -z
    ]
    for q in quads:
        b = len(verts)
        verts.extend(q)
        uvs.extend([[0, 0], [1, 0], [1, 1], [0, 1]])
        faces.extend([[b, b + 1, b + 2], [b, b + 2, b + 3]])
    return np.asarray(verts, np.float64), np.asarray(uvs, np.float32), np.asarray(faces, np.int32)


class _Loader:
    def __init__(self):
        self.desc = SceneDescription()

    # ---- parameters
    def _add_param(self, key, kind, data, twostate):
        arr = _to_numpy(data)
        if kind == PARAM_VALUE:
            arr = arr.reshape(-1)
            if arr.shape[0] not in (1, 3):
                raise ValueError(f"{key}: a value parameter has 1 or 3 channels, got {arr.shape}")
        elif kind == PARAM_TEXTURE:
            if arr.ndim == 2:
                arr = arr[..., None]
            if arr.ndim != 3 or arr.shape[2] not in (1, 3):
                raise ValueError(f"{key}: bitmap data must be (H, W, 3), got {arr.shape}")
        else:
            if arr.ndim == 3:
                arr = arr[..., None]
            if arr.ndim != 4 or arr.shape[3] not in (1, 3):
                raise ValueError(f"{key}: grid data must be (D, H, W, C), got {arr.shape}")
        self.desc.params.append(ParamDesc(key, kind, tuple(arr.shape), arr.copy(), twostate))
        return len(self.desc.params) - 1

    # ---- BSDFs
    def _bsdf(self, prefix, d):
        t = d.get("type")
        twostate = False
        if t == "twostate":
            for k in ("old", "new", "incoming"):
                if k not in d:
                    raise ValueError(f"{prefix}: twostate BSDF needs 'old', 'new' and 'incoming'")
            types = {d[k].get("type") for k in ("old", "new", "incoming")}
            if len(types) != 1:
                raise ValueError(f"{prefix}: twostate old/new/incoming must share one BSDF type")
            twostate = True
            d = d["incoming"]
            t = d.get("type")
        if t == "null":
            self.desc.bsdfs.append(BsdfDesc(BSDF_NULL, (0.0, 0.0, 0.0), -1, twostate))
        elif t == "diffuse":
            refl = d.get("reflectance", 0.5)
            tex = -1
            rgb = (0.0, 0.0, 0.0)
            if isinstance(refl, dict) and refl.get("type") == "bitmap":
                tex = self._add_param(f"{prefix}.reflectance.data", PARAM_TEXTURE, _bitmap_data(refl, prefix),
                                      twostate)
                if self.desc.params[tex].shape[2] != 3:
                    raise ValueError(f"{prefix}: reflectance bitmap must have 3 channels")
            else:
                rgb = _as_rgb(refl)
            self.desc.bsdfs.append(BsdfDesc(BSDF_DIFFUSE, rgb, tex, twostate))
        else:
            raise ValueError(f"{prefix}: unsupported BSDF type {t!r}")
        return len(self.desc.bsdfs) - 1

    # ---- media
    def _medium(self, prefix, d):
        t = d.get("type")
        twostate = False
        if t == "twostatemedium":
            for k in ("old", "new", "incoming"):
                if k not in d:
                    raise ValueError(f"{prefix}: twostatemedium needs 'old', 'new' and 'incoming'")
            twostate = True
            d = d["incoming"]
            t = d.get("type")
        if t not in ("heterogeneous", "homogeneous"):
            raise ValueError(f"{prefix}: unsupported medium type {t!r} (heterogeneous, homogeneous)")
        if d.get("has_spectral_extinction", False):
            raise ValueError(f"{prefix}: spectrally varying extinction is not supported")
        phase = d.get("phase", {"type": "isotropic"})
        if phase.get("type") != "isotropic":
            raise ValueError(f"{prefix}: only the isotropic phase function is supported")
        if t == "homogeneous":
            return self._homogeneous(prefix, d, twostate)
        sig = d.get("sigma_t")
        if not (isinstance(sig, dict) and sig.get("type") == "gridvolume"):
            raise ValueError(f"{prefix}: sigma_t must be a gridvolume")
        sig_data, sig_box = _grid_data(sig, prefix)
        tw = _transform(sig).matrix if sig_box is None else _box_matrix(*sig_box)
        if np.any(np.abs(tw[:3, :3] - np.diag(np.diag(tw[:3, :3]))) > 1e-12) or np.any(np.diag(tw[:3, :3]) <= 0):
            raise ValueError(f"{prefix}: grid to_world must be an axis-aligned positive scale + translation")
        lo = tw[:3, 3]
        hi = lo + np.diag(tw[:3, :3])
        sp = self._add_param(f"{prefix}.sigma_t.data", PARAM_GRID, sig_data, twostate)
        if self.desc.params[sp].shape[3] != 1:
            raise ValueError(f"{prefix}: sigma_t grid must have 1 channel")
        alb = d.get("albedo", 0.75)
        ap = -1
        argb = (0.0, 0.0, 0.0)
        if isinstance(alb, dict) and alb.get("type") == "gridvolume":
            alb_data, alb_box = _grid_data(alb, prefix)
            ap = self._add_param(f"{prefix}.albedo.data", PARAM_GRID, alb_data, twostate)
            if self.desc.params[ap].shape[3] != 3:
                raise ValueError(f"{prefix}: albedo grid must have 3 channels")
            aw = _transform(alb).matrix if alb_box is None else _box_matrix(*alb_box)
            if not np.allclose(aw, tw):
                raise ValueError(f"{prefix}: albedo and sigma_t grids must share to_world")
        else:
            argb = _as_rgb(alb)
        self.desc.media.append(MediumDesc(
            sp, ap, argb, float(d.get("scale", 1.0)), tuple(float(x) for x in lo), tuple(float(x) for x in hi),
            bool(d.get("sample_emitters", True)), int(d.get("majorant_resolution_factor", 0)), twostate))
        return len(self.desc.media) - 1

    def _homogeneous(self, prefix, d, twostate):
        """Mitsuba's 'homogeneous' medium: constant sigma_t (grey) and albedo,
        keys ``<prefix>.sigma_t.value.value`` and ``<prefix>.albedo.value.value``."""
        if any(m.homogeneous for m in self.desc.media):
            raise ValueError(f"{prefix}: at most one homogeneous medium per scene")
        sig = _as_rgb(d.get("sigma_t", 1.0))
        if not (sig[0] == sig[1] == sig[2]):
            raise ValueError(f"{prefix}: sigma_t must be grey (spectrally varying extinction is not supported)")
        sp = self._add_param(f"{prefix}.sigma_t.value.value", PARAM_VALUE, np.float32([sig[0]]), twostate)
        alb = d.get("albedo", 0.75)
        if isinstance(alb, dict) and alb.get("type") not in ("rgb", "spectrum"):
            raise ValueError(f"{prefix}: a homogeneous medium's albedo is a constant (float, rgb or spectrum)")
        ap = self._add_param(f"{prefix}.albedo.value.value", PARAM_VALUE, np.float32(_as_rgb(alb)), twostate)
        self.desc.media.append(MediumDesc(
            sp, ap, (0.0, 0.0, 0.0), float(d.get("scale", 1.0)), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0),
            bool(d.get("sample_emitters", True)), 0, twostate, homogeneous=True))
        return len(self.desc.media) - 1

    # ---- shapes
    def _shape(self, name, d):
        t = d["type"]
        bsdf = self._bsdf(f"{name}.bsdf", d.get("bsdf", {"type": "diffuse", "reflectance": 0.5}))
        emitter = -1
        if "emitter" in d:
            e = d["emitter"]
            if e.get("type") != "area":
                raise ValueError(f"{name}: shape emitters must be of type 'area'")
            self.desc.emitters.append(EmitterDesc(EMITTER_AREA, _as_rgb(e.get("radiance", 1.0))))
            emitter = len(self.desc.emitters) - 1
        interior = self._medium(f"{name}.interior_medium", d["interior"]) if "interior" in d else -1
        exterior = self._medium(f"{name}.exterior_medium", d["exterior"]) if "exterior" in d else -1
        tw = _transform(d)
        if t == "sphere":
            if emitter >= 0:
                raise ValueError(f"{name}: sphere emitters are not supported")
            c = np.asarray(d.get("center", [0, 0, 0]), np.float64)
            r = float(d.get("radius", 1.0))
            c = tw.transform_point(c)
            m = tw.matrix[:3, :3]
            s = np.linalg.svd(m, compute_uv=False)
            if not np.allclose(s, s[0], rtol=1e-6):
                raise ValueError(f"{name}: sphere to_world must be a similarity transform")
            self.desc.shapes.append(ShapeDesc(name, "sphere", bsdf, emitter, interior, exterior,
                                              center=tuple(float(x) for x in c), radius=r * float(s[0])))
            return
        if t == "rectangle":
            v, uv, f = _rectangle()
        elif t == "cube":
            v, uv, f = _cube()
        elif t == "mesh":
            v = np.asarray(d["vertices"], np.float64).reshape(-1, 3)
            f = np.asarray(d["faces"], np.int32).reshape(-1, 3)
            uv = np.asarray(d["uv"], np.float32).reshape(-1, 2) if "uv" in d else None
        else:
            raise ValueError(f"{name}: unsupported shape type {t!r}")
        vw = tw.transform_point(v).astype(np.float32)
        if np.linalg.det(tw.matrix[:3, :3]) < 0:
            f = f[:, ::-1].copy()
        # a non-emitting cube is also described as an analytic box (the GPU
        # traces it as one primitive; the triangles describe the same surface)
        box = np.ascontiguousarray(tw.matrix[:3, :4], np.float32) if t == "cube" and emitter < 0 else None
        self.desc.shapes.append(ShapeDesc(name, "mesh", bsdf, emitter, interior, exterior,
                                          vertices=vw, uv=uv, faces=np.ascontiguousarray(f, np.int32), box=box))

    # ---- sensors
    def _sensor(self, name, d):
        film = d.get("film", {"type": "hdrfilm"})
        rf = film.get("rfilter", {"type": "box"})
        if rf.get("type", "box") != "box":
            raise ValueError(f"{name}: only the box reconstruction filter is supported")
        w, h = int(film.get("width", 768)), int(film.get("height", 576))
        spp = int(d.get("sampler", {}).get("sample_count", 4))
        if d["type"] == "perspective":
            views = [ViewDesc(_transform(d).matrix.copy(), float(d.get("fov", 45.0)))]
        else:
            subs = [(k, v) for k, v in d.items() if isinstance(v, dict) and v.get("type") == "perspective"]
            if not subs:
                raise ValueError(f"{name}: batch sensor without perspective sub-sensors")
            views = [ViewDesc(_transform(v).matrix.copy(), float(v.get("fov", 45.0))) for _, v in subs]
            if w % len(views):
                raise ValueError(f"{name}: batch film width {w} is not a multiple of {len(views)} sensors")
        if d.get("fov_axis", "x") != "x":
            raise ValueError(f"{name}: only fov_axis 'x' is supported")
        self.desc.sensors.append(SensorDesc(w, h, views, spp))

    def load(self, d: Dict[str, Any]) -> SceneDescription:
        if d.get("type") != "scene":
            raise ValueError("scene dict must have type 'scene'")
        for name, obj in d.items():
            if name == "type":
                continue
            if not isinstance(obj, dict) or "type" not in obj:
                raise ValueError(f"{name}: expected a dict with a 'type'")
            t = obj["type"]
            if t in SHAPE_TYPES:
                self._shape(name, obj)
            elif t in SENSOR_TYPES:
                self._sensor(name, obj)
            elif t == "constant":
                self.desc.emitters.append(EmitterDesc(EMITTER_CONSTANT, _as_rgb(obj.get("radiance", 1.0))))
            elif t in ("path", "prb", "prbvolpath", "volpath", "twostateprb", "twostateprbvolpath", "cv", "meta"):
                continue  # integrators are passed to render() explicitly
            else:
                raise ValueError(f"{name}: unsupported object type {t!r}")
        if not self.desc.sensors:
            raise ValueError("the scene has no sensor")
        if sum(e.type == EMITTER_CONSTANT for e in self.desc.emitters) > 1:
            raise ValueError("at most one constant emitter")
        return self.desc


def describe(scene_dict: Dict[str, Any]) -> SceneDescription:
    return _Loader().load(scene_dict)
