"""Host-side logic on the CPU: the C ABI library loads and exports every
symbol include/rcv.h declares; scene dicts, parameter keys and plugin
property handling behave like the reference's."""
import os
import re

import numpy as np
import pytest

import recursive_control_variates_amd as rcv
from recursive_control_variates_amd import _native as N
from recursive_control_variates_amd import scene_desc, scenes
from recursive_control_variates_amd.transform import T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "rcv.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(rcv_\w+)\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert lib.rcv_version() >= 1
    # every declared entry point has a ctypes signature in the binding (the
    # batch entry points register theirs when the batch module loads)
    import recursive_control_variates_amd.batch  # noqa: F401
    for s in syms:
        assert s in N.SIGNATURES or s in ("rcv_version", "rcv_last_error"), s


def test_volume_scene_description():
    g = scenes.blob_grid(16, 1)
    d = scene_desc.describe(scenes.volume_scene_dict(g, film_res=32, spp=4))
    assert [p.key for p in d.params] == [scenes.VOLUME_KEY]
    assert d.params[0].shape == (16, 16, 16, 1) and d.params[0].twostate
    m = d.media[0]
    assert m.bbox_min == (0.0, 0.0, 0.0) and m.bbox_max == (1.0, 1.0, 1.0) and m.scale == 20.0
    assert d.sensors[0].film_width == 128 and len(d.sensors[0].views) == 4
    tv, tu, ts = d.triangles()
    assert tv.shape == (12, 9)
    # cube spans [0,1]^3 with outward normals
    v = tv.reshape(-1, 3, 3)
    assert np.allclose(v.min((0, 1)), 0) and np.allclose(v.max((0, 1)), 1)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    c = v.mean(1) - 0.5
    assert np.all(np.sum(n * c, 1) > 0)


def test_cornell_scene_description():
    d = scene_desc.describe(scenes.cornell_scene_dict(scenes.texture_target(), film_res=16))
    assert [p.key for p in d.params] == [scenes.TEXTURE_KEY]
    assert d.params[0].shape == (64, 64, 3)
    # all wall normals point into the box
    tv, _, ts = d.triangles()
    v = tv.reshape(-1, 3, 3)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    walls = [i for i, s in enumerate(d.shapes) if s.name != "light" and s.kind == "mesh"]
    for t in range(len(ts)):
        if ts[t] in walls:
            assert np.dot(n[t], -v[t].mean(0)) > 0
    sp, ss = d.spheres()
    assert sp.shape == (1, 4) and np.isclose(sp[0, 3], 0.3)
    assert d.emitters[0].rgb == (15.0, 15.0, 15.0)


def test_scene_desc_errors():
    g = np.zeros((4, 4, 4, 1), np.float32)
    base = scenes.volume_scene_dict(g)
    bad = dict(base)
    bad["object"] = dict(base["object"], interior={"type": "nanovdb"})
    with pytest.raises(ValueError, match="heterogeneous, homogeneous"):
        scene_desc.describe(bad)
    bad["object"] = dict(base["object"], interior={"type": "homogeneous", "sigma_t": {"type": "rgb",
                                                                                     "value": [1, 2, 3]}})
    with pytest.raises(ValueError, match="grey"):
        scene_desc.describe(bad)
    bad["object2"] = dict(base["object"], interior={"type": "homogeneous"})
    bad["object"] = dict(base["object"], interior={"type": "homogeneous"})
    with pytest.raises(ValueError, match="at most one homogeneous"):
        scene_desc.describe(bad)
    bad = dict(base)
    bad["envmap2"] = {"type": "constant"}
    with pytest.raises(ValueError, match="one constant"):
        scene_desc.describe(bad)
    bad = dict(base)
    bad["sensor"] = dict(base["sensor"], film={"type": "hdrfilm", "width": 10, "height": 10,
                                                 "rfilter": {"type": "gaussian"}})
    with pytest.raises(ValueError, match="box"):
        scene_desc.describe(bad)
    with pytest.raises(ValueError, match="type 'scene'"):
        scene_desc.describe({"type": "cube"})


def test_homogeneous_medium_description():
    # the tutorial's medium (tutorial.ipynb:77-88) inside a null-BSDF sphere,
    # wrapped in the two-state adapter (tutorial.ipynb:310-316)
    medium = {"type": "homogeneous", "scale": 5, "sigma_t": {"type": "spectrum", "value": 2.5},
              "albedo": {"type": "rgb", "value": (0.85, 0.97, 0.999999)}}
    d = scene_desc.describe({
        "type": "scene",
        "bunny": {"type": "sphere", "radius": 0.5, "bsdf": {"type": "null"},
                  "interior": {"type": "twostatemedium", "old": medium, "new": medium, "incoming": medium}},
        "env": {"type": "constant"},
        "sensor": {"type": "perspective", "film": {"type": "hdrfilm", "width": 8, "height": 8}},
    })
    keys = [p.key for p in d.params]
    assert keys == ["bunny.interior_medium.sigma_t.value.value", "bunny.interior_medium.albedo.value.value"]
    assert [p.kind for p in d.params] == [scene_desc.PARAM_VALUE] * 2
    assert d.params[0].shape == (1,) and d.params[1].shape == (3,)
    np.testing.assert_array_equal(d.params[0].init, np.float32([2.5]))
    np.testing.assert_array_equal(d.params[1].init, np.float32([0.85, 0.97, 0.999999]))
    m = d.media[0]
    assert m.homogeneous and m.twostate and m.scale == 5.0 and m.majorant_factor == 0
    # a plain float sigma_t and albedo
    d = scene_desc.describe({
        "type": "scene", "s": {"type": "sphere", "bsdf": {"type": "null"},
                               "interior": {"type": "homogeneous", "sigma_t": 1.5, "albedo": 0.5}},
        "sensor": {"type": "perspective"}})
    assert d.params[0].init[0] == np.float32(1.5) and np.all(d.params[1].init == np.float32(0.5))
    assert not d.media[0].twostate


def test_transform_conventions():
    m = T.look_at(origin=[0, 0, 3], target=[0, 0, 0], up=[0, 1, 0]).matrix
    assert np.allclose(m[:3, 2], [0, 0, -1]) and np.allclose(m[:3, 3], [0, 0, 3])
    assert np.allclose(m[:3, 1], [0, 1, 0]) and np.allclose(m[:3, 0], [-1, 0, 0])
    # translate(a).scale(b) applies scale first
    p = T.translate([1, 2, 3]).scale(2.0).transform_point([1, 1, 1])
    assert np.allclose(p, [3, 4, 5])


def test_plugin_registry_and_props():
    i = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 16, "rr_depth": 16})
    assert i.max_depth == 16 and i.rr_depth == 16 and not i.hide_emitters
    p = rcv.load_dict({"type": "prb"})
    assert p.max_depth == 6 and p.rr_depth == 5           # reference plugin defaults
    cv = rcv.load_dict({"type": "cv", "integrator": {"type": "twostateprb", "max_depth": 4}})
    assert cv.warmup == 1 and cv.beta1 == 0.9 and cv.beta2 == 0.999
    assert cv.adjoint_integrator is cv.integrator
    with pytest.raises(ValueError, match="twostate"):
        rcv.load_dict({"type": "cv", "integrator": {"type": "prb"}})
    m = rcv.load_dict({"type": "meta", "method": "cv_ps", "integrator": {"type": "twostateprbvolpath"},
                       "pre_update": True})
    assert m.method == "cv_ps" and m.pre_update
    with pytest.raises(ValueError):
        rcv.load_dict({"type": "meta", "method": "cv_xx", "integrator": {"type": "prb"}})
    with pytest.raises(NotImplementedError):
        rcv.load_dict({"type": "meta", "denoise": True, "integrator": {"type": "prb"}})
    with pytest.raises(ValueError):
        rcv.load_dict({"type": "nope"})
    # reset() drops the CV state (upsampling calls it, utils.py:192)
    cv.init = True
    cv.reset()
    assert cv.init is False


def test_render_argument_checks():
    from recursive_control_variates_amd.rendering import render, sample_tea_32

    class FakeScene:
        def sensors(self):
            class S:
                spp = 4
            return [S()]

    integ = rcv.load_dict({"type": "prb"})
    with pytest.raises(ValueError, match="different"):
        render(FakeScene(), None, 0, integ, seed=5, seed_grad=5)
    with pytest.raises(ValueError, match="integrator"):
        render(FakeScene(), None, 0, None)
    with pytest.raises(TypeError):
        render(FakeScene(), None, "0", integ)
    # default seed_grad = TEA(seed, 1)[0] decorrelates primal and adjoint
    assert sample_tea_32(0, 1)[0] != 0


def test_adam_host_state():
    from recursive_control_variates_amd.optim import Adam
    opt = Adam(lr=0.1)
    import torch
    opt["a"] = torch.zeros(3)
    assert opt["a"].requires_grad and opt.t["a"] == 0
    opt.set_learning_rate({"a": 0.5})
    assert opt.lr_k["a"] == 0.5
    opt.set_learning_rate(0.2)
    assert opt.lr == 0.2 and opt.lr_k == {"a": 0.5}
    opt["a"] = torch.ones(3)        # same shape: state kept
    opt["a"] = torch.ones(4)        # new shape: state reset
    assert opt.state["a"][0].shape == (4,)


def test_adjust_majorant_res_factor_rule():
    # utils.py:27-43: largest factor <= 8 with >= 4 cells on the shortest side
    from recursive_control_variates_amd.optimize import adjust_majorant_res_factor
    from recursive_control_variates_amd.scene_desc import MediumDesc

    class _Scene:
        def __init__(self):
            self.desc = type("D", (), {})()
            self.desc.media = [MediumDesc(0, -1, (0.8,) * 3, 1.0, (0,) * 3, (1,) * 3, True, 0, False)]
            self.calls = []

        def majorant_resolution_factor(self, m):
            return self.desc.media[m].majorant_factor

        def set_majorant_factor(self, m, k):
            self.calls.append(k)
            self.desc.media[m].majorant_factor = k

    expect = {4: 0, 7: 0, 8: 2, 11: 2, 12: 3, 16: 4, 20: 5, 31: 7, 32: 8, 64: 8, 256: 8}
    for side, k in expect.items():
        s = _Scene()
        adjust_majorant_res_factor(s, (side, side + 5, side + 9, 1))
        assert s.majorant_resolution_factor(0) == k, (side, k)
        assert s.calls == ([] if k == 0 else [k])


def test_cube_shapes_carry_an_analytic_box():
    # a non-emitting cube is uploaded as one oriented box (its triangles stay
    # in the description for the oracle); emitting cubes and other shapes
    # stay meshes
    from recursive_control_variates_amd.scene_desc import describe
    from recursive_control_variates_amd.transform import T
    tw = T.translate([0.5, 0.5, 0.5]).rotate([0, 1, 0], 30).scale([0.5, 0.25, 0.5])
    d = {"type": "scene",
         "sensor": {"type": "perspective", "film": {"type": "hdrfilm", "width": 4, "height": 4}},
         "a": {"type": "cube", "to_world": tw, "bsdf": {"type": "null"}},
         "b": {"type": "cube", "emitter": {"type": "area", "radiance": {"type": "rgb", "value": 1.0}}},
         "c": {"type": "rectangle"}}
    desc = describe(d)
    by_name = {s.name: s for s in desc.shapes}
    box = by_name["a"].box
    assert box is not None and box.shape == (3, 4)
    np.testing.assert_allclose(box, np.asarray(tw.matrix, np.float64)[:3, :4], rtol=1e-6, atol=1e-7)
    # the box maps the mesh's object-space corners onto its world vertices
    v = by_name["a"].vertices
    obj = np.array([[1, -1, -1], [1, 1, -1], [1, 1, 1], [1, -1, 1]], np.float64)
    np.testing.assert_allclose(obj @ box[:, :3].T.astype(np.float64) + box[:, 3], v[:4], rtol=1e-6, atol=1e-6)
    assert by_name["b"].box is None and by_name["c"].box is None


def test_adam_scalar_rate_keeps_per_key_rates():
    # mi.ad.Optimizer: a scalar set_learning_rate replaces the default rate
    # only; a key given its own rate (lr_factor, utils.py:82-83) keeps it
    # through the schedule's halvings (optimize.py:139-141)
    import torch
    from recursive_control_variates_amd.optim import Adam
    lr = 0.01
    opt = Adam(lr=lr)
    opt["a"] = torch.zeros(2)
    opt["b"] = torch.zeros(2)
    opt.set_learning_rate({"a": 2 * lr})
    opt.set_learning_rate(lr / 2)
    assert opt.lr_k.get("a", opt.lr) == 2 * lr
    assert opt.lr_k.get("b", opt.lr) == lr / 2


def test_integration_binding_matches_native_structs():
    # the reference-side ctypes binding shown in INTEGRATION.md declares the
    # same IntegratorDesc fields as the package's binding (and rcv.h)
    import ctypes as C
    import re
    from recursive_control_variates_amd import _native
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"class IntegratorDesc\(C\.Structure\):\n(.*?)\n\n", text, re.S)
    assert m, "INTEGRATION.md lost its IntegratorDesc binding"
    ns = {"C": C}
    exec("class IntegratorDesc(C.Structure):\n" + m.group(1), ns)
    shown = [(n, t) for n, t in ns["IntegratorDesc"]._fields_]
    assert shown == list(_native.IntegratorDesc._fields_)
    hdr = open(os.path.join(ROOT, "include", "rcv.h")).read()
    body = re.search(r"typedef struct rcv_integrator_desc \{(.*?)\} rcv_integrator_desc;", hdr, re.S).group(1)
    names = re.findall(r"int32_t (\w+);", body)
    assert names == [n for n, _ in shown]
    # the medium descriptor likewise (field names, order and sizes)
    m = re.search(r"class MediumDesc\(C\.Structure\):.*?\n(.*?)\n\n", text, re.S)
    assert m, "INTEGRATION.md lost its MediumDesc binding"
    exec("class MediumDesc(C.Structure):\n" + m.group(1), ns)
    shown = [(n, t) for n, t in ns["MediumDesc"]._fields_]
    ref = list(_native.MediumDesc._fields_)
    assert [n for n, _ in shown] == [n for n, _ in ref]
    assert [C.sizeof(t) for _, t in shown] == [C.sizeof(t) for _, t in ref]
    body = re.search(r"typedef struct rcv_medium_desc \{(.*?)\} rcv_medium_desc;", hdr, re.S).group(1)
    assert re.findall(r"(?:int32_t|float) (\w+)(?:\[3\])?;", body) == [n for n, _ in shown]


def test_run_opt_rejects_unsupported_keys():
    from recursive_control_variates_amd.optimize import run_opt
    base = {"scene": object(), "integrator": object(), "variables": {"k": {"init": 0.5}}}
    with pytest.raises(ValueError, match="final_res"):
        run_opt(dict(base, final_res=256))
    with pytest.raises(ValueError, match="largesteps"):
        run_opt(dict(base, variables={"k": {"init": 0.5, "largesteps": object()}}))
    with pytest.raises(ValueError, match="VGG"):
        run_opt(dict(base, loss="VGG"))
    with pytest.raises(NotImplementedError):
        run_opt(dict(base, denoise=True))
