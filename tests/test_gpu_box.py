"""The 'cube' shape traced as one analytic oriented box (rcv_scene_add_box)
against the same cube traced as its 12-triangle mesh, with the same random
numbers: a textured diffuse cube under an area light (hit points, outward
normals and the per-face uv layout, for a rotated and non-uniformly scaled
and for a mirrored transform) and a heterogeneous volume behind a null-BSDF
cube boundary (the reference's volume scenes, experiments.py:74-79)."""
import numpy as np
import pytest
import torch

from conftest import announce
from helpers import cam, volume_box_scene
import recursive_control_variates_amd as rcv
from recursive_control_variates_amd.scene import Scene
from recursive_control_variates_amd.scene_desc import describe
from recursive_control_variates_amd.transform import T

pytestmark = pytest.mark.gpu


def _both(d):
    """(box scene, mesh scene) of one scene dict."""
    desc = describe(d)
    assert any(s.box is not None for s in desc.shapes)
    mesh = describe(d)
    for s in mesh.shapes:
        s.box = None
    return Scene(desc), Scene(mesh)


def _cube_scene(to_world, w=24, h=20, spp=16):
    tex = np.random.default_rng(3).uniform(0.1, 0.9, (8, 8, 3)).astype(np.float32)
    return {
        "type": "scene",
        "sensor": cam((2.2, 1.9, 2.8), (0.0, 0.4, 0.0), fov=45, w=w, h=h, spp=spp),
        "floor": {"type": "rectangle", "to_world": T.rotate([1, 0, 0], -90).scale(3.0),
                  "bsdf": {"type": "diffuse", "reflectance": {"type": "rgb", "value": 0.5}}},
        "light": {"type": "rectangle", "to_world": T.translate([0.3, 2.6, 0.2]).rotate([1, 0, 0], 90).scale(0.6),
                  "bsdf": {"type": "diffuse", "reflectance": {"type": "rgb", "value": 0.0}},
                  "emitter": {"type": "area", "radiance": {"type": "rgb", "value": 9.0}}},
        "cube": {"type": "cube", "to_world": to_world,
                 "bsdf": {"type": "diffuse", "reflectance": {"type": "bitmap", "data": tex}}},
    }


@pytest.mark.parametrize("mirrored", [False, True])
def test_textured_cube_box_matches_mesh(cuda, mirrored):
    announce(f"box vs mesh, textured cube (mirrored={mirrored})")
    tw = T.translate([0.0, 0.55, 0.0]).rotate([0, 1, 0], 33).rotate([1, 0, 0], 12).scale(
        [-0.5 if mirrored else 0.5, 0.35, 0.45])
    sb, sm = _both(_cube_scene(tw))
    integ = rcv.load_dict({"type": "prb", "max_depth": 5})
    a = integ.render(sb, 0, 11, 16).cpu().numpy()
    b = integ.render(sm, 0, 11, 16).cpu().numpy()
    close = np.abs(a - b) <= 1e-4 + 1e-3 * np.abs(b)
    assert close.mean() > 0.98, close.mean()
    assert abs(a.mean() - b.mean()) < 2e-3 * b.mean()


def test_volume_boundary_box_matches_mesh(cuda):
    announce("box vs mesh, volume boundary")
    g = np.random.default_rng(4).uniform(0.2, 1.0, (6, 6, 6, 1)).astype(np.float32)
    sb, sm = _both(volume_box_scene(g, scale=6.0, w=16, h=12, spp=16))
    integ = rcv.load_dict({"type": "twostateprbvolpath", "max_depth": 12, "rr_depth": 4})
    for seed in (3, 4):
        ha, fa = [x.cpu().numpy() for x in integ.render_twostates(sb, 0, seed, 16)]
        hb, fb = [x.cpu().numpy() for x in integ.render_twostates(sm, 0, seed, 16)]
        for a, b in ((ha, hb), (fa, fb)):
            close = np.abs(a - b) <= 1e-4 + 1e-3 * np.abs(b)
            assert close.mean() > 0.97, close.mean()
            assert abs(a.mean() - b.mean()) < 3e-3 * b.mean()
