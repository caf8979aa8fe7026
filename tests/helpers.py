"""Small scenes with known answers, shared by CPU and GPU tests."""
import numpy as np

from recursive_control_variates_amd.transform import T


def cam(origin, target, fov=45.0, w=8, h=8, spp=4, up=(0, 1, 0)):
    return {"type": "perspective", "fov": fov, "to_world": T.look_at(origin=list(origin), target=list(target),
                                                                        up=list(up)),
            "film": {"type": "hdrfilm", "width": w, "height": h, "rfilter": {"type": "box"}},
            "sampler": {"type": "independent", "sample_count": spp}}


def medium_dict(grid, albedo, scale, twostate=False, majorant_factor=0):
    m = {"type": "heterogeneous", "sigma_t": {"type": "gridvolume", "data": grid, "to_world": T.scale(1.0)},
         "albedo": albedo, "scale": scale, "majorant_resolution_factor": majorant_factor}
    return {"type": "twostatemedium", "old": m, "new": m, "incoming": m} if twostate else m


def volume_box_scene(grid, albedo=0.8, scale=5.0, env=1.0, w=8, h=8, spp=4, twostate=True, fov=40.0,
                     origin=(0.5, 0.5, 2.6), majorant_factor=0):
    """Grid medium in a null unit cube [0,1]^3 under a constant environment,
    one camera on +z looking at the cube centre (``majorant_factor`` > 0:
    majorant supergrid of ceil(n / factor) cells per axis)."""
    return {
        "type": "scene",
        "box": {"type": "cube", "bsdf": {"type": "null"}, "to_world": T.translate([0.5, 0.5, 0.5]).scale(0.5),
                "interior": medium_dict(grid, albedo, scale, twostate, majorant_factor)},
        "env": {"type": "constant", "radiance": {"type": "rgb", "value": env}},
        "sensor": cam(origin, (0.5, 0.5, 0.5), fov=fov, w=w, h=h, spp=spp),
    }


def furnace_box_scene(albedo=0.5, radiance=1.0, w=6, h=6, spp=4, textured=False, tex=None):
    """Closed box of six inward-facing emitting diffuse walls; camera inside.
    E[L] = radiance * sum_{k < max_depth} albedo^k."""
    refl = {"type": "rgb", "value": albedo}
    if textured:
        refl = {"type": "bitmap", "data": tex}
    walls = {
        "floor": T.translate([0, -1, 0]).rotate([1, 0, 0], -90),
        "ceiling": T.translate([0, 1, 0]).rotate([1, 0, 0], 90),
        "back": T.translate([0, 0, -1]),
        "front": T.translate([0, 0, 1]).rotate([0, 1, 0], 180),
        "right": T.translate([1, 0, 0]).rotate([0, 1, 0], -90),
        "left": T.translate([-1, 0, 0]).rotate([0, 1, 0], 90),
    }
    d = {"type": "scene", "sensor": cam((0.1, 0.05, 0.3), (0.2, -0.3, -1.0), fov=60, w=w, h=h, spp=spp)}
    for name, tw in walls.items():
        bsdf = {"type": "diffuse", "reflectance": refl}
        if textured and name == "back":
            bsdf = {"type": "twostate", "old": bsdf, "new": bsdf, "incoming": bsdf}
        elif textured:
            bsdf = {"type": "diffuse", "reflectance": {"type": "rgb", "value": albedo}}
        d[name] = {"type": "rectangle", "to_world": tw, "bsdf": bsdf,
                   "emitter": {"type": "area", "radiance": {"type": "rgb", "value": radiance}}}
    return d


def floor_light_scene(w=8, h=8, spp=4, light_y=1.0, light_half=0.2, radiance=10.0, albedo=0.6):
    """Diffuse floor (y = 0, 2x2) under a small downward-facing square light,
    camera looking down at the floor from the side."""
    return {
        "type": "scene",
        "floor": {"type": "rectangle", "to_world": T.rotate([1, 0, 0], -90),
                  "bsdf": {"type": "diffuse", "reflectance": {"type": "rgb", "value": albedo}}},
        "light": {"type": "rectangle",
                  "to_world": T.translate([0, light_y, 0]).rotate([1, 0, 0], 90).scale(light_half),
                  "bsdf": {"type": "diffuse", "reflectance": {"type": "rgb", "value": 0.0}},
                  "emitter": {"type": "area", "radiance": {"type": "rgb", "value": radiance}}},
        "sensor": cam((0.0, 1.5, 1.6), (0.0, 0.0, 0.0), fov=30, w=w, h=h, spp=spp),
    }


def camera_rays(view_to_world, fov, w, h, px, py, jx, jy):
    """Same pinhole convention as the renderers: x = left, y = up, z = fwd;
    film u grows right, v down."""
    m = np.asarray(view_to_world, np.float64)
    tan_half = np.tan(np.radians(fov) * 0.5)
    fx = (px + jx) / w
    fy = (py + jy) / h
    cx = (1 - 2 * fx) * tan_half
    cy = (1 - 2 * fy) * tan_half / (w / h)
    d = m[:3, 0][None] * cx[:, None] + m[:3, 1][None] * cy[:, None] + m[:3, 2][None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.broadcast_to(m[:3, 3], d.shape), d


def box_chord(o, d, lo=(0, 0, 0), hi=(1, 1, 1)):
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0 = (np.asarray(lo) - o) * inv
        t1 = (np.asarray(hi) - o) * inv
    tn = np.max(np.minimum(t0, t1), axis=1)
    tf = np.min(np.maximum(t0, t1), axis=1)
    tn = np.maximum(tn, 0)
    return np.where(tf > tn, tf - tn, 0.0)


def homogeneous_box_scene(sigma_t=1.0, albedo=0.8, scale=2.0, env=1.0, w=8, h=8, spp=4, twostate=True, fov=40.0,
                          origin=(0.5, 0.5, 2.6)):
    """A homogeneous medium (constant sigma_t value x scale and albedo,
    tutorial.ipynb:77-88) in a null unit cube [0,1]^3 under a constant
    environment, one camera on +z looking at the cube centre."""
    m = {"type": "homogeneous", "sigma_t": {"type": "spectrum", "value": sigma_t},
         "albedo": {"type": "rgb", "value": albedo}, "scale": scale}
    interior = {"type": "twostatemedium", "old": m, "new": m, "incoming": m} if twostate else m
    return {
        "type": "scene",
        "box": {"type": "cube", "bsdf": {"type": "null"}, "to_world": T.translate([0.5, 0.5, 0.5]).scale(0.5),
                "interior": interior},
        "env": {"type": "constant", "radiance": {"type": "rgb", "value": env}},
        "sensor": cam(origin, (0.5, 0.5, 0.5), fov=fov, w=w, h=h, spp=spp),
    }


def pixel_chords(view, w, h, n=24):
    """Per pixel, the chords through the unit cube of an n x n grid of
    sub-pixel camera rays, shape (h, w, n * n)."""
    jj = (np.arange(n) + 0.5) / n
    jx, jy = [a.ravel() for a in np.meshgrid(jj, jj)]
    out = np.zeros((h, w, jx.size))
    for py in range(h):
        for px in range(w):
            o, dd = camera_rays(view.to_world, view.fov, w, h, np.full(jx.shape, px), np.full(jx.shape, py), jx, jy)
            out[py, px] = box_chord(o, dd)
    return out
