"""Procedural scenes of the hot path's workloads (the reference's own scenes
are downloaded data that is not in the tree, README.md:114-123).

Choices not fixed by the workload definitions, stated here once:

* Cornell box (surfaces): walls are 2x2 rectangles bounding [-1, 1]^3 with
  inward normals (floor, ceiling and back 0.725 grey, left red
  (0.63, 0.065, 0.05), right green (0.14, 0.45, 0.091)); sphere r = 0.3 at
  (0.35, -0.7, 0.2), white 0.8; area light 0.3 x 0.3 at y = 0.99 facing
  down, radiance 15; camera at (0, 0, 3.9) looking at the origin, fov 39.3
  degrees. The back wall carries the optimised 64x64 RGB texture.
* Volumes: the grid spans the unit cube [0, 1]^3 inside a null-BSDF cube
  (experiments.py:74-79); constant environment of radiance 1; views on a
  circle of radius 2 around (0.5, 0.5, 0.5) at azimuths 0/90/180/270
  degrees, height 0.5, fov 45 degrees, side by side on one batch film
  (experiments.py:28-50); rr_depth = max_depth, as the reference's volume
  experiments set it (experiments.py:130).
"""
from __future__ import annotations

import numpy as np

from .transform import T


def blob_grid(res: int, seed: int, n_blobs: int = 8) -> np.ndarray:
    """Sum of isotropic Gaussian blobs sampled at voxel centres: centres
    U[0.2, 0.8]^3, sigma U[0.05, 0.15], amplitude U[0.5, 1] drawn from
    numpy.random.default_rng(seed) in that order per blob. Shape (res,)*3+(1,)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.2, 0.8, size=(n_blobs, 3))
    s = rng.uniform(0.05, 0.15, size=n_blobs)
    a = rng.uniform(0.5, 1.0, size=n_blobs)
    x = (np.arange(res) + 0.5) / res
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    g = np.zeros((res, res, res), np.float64)
    for k in range(n_blobs):
        d2 = (xx - c[k, 0]) ** 2 + (yy - c[k, 1]) ** 2 + (zz - c[k, 2]) ** 2
        g += a[k] * np.exp(-0.5 * d2 / s[k] ** 2)
    return g.astype(np.float32)[..., None]


def albedo_grid(res: int, seed: int) -> np.ndarray:
    """Smooth RGB albedo in [0.5, 0.95]: per channel, a sum of 4 blobs."""
    out = []
    for ch in range(3):
        g = blob_grid(res, seed * 7 + ch, n_blobs=4)[..., 0]
        g = g / max(g.max(), 1e-6)
        out.append(0.5 + 0.45 * g)
    return np.stack(out, axis=-1).astype(np.float32)


def texture_target(res: int = 64, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).uniform(0.1, 0.9, size=(res, res, 3)).astype(np.float32)


def _views(n_views, radius=2.0, center=(0.5, 0.5, 0.5), height=0.5, fov=45.0):
    out = {}
    for i in range(n_views):
        theta = 2.0 * np.pi * i / n_views
        origin = [center[0] + radius * np.sin(theta), height, center[2] + radius * np.cos(theta)]
        out[f"sensor_{i:02d}"] = {"type": "perspective", "fov": fov,
                                  "to_world": T.look_at(origin=origin, target=list(center), up=[0, 1, 0])}
    return out


def volume_scene_dict(grid: np.ndarray, twostate: bool = True, film_res: int = 256, n_views: int = 4,
                      spp: int = 32, scale: float = 20.0, albedo=0.8, albedo_grid_data=None,
                      env_radiance: float = 1.0, majorant_resolution_factor: int = 0):
    """Heterogeneous-volume scene (the 64^3 / 128^3 blob workloads)."""
    sensor = {"type": "batch",
              "film": {"type": "hdrfilm", "width": film_res * n_views, "height": film_res,
                       "rfilter": {"type": "box"}},
              "sampler": {"type": "independent", "sample_count": spp}}
    sensor.update(_views(n_views))
    medium = {"type": "heterogeneous",
              "sigma_t": {"type": "gridvolume", "data": grid, "to_world": T.scale(1.0)},
              "scale": scale, "sample_emitters": True,
              "majorant_resolution_factor": majorant_resolution_factor}
    if albedo_grid_data is not None:
        medium["albedo"] = {"type": "gridvolume", "data": albedo_grid_data, "to_world": T.scale(1.0)}
    else:
        medium["albedo"] = albedo
    interior = ({"type": "twostatemedium", "old": medium, "new": medium, "incoming": medium}
                if twostate else medium)
    return {
        "type": "scene",
        "object": {"type": "cube", "bsdf": {"type": "null"},
                   "to_world": T.translate([0.5, 0.5, 0.5]).scale(0.5), "interior": interior},
        "envmap": {"type": "constant", "radiance": {"type": "rgb", "value": env_radiance}},
        "sensor": sensor,
    }


def cornell_scene_dict(texture: np.ndarray, twostate: bool = True, film_res: int = 128, spp: int = 16):
    """Cornell box with a textured back wall (the 64x64 texture workload)."""
    white = {"type": "diffuse", "reflectance": {"type": "rgb", "value": [0.725, 0.71, 0.68]}}
    back = {"type": "diffuse", "reflectance": {"type": "bitmap", "data": texture}}
    back_bsdf = {"type": "twostate", "old": back, "new": back, "incoming": back} if twostate else back
    return {
        "type": "scene",
        "floor": {"type": "rectangle", "to_world": T.translate([0, -1, 0]).rotate([1, 0, 0], -90), "bsdf": white},
        "ceiling": {"type": "rectangle", "to_world": T.translate([0, 1, 0]).rotate([1, 0, 0], 90), "bsdf": white},
        "back": {"type": "rectangle", "to_world": T.translate([0, 0, -1]), "bsdf": back_bsdf},
        "green-wall": {"type": "rectangle", "to_world": T.translate([1, 0, 0]).rotate([0, 1, 0], -90),
                       "bsdf": {"type": "diffuse", "reflectance": {"type": "rgb", "value": [0.14, 0.45, 0.091]}}},
        "red-wall": {"type": "rectangle", "to_world": T.translate([-1, 0, 0]).rotate([0, 1, 0], 90),
                     "bsdf": {"type": "diffuse", "reflectance": {"type": "rgb", "value": [0.63, 0.065, 0.05]}}},
        "sphere": {"type": "sphere", "center": [0.35, -0.7, 0.2], "radius": 0.3,
                   "bsdf": {"type": "diffuse", "reflectance": {"type": "rgb", "value": 0.8}}},
        "light": {"type": "rectangle",
                  "to_world": T.translate([0, 0.99, 0]).rotate([1, 0, 0], 90).scale(0.15),
                  "bsdf": {"type": "diffuse", "reflectance": {"type": "rgb", "value": 0.0}},
                  "emitter": {"type": "area", "radiance": {"type": "rgb", "value": 15.0}}},
        "sensor": {"type": "perspective", "fov": 39.3077,
                   "to_world": T.look_at(origin=[0, 0, 3.9], target=[0, 0, 0], up=[0, 1, 0]),
                   "film": {"type": "hdrfilm", "width": film_res, "height": film_res, "rfilter": {"type": "box"}},
                   "sampler": {"type": "independent", "sample_count": spp}},
    }


def homogeneous_cornell_scene_dict(sigma_t: float = 2.5, twostate: bool = True, film_res: int = 128, spp: int = 2,
                                   albedo=(0.85, 0.97, 0.999999), scale: float = 5.0):
    """The tutorial's scene (tutorial.ipynb:60-95, 310-327): a Cornell box
    (without the two inner boxes) holding an object whose interior is the
    homogeneous medium being optimised (scale 5, sigma_t 2.5, albedo (0.85,
    0.97, 0.999999)), wrapped in the two-state adapter for the cv integrator.
    The tutorial's object is the Stanford bunny (a .ply file not in the tree)
    with a rough-dielectric boundary; here it is a sphere of radius 0.45 at
    (0, -0.55, 0) with a null boundary. Key of the optimised value:
    ``bunny.interior_medium.sigma_t.value.value`` (tutorial.ipynb:181)."""
    d = cornell_scene_dict(np.full((4, 4, 3), 0.725, np.float32), twostate=False, film_res=film_res, spp=spp)
    d["back"]["bsdf"] = {"type": "diffuse", "reflectance": {"type": "rgb", "value": [0.725, 0.71, 0.68]}}
    del d["sphere"]
    medium = {"type": "homogeneous", "scale": scale, "sigma_t": {"type": "spectrum", "value": sigma_t},
              "albedo": {"type": "rgb", "value": tuple(albedo)}}
    d["bunny"] = {"type": "sphere", "center": [0.0, -0.55, 0.0], "radius": 0.45, "bsdf": {"type": "null"},
                  "interior": ({"type": "twostatemedium", "old": medium, "new": medium, "incoming": medium}
                               if twostate else medium)}
    return d


VOLUME_KEY = "object.interior_medium.sigma_t.data"
HOMOGENEOUS_KEY = "bunny.interior_medium.sigma_t.value.value"
ALBEDO_KEY = "object.interior_medium.albedo.data"
TEXTURE_KEY = "back.bsdf.reflectance.data"


# ------------------------------------------------------------ job mixes
def _side(rng):
    return int(rng.integers(2, 17)) * 32          # U{64, 96, ..., 512}


def volume_job_params(seed: int):
    """Size parameters of one volume job of the batched / served mix: single
    view of side^2 (side U{64..512} step 32), grid side U{32, 64, 128}, spp
    U{4..64}, max_depth U{8..32}, drawn from default_rng(seed + 7919)."""
    rng = np.random.default_rng(seed + 7919)
    side = _side(rng)
    res = int(rng.choice([32, 64, 128]))
    spp = int(rng.integers(4, 65))
    depth = int(rng.integers(8, 33))
    return dict(kind="volume", side=side, grid=res, spp=spp, max_depth=depth, key=VOLUME_KEY)


def surface_job_params(seed: int):
    """Size parameters of one surface job: side^2 film, spp U{4..64},
    max_depth 4, drawn from default_rng(seed + 104729)."""
    rng = np.random.default_rng(seed + 104729)
    side = _side(rng)
    spp = int(rng.integers(4, 65))
    return dict(kind="surface", side=side, grid=64, spp=spp, max_depth=4, key=TEXTURE_KEY)


def volume_job_dict(seed: int, twostate=True):
    """Scene and description of one volume job (blob target from ``seed``,
    sigma_t scale 20, albedo 0.8, initial grid 0.1)."""
    m = volume_job_params(seed)
    m["target"] = blob_grid(m["grid"], seed)
    m["init"] = np.full_like(m["target"], 0.1)
    d = volume_scene_dict(m["init"], twostate=twostate, film_res=m["side"], n_views=1, spp=m["spp"])
    return d, m


def surface_job_dict(seed: int, twostate=True):
    """Scene and description of one surface job: the textured Cornell box with
    a 64x64 texture target from ``seed``, initial texture 0.5."""
    m = surface_job_params(seed)
    m["target"] = texture_target(64, seed)
    m["init"] = np.full_like(m["target"], 0.5)
    d = cornell_scene_dict(m["init"], twostate=twostate, film_res=m["side"], spp=m["spp"])
    return d, m
