"""MI355X-native recursive control variates for inverse rendering.

Plugin registry with the reference's names (plugins/__init__.py:11-17):

=====================  ==============================================
integrator ``cv``       CVIntegrator (plugins/cv_integrator.py)
integrator ``meta``     MetaIntegrator (plugins/metaintegrator.py)
``twostateprb``         TwoStatePRBIntegrator (plugins/twostatepath.py)
``twostateprbvolpath``  TwoStatePRBVolpathIntegrator (plugins/twostatevolpath.py)
``volpathsimple``       VolpathSimpleIntegrator (plugins/volpathsimple.py)
``prb``/``prbvolpath``  plain single-state renders (upstream base classes)
bsdf ``twostate``       handled by the scene loader (plugins/twostatebsdf.py)
medium ``twostatemedium`` handled by the scene loader (plugins/twostatemedium.py)
=====================  ==============================================

All computation runs in librcv.so (HIP, gfx950); there is no CPU path.
"""
from __future__ import annotations

from .transform import T, Transform4f

_INTEGRATORS = {}


def register_integrator(name, ctor):
    _INTEGRATORS[name] = ctor


def _register_defaults():
    from .cv_integrator import CVIntegrator, MetaIntegrator
    from .integrators import (PRBIntegrator, PRBVolpathIntegrator, TwoStatePRBIntegrator,
                              TwoStatePRBVolpathIntegrator, VolpathSimpleIntegrator)
    register_integrator("prb", PRBIntegrator)
    register_integrator("path", PRBIntegrator)
    register_integrator("prbvolpath", PRBVolpathIntegrator)
    register_integrator("volpath", PRBVolpathIntegrator)
    register_integrator("twostateprb", TwoStatePRBIntegrator)
    register_integrator("twostateprbvolpath", TwoStatePRBVolpathIntegrator)
    register_integrator("volpathsimple", VolpathSimpleIntegrator)
    register_integrator("meta", MetaIntegrator)
    register_integrator("cv", CVIntegrator)


_register_defaults()


def load_dict(d, device="cuda"):
    """Builds a scene ('type': 'scene') or an integrator from a dict, like
    mi.load_dict. Nested integrator dicts are loaded recursively."""
    t = d.get("type")
    if t == "scene":
        from .scene import load_scene
        return load_scene(d, device)
    if t in _INTEGRATORS:
        props = {}
        for k, v in d.items():
            if k == "type":
                continue
            if isinstance(v, dict) and v.get("type") in _INTEGRATORS:
                v = load_dict(v, device)
            props[k] = v
        return _INTEGRATORS[t](props)
    raise ValueError(f"load_dict: unknown type {t!r}")


def traverse(scene):
    from .scene import traverse as _traverse
    return _traverse(scene)


def render(scene, params=None, sensor=0, integrator=None, seed=0, seed_grad=0, spp=0, spp_grad=0):
    from .rendering import render as _render
    return _render(scene, params, sensor, integrator, seed, seed_grad, spp, spp_grad)


__all__ = ["T", "Transform4f", "load_dict", "traverse", "render", "register_integrator"]
