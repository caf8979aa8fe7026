"""``render()``: the mi.render contract as a torch.autograd.Function.

Forward evaluates ``integrator.render(scene, sensor, seed, spp)`` without
gradients; backward calls the integrator's adjoint with the image gradient,
``seed_grad`` and ``spp_grad`` (contract restated at
plugins/twostatepath.py:22-57,62-95): ``spp_grad`` defaults to ``spp`` and
``seed_grad`` to the first output of TEA(seed, 1) so primal and adjoint use
decorrelated random numbers.
"""
from __future__ import annotations

import torch

from .scene import SceneParameters


def _tea(v0: int, v1: int, rounds: int = 4):
    """TEA (Wheeler & Needham 1994) with the seed-scrambling keys used for the
    per-lane generators; returns (v0, v1)."""
    m = 0xFFFFFFFF
    s = 0
    v0 &= m
    v1 &= m
    for _ in range(rounds):
        s = (s + 0x9E3779B9) & m
        v0 = (v0 + ((((v1 << 4) + 0xA341316C) & m) ^ ((v1 + s) & m) ^ (((v1 >> 5) + 0xC8013EA4) & m))) & m
        v1 = (v1 + ((((v0 << 4) + 0xAD90777D) & m) ^ ((v0 + s) & m) ^ (((v0 >> 5) + 0x7E95761E) & m))) & m
    return v0, v1


def sample_tea_32(v0: int, v1: int, rounds: int = 4):
    return _tea(v0, v1, rounds)


class _RenderOp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scene, params, integrator, sensor, seed, seed_grad, spp, spp_grad, *diff):
        ctx.scene = scene
        ctx.params = params
        ctx.integrator = integrator
        ctx.sensor = sensor
        ctx.seed_grad = seed_grad
        ctx.spp_grad = spp_grad
        with torch.no_grad():
            img = integrator.render(scene, sensor, seed, spp)
        return img

    @staticmethod
    def backward(ctx, grad_out):
        params = ctx.params
        if params is None:
            raise RuntimeError("render(): pass scene parameters to differentiate")
        grads = ctx.integrator.adjoint(ctx.scene, params, grad_out.contiguous(), ctx.sensor, ctx.seed_grad,
                                       ctx.spp_grad)
        keys = [k for k, _ in params.differentiable()]
        return (None,) * 8 + tuple(grads[k] for k in keys)


def render(scene, params: SceneParameters = None, sensor: int = 0, integrator=None, seed: int = 0,
           seed_grad: int = 0, spp: int = 0, spp_grad: int = 0) -> torch.Tensor:
    """Renders ``scene`` with ``integrator``; differentiable w.r.t. the
    parameters assigned to ``params`` with ``requires_grad``."""
    if integrator is None:
        raise ValueError("render(): an integrator is required")
    if params is not None and not isinstance(params, SceneParameters):
        raise TypeError("params should be the result of traverse(scene)")
    if not isinstance(sensor, int):
        raise TypeError("sensor must be an int index")
    if spp == 0:
        spp = scene.sensors()[sensor].spp
    if spp_grad == 0:
        spp_grad = spp
    if seed_grad == 0:
        seed_grad = sample_tea_32(seed, 1)[0]
    elif seed_grad == seed:
        raise ValueError("The primal and differential seed should be different "
                         "to ensure unbiased gradient computation!")
    diff = [t for _, t in params.differentiable()] if params is not None else []
    return _RenderOp.apply(scene, params, integrator, sensor, seed, seed_grad, spp, spp_grad, *diff)
