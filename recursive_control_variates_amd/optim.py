"""Image losses and the Adam optimizer on librcv kernels.

* ``l1_loss`` / ``l2_loss`` (utils.py:21-25): one fused kernel computes the
  mean and dLoss/dimg; wrapped as an autograd Function.
* ``Adam`` mirrors the mi.ad.Adam surface used by the reference driver
  (optimize.py:90,107,127; utils.py:62-87,167-175): ``opt[key] = value``,
  ``step()``, ``set_learning_rate``; per-key clamping is fused into the
  update (``set_clamp``). Each step writes a fresh tensor, so scene states
  bound by reference stay intact.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict

import torch

from . import _native as N


class _LossOp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, ref, kind):
        img_c = img.detach().contiguous()
        ref_c = ref.detach().to(img.device, torch.float32).contiguous()
        if img_c.shape != ref_c.shape:
            raise ValueError(f"loss: shapes {tuple(img_c.shape)} and {tuple(ref_c.shape)} differ")
        out = torch.empty(1, dtype=torch.float32, device=img.device)
        grad = torch.empty_like(img_c) if img.requires_grad else None
        N.call("rcv_loss", C.c_void_p(img_c.data_ptr()), C.c_void_p(ref_c.data_ptr()), img_c.numel(), kind,
               C.c_void_p(out.data_ptr()), C.c_void_p(grad.data_ptr()) if grad is not None else None,
               C.c_void_p(N.stream_handle(img.device)))
        ctx.save_for_backward(grad if grad is not None else out)
        ctx.has_grad = grad is not None
        return out[0]

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        if not ctx.has_grad:
            return None, None, None
        return grad * g, None, None


def l1_loss(x, y):
    return _LossOp.apply(x, y, N.LOSS_L1)


def l2_loss(x, y):
    return _LossOp.apply(x, y, N.LOSS_L2)


class Adam:
    def __init__(self, lr=1e-2, beta_1=0.9, beta_2=0.999, epsilon=1e-8):
        if not (0 <= beta_1 < 1 and 0 <= beta_2 < 1):
            raise ValueError("betas must lie in [0, 1)")
        self.lr = float(lr)
        self.beta_1 = float(beta_1)
        self.beta_2 = float(beta_2)
        self.epsilon = float(epsilon)
        self.variables: Dict[str, torch.Tensor] = {}
        self.state: Dict[str, tuple] = {}
        self.t: Dict[str, int] = {}
        self.lr_k: Dict[str, float] = {}
        self.clamp_k: Dict[str, tuple] = {}

    def keys(self):
        return self.variables.keys()

    def items(self):
        return self.variables.items()

    def __contains__(self, key):
        return key in self.variables

    def __getitem__(self, key):
        return self.variables[key]

    def __setitem__(self, key, value):
        if not isinstance(value, torch.Tensor):
            value = torch.as_tensor(value, dtype=torch.float32)
        value = value.detach().float().contiguous()
        old = self.variables.get(key)
        if old is None or old.shape != value.shape or old.device != value.device:
            self.state[key] = (torch.zeros_like(value), torch.zeros_like(value))
            self.t[key] = 0
        self.variables[key] = value.requires_grad_(True)

    def set_learning_rate(self, lr):
        """A dict sets per-key rates; a number sets the default rate only,
        and keys given their own rate keep it (mi.ad.Optimizer semantics:
        a scalar replaces ``lr_default``, per-key entries stay)."""
        if isinstance(lr, dict):
            for k, v in lr.items():
                self.lr_k[k] = float(v)
        else:
            self.lr = float(lr)

    def set_clamp(self, key, lo, hi):
        self.clamp_k[key] = (float(lo), float(hi))

    def step(self):
        for key, p in list(self.variables.items()):
            g = p.grad
            if g is None:
                continue
            g = g.detach().contiguous().float()
            m, v = self.state[key]
            self.t[key] += 1
            lo, hi = self.clamp_k.get(key, (-float("inf"), float("inf")))
            out = torch.empty_like(p)
            N.call("rcv_adam_step", C.c_void_p(out.data_ptr()), C.c_void_p(p.data_ptr()),
                   C.c_void_p(g.data_ptr()), C.c_void_p(m.data_ptr()), C.c_void_p(v.data_ptr()), p.numel(),
                   self.t[key], C.c_float(self.lr_k.get(key, self.lr)), C.c_float(self.beta_1),
                   C.c_float(self.beta_2), C.c_float(self.epsilon), C.c_float(lo), C.c_float(hi),
                   C.c_void_p(N.stream_handle(p.device)))
            self.variables[key] = out.requires_grad_(True)
