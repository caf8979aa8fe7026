"""Affine 4x4 transforms for scene descriptions (the subset of
``mitsuba.scalar_rgb.Transform4f`` the reference configs use:
``look_at``, ``translate``, ``scale``, ``rotate``; experiments.py:15,45-48).

Chaining follows the reference's convention: ``T.translate(a).scale(b)``
applies ``scale`` first, then ``translate``.
"""
from __future__ import annotations

import math

import numpy as np


class Transform4f:
    def __init__(self, matrix=None):
        self.matrix = np.eye(4) if matrix is None else np.asarray(matrix, dtype=np.float64).reshape(4, 4)

    # -- constructors (usable on the class or chained on an instance) --
    @staticmethod
    def _translate_m(v):
        v = np.broadcast_to(np.asarray(v, np.float64), (3,))
        m = np.eye(4)
        m[:3, 3] = v
        return m

    @staticmethod
    def _scale_m(v):
        v = np.broadcast_to(np.asarray(v, np.float64), (3,))
        return np.diag([v[0], v[1], v[2], 1.0])

    @staticmethod
    def _rotate_m(axis, angle_deg):
        a = np.asarray(axis, np.float64)
        a = a / np.linalg.norm(a)
        t = math.radians(angle_deg)
        c, s = math.cos(t), math.sin(t)
        x, y, z = a
        r = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                      [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                      [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
        m = np.eye(4)
        m[:3, :3] = r
        return m

    @staticmethod
    def _look_at_m(origin, target, up):
        o = np.asarray(origin, np.float64)
        d = np.asarray(target, np.float64) - o
        d = d / np.linalg.norm(d)
        left = np.cross(np.asarray(up, np.float64), d)
        ln = np.linalg.norm(left)
        if ln < 1e-12:
            raise ValueError("look_at: up is parallel to the viewing direction")
        left = left / ln
        new_up = np.cross(d, left)
        m = np.eye(4)
        m[:3, 0] = left
        m[:3, 1] = new_up
        m[:3, 2] = d
        m[:3, 3] = o
        return m

    def _chain(self, m):
        return Transform4f(self.matrix @ m)

    def translate(self, v):
        return self._chain(self._translate_m(v))

    def scale(self, v):
        return self._chain(self._scale_m(v))

    def rotate(self, axis, angle):
        return self._chain(self._rotate_m(axis, angle))

    def look_at(self, origin, target, up):
        return self._chain(self._look_at_m(origin, target, up))

    def __matmul__(self, other):
        return Transform4f(self.matrix @ other.matrix)

    def transform_point(self, p):
        p = np.asarray(p, np.float64)
        return (p @ self.matrix[:3, :3].T) + self.matrix[:3, 3]

    def transform_vector(self, v):
        return np.asarray(v, np.float64) @ self.matrix[:3, :3].T

    def __repr__(self):
        return f"Transform4f({self.matrix.tolist()})"


class _TFactory:
    """``T.translate(...)`` etc. starting from the identity."""

    def __getattr__(self, name):
        return getattr(Transform4f(), name)


T = _TFactory()
