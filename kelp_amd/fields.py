"""Velocity snapshots on the host, mirroring src/kelp/core/io.py.

``VelocityField`` keeps the reference's field names and shapes
(core/io.py:180): u, v node or element values, ``(N,)`` in 2-D or
``(K-1, N)`` per layer in 3-D; w at the K interfaces ``(K, N)``, positive
down.  ``LazyLerpField`` is the reference's bracket view (core/io.py:386):
the pair ``(vel_a, vel_b)`` and the fraction ``alpha`` of the way from a to
b.  The device path never materialises the blend: the step kernel blends the
two resident snapshots per vertex, with the same arithmetic.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np


def _blend(a, b, alpha):
    """``(1 - alpha) * a + alpha * b`` (core/io.py:68 _blend)."""
    beta = 1.0 - alpha
    return beta * np.asarray(a, dtype=np.float64) + alpha * np.asarray(b, dtype=np.float64)


@dataclass
class VelocityField:
    u: np.ndarray
    v: np.ndarray
    eta: Optional[np.ndarray] = None
    w: Optional[np.ndarray] = None
    t: Optional[float] = None
    kh_zonal: Optional[np.ndarray] = None
    kh_meridional: Optional[np.ndarray] = None
    kh_layout: Optional[str] = None
    uv_layout: str = "element"
    scalars: Optional[Dict[str, np.ndarray]] = None
    az: Optional[np.ndarray] = None
    uv_vertical: str = "layer"

    def __post_init__(self) -> None:
        if self.uv_layout not in ("element", "node"):
            raise ValueError(f"VelocityField: unknown uv_layout {self.uv_layout!r}")
        u = np.asarray(self.u)
        if np.shape(self.v) != u.shape:
            raise ValueError(f"VelocityField: u {u.shape} and v {np.shape(self.v)} differ")
        if self.w is not None and u.ndim == 2 and np.ndim(self.w) == 2:
            if np.shape(self.w)[0] != u.shape[0] + 1:
                raise ValueError(
                    f"VelocityField: w has {np.shape(self.w)[0]} interfaces for {u.shape[0]} layers")

    @property
    def is_3d(self) -> bool:
        return self.w is not None

    def lerp(self, other: "VelocityField", alpha: float) -> "VelocityField":
        """Eager blend (core/io.py:289)."""
        if self.uv_layout != other.uv_layout:
            raise ValueError("VelocityField.lerp: uv_layout mismatch "
                             f"{self.uv_layout!r} vs {other.uv_layout!r}")
        w = None if self.w is None or other.w is None else _blend(self.w, other.w, alpha)
        eta = None if self.eta is None or other.eta is None else _blend(self.eta, other.eta, alpha)
        return VelocityField(u=_blend(self.u, other.u, alpha), v=_blend(self.v, other.v, alpha),
                             eta=eta, w=w, uv_layout=self.uv_layout, uv_vertical=self.uv_vertical)


class LazyLerpField:
    """The bracket ``(vel_a, vel_b, alpha)`` (core/io.py:386).  ``u``, ``v``
    and ``w`` blend on first access (host arrays, for callers that read
    them); the device step reads the two snapshots instead."""

    def __init__(self, vel_a: VelocityField, vel_b: VelocityField, alpha: float) -> None:
        if vel_a.uv_layout != vel_b.uv_layout:
            raise ValueError("LazyLerpField: uv_layout mismatch "
                             f"{vel_a.uv_layout!r} vs {vel_b.uv_layout!r}")
        for name in ("u", "v"):
            if np.shape(getattr(vel_a, name)) != np.shape(getattr(vel_b, name)):
                raise ValueError(f"LazyLerpField: {name} shape mismatch")
        if (vel_a.w is None) != (vel_b.w is None):
            raise ValueError("LazyLerpField: w present on one side only")
        if vel_a.w is not None and np.shape(vel_a.w) != np.shape(vel_b.w):
            raise ValueError("LazyLerpField: w shape mismatch")
        self.vel_a = vel_a
        self.vel_b = vel_b
        self.alpha = float(alpha)
        self.uv_layout = vel_a.uv_layout
        self.uv_vertical = vel_a.uv_vertical
        self.t = None
        self._cache: Dict[str, np.ndarray] = {}

    def _get(self, name):
        if name not in self._cache:
            a, b = getattr(self.vel_a, name), getattr(self.vel_b, name)
            self._cache[name] = None if a is None or b is None else _blend(a, b, self.alpha)
        return self._cache[name]

    @property
    def u(self):
        return self._get("u")

    @property
    def v(self):
        return self._get("v")

    @property
    def w(self):
        if self.vel_a.w is None:
            return None
        return self._get("w")

    @property
    def eta(self):
        return self._get("eta")

    @property
    def is_3d(self) -> bool:
        return self.vel_a.w is not None

    def materialize(self) -> VelocityField:
        return VelocityField(u=self.u, v=self.v, eta=self.eta, w=self.w,
                             uv_layout=self.uv_layout, uv_vertical=self.uv_vertical)
