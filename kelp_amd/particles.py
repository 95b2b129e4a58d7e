"""Particle state (host side), mirroring src/kelp/core/particles.py.

``ParticleStatus`` / ``FailureReason`` codes are the reference's
(core/particles.py:18-73); ``ParticleState`` keeps the reference's field names
(core/particles.py:153) so code written against the reference reads the same.
"""
from __future__ import annotations

from dataclasses import dataclass, fields, replace as dc_replace
from enum import IntEnum
from typing import Optional

import numpy as np


class ParticleStatus(IntEnum):
    ALIVE = 1
    BEACHED = 2
    OUTSIDE = 3
    NOT_RELEASED = 4


class FailureReason(IntEnum):
    NONE = 0
    INTERPOLATION_FAILED = 1
    OUT_OF_BOUNDS = 2
    SURFACE_HIT = 3
    BOTTOM_HIT = 4


@dataclass
class ParticleState:
    lon: np.ndarray
    lat: np.ndarray
    z: np.ndarray
    status: np.ndarray
    nearest_node: np.ndarray
    u: Optional[np.ndarray] = None
    v: Optional[np.ndarray] = None
    w: Optional[np.ndarray] = None
    failure_reason: Optional[np.ndarray] = None
    release_step: Optional[np.ndarray] = None
    elem: Optional[np.ndarray] = None
    release_lon: Optional[np.ndarray] = None
    release_lat: Optional[np.ndarray] = None
    release_z: Optional[np.ndarray] = None

    @property
    def n_particles(self) -> int:
        return int(self.lon.shape[0])

    @property
    def alive_mask(self) -> np.ndarray:
        return self.status == ParticleStatus.ALIVE

    def replace(self, **kw) -> "ParticleState":
        return dc_replace(self, **kw)

    def copy(self) -> "ParticleState":
        kw = {}
        for f in fields(self):
            a = getattr(self, f.name)
            kw[f.name] = None if a is None else np.array(a, copy=True)
        return ParticleState(**kw)


def make_state(lon, lat, z=None) -> ParticleState:
    """core/particles.py:384 make_state."""
    lon = np.asarray(lon, dtype=np.float64)
    lat = np.asarray(lat, dtype=np.float64)
    P = lon.shape[0]
    z = np.zeros(P, dtype=np.float64) if z is None else np.asarray(z, dtype=np.float64)
    return ParticleState(lon=lon, lat=lat, z=z,
                         status=np.full(P, ParticleStatus.ALIVE, dtype=np.int32),
                         nearest_node=np.full(P, -1, dtype=np.int32))


def count_by_status(status: np.ndarray) -> dict:
    """core/particles.py:99 count_by_status."""
    return {s.name.lower(): int(np.count_nonzero(status == s)) for s in ParticleStatus}
