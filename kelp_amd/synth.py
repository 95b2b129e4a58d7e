"""Synthetic meshes, snapshots and seed clouds of the tracking workloads.

No model files are read: every workload is generated from a recorded seed.

* ``double_gyre_lattice`` -- jittered 0.25-degree lon-lat lattice over
  0-40E x 30-60N (161 x 121 = 19,481 nodes), each quad split into two
  triangles (38,400 elements), steady float32 double gyre at the nodes,
  10,000 seeds uniform in the interior; RK2, dt 600 s, 1000 steps.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import Mesh, make_mesh

C1_SEED = 12345


def lattice_mesh(lon0=0.0, lon1=40.0, lat0=30.0, lat1=60.0, step=0.25, jitter=0.05,
                 seed=C1_SEED, name="lattice") -> Mesh:
    """Jittered lon-lat lattice, each quad split along its (i,j)-(i+1,j+1)
    diagonal into two counter-clockwise triangles."""
    nlon = int(round((lon1 - lon0) / step)) + 1
    nlat = int(round((lat1 - lat0) / step)) + 1
    lon = lon0 + step * np.arange(nlon)
    lat = lat0 + step * np.arange(nlat)
    LON, LAT = np.meshgrid(lon, lat)                  # (nlat, nlon), row = latitude
    rng = np.random.default_rng(seed)
    LON = LON + rng.uniform(-jitter, jitter, LON.shape)
    LAT = LAT + rng.uniform(-jitter, jitter, LAT.shape)
    nid = np.arange(nlat * nlon, dtype=np.int64).reshape(nlat, nlon)
    a = nid[:-1, :-1].ravel()
    b = nid[:-1, 1:].ravel()
    c = nid[1:, 1:].ravel()
    d = nid[1:, :-1].ravel()
    tri = np.empty((2 * a.size, 3), dtype=np.int32)
    tri[0::2] = np.stack([a, b, c], axis=1)
    tri[1::2] = np.stack([a, c, d], axis=1)
    return make_mesh(name, LON.ravel(), LAT.ravel(), tri)


def double_gyre(mesh: Mesh, U=0.5, lon0=0.0, lon1=40.0, lat0=30.0, lat1=60.0, dtype=np.float32):
    """u = -U sin(pi x) cos(pi y), v = U cos(pi x) sin(pi y) at the nodes,
    x, y the domain-normalised node coordinates."""
    x = (mesh.nodes_lon - lon0) / (lon1 - lon0)
    y = (mesh.nodes_lat - lat0) / (lat1 - lat0)
    u = -U * np.sin(np.pi * x) * np.cos(np.pi * y)
    v = U * np.cos(np.pi * x) * np.sin(np.pi * y)
    return u.astype(dtype), v.astype(dtype)


def uniform_seeds(n, lon0, lon1, lat0, lat1, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(lon0, lon1, n), rng.uniform(lat0, lat1, n)


@dataclass
class Workload2D:
    mesh: Mesh
    u: np.ndarray
    v: np.ndarray
    lon: np.ndarray
    lat: np.ndarray
    order: int
    dt: float
    n_steps: int
    output_every: int


def double_gyre_workload(n_particles=10_000, seed=C1_SEED) -> Workload2D:
    """The 2-D double-gyre workload (RK2, dt 600 s, 1000 steps, output /10)."""
    mesh = lattice_mesh(seed=seed)
    u, v = double_gyre(mesh)
    lon, lat = uniform_seeds(n_particles, 0.5, 39.5, 30.5, 59.5, seed + 1)
    return Workload2D(mesh, u, v, lon, lat, order=2, dt=600.0, n_steps=1000, output_every=10)
