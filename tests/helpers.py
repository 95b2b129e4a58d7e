"""Shared helpers of the test-suite (small meshes, brute-force references)."""
from __future__ import annotations

import numpy as np

from kelp_amd.mesh import make_mesh


def brute_nearest(node_xyz, lon, lat):
    """argmin of the unit-sphere chord over all nodes, the reference's
    brute-force oracle (tests/test_exact_nearest.py:28-35)."""
    lr = np.radians(np.asarray(lon, np.float64))
    la = np.radians(np.asarray(lat, np.float64))
    q = np.stack([np.cos(la) * np.cos(lr), np.cos(la) * np.sin(lr), np.sin(la)], 1)
    out = np.empty(q.shape[0], dtype=np.int64)
    d2min = np.empty(q.shape[0])
    for s in range(0, q.shape[0], 256):
        d2 = ((node_xyz[None, :, :] - q[s:s + 256, None, :]) ** 2).sum(-1)
        out[s:s + 256] = d2.argmin(1)
        d2min[s:s + 256] = d2.min(1)
    return out, d2min


def chord2(node_xyz, nodes, lon, lat):
    lr = np.radians(np.asarray(lon, np.float64))
    la = np.radians(np.asarray(lat, np.float64))
    q = np.stack([np.cos(la) * np.cos(lr), np.cos(la) * np.sin(lr), np.sin(la)], 1)
    return ((node_xyz[nodes] - q) ** 2).sum(-1)


def global_latlon_mesh(step=6.0, jitter=0.0, seed=3):
    """Near-global triangulated lon-lat grid (exercises the wrap grid)."""
    lon = np.arange(-180.0, 180.0, step)
    lat = np.arange(-84.0, 84.0 + 1e-9, step)
    LON, LAT = np.meshgrid(lon, lat)
    rng = np.random.default_rng(seed)
    if jitter:
        LON = LON + rng.uniform(-jitter, jitter, LON.shape)
        LAT = LAT + rng.uniform(-jitter, jitter, LAT.shape)
    nlat, nlon = LON.shape
    nid = np.arange(nlat * nlon).reshape(nlat, nlon)
    a = nid[:-1, :]
    b = np.roll(nid[:-1, :], -1, axis=1)
    c = np.roll(nid[1:, :], -1, axis=1)
    d = nid[1:, :]
    tri = np.concatenate([np.stack([a.ravel(), b.ravel(), c.ravel()], 1),
                          np.stack([a.ravel(), c.ravel(), d.ravel()], 1)]).astype(np.int32)
    return make_mesh("global", LON.ravel(), LAT.ravel(), tri)
