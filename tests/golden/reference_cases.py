"""Inputs of the reference's 3-D golden cases, regenerated without the reference.

Restates the recipe of the reference's golden generator for the node-layout
HEALPix cases (tests/_gen_golden_3d.py:197-300 with the fixture helpers of
tests/test_fused_3d_bary.py:28-70): an all-wet nside-8 NEST dual mesh with
10 z-levels and a banded bathymetry, random-normal node snapshots, 48 seeds,
eight 6-hour steps bracketed by one snapshot pair.  ``run_case`` drives the
oracle over those steps and returns the per-step states in the layout of the
stored ``.npz`` files.
"""
from __future__ import annotations

import numpy as np

from kelp_amd import healpix

NSIDE = 8
NPIX = 12 * NSIDE * NSIDE
Z_LEVELS = np.array([0.0, 10.0, 25.0, 50.0, 100.0, 250.0, 500.0, 1000.0, 2000.0, 4000.0])
K = Z_LEVELS.size
NODE_DT = 21600.0
NODE_STEPS = 8
NODE_P = 48

#: case -> (integrator order, snapshot seeds, diffusion kwargs or None)
CASES = {
    "fesom2_node_rk2_fused_lerp": (2, (1, 2), None),
    "fesom2_node_rk4_unfused": (4, (21, 22), None),
    "split_uniform_diffusion_rk2": (2, (31, 32), dict(kh_x=50.0, kh_y=20.0, seed=2026)),
    # steady single snapshot, vertical Az kick only (tests/_gen_golden_3d.py:305-318)
    "az_const_euler": (1, (41,), dict(kh_x=0.0, kh_y=0.0, seed=99, az_const=5e-3)),
}


def dual_mesh():
    lon, lat = healpix.pix2ang_nest(NSIDE, np.arange(NPIX), lonlat=True)
    depth = 60.0 + 3500.0 * np.abs(np.sin(np.deg2rad(lat * 2.0)))
    depth[::17] = 35.0
    return healpix.dual_mesh(NSIDE, depth=depth, z_levels=Z_LEVELS)


def field(mesh, seed, scale=0.4):
    rng = np.random.default_rng(seed)
    N = mesh.n_nodes
    u = scale * rng.standard_normal((K - 1, N))
    v = scale * rng.standard_normal((K - 1, N))
    w = 1e-4 * rng.standard_normal((K, N))
    return u, v, w


def seeds(n=NODE_P, seed=5):
    rng = np.random.default_rng(seed)
    lon = rng.uniform(-179.0, 179.0, n)
    lat = np.degrees(np.arcsin(rng.uniform(-0.95, 0.95, n)))
    z = rng.uniform(0.0, 3000.0, n)
    z[::7] = 20.0
    z[1::13] = 3900.0
    return lon, lat, z


def payload(step, direction):
    """(use_b, alpha0, target_is_b, dt_to_next) of step ``step``: the
    snapshot pair brackets the whole run, [0, T] (tests/_gen_golden_3d.py:99)."""
    T = NODE_STEPS * NODE_DT
    t0 = 0.0 if direction > 0 else T
    s = t0 + direction * step * NODE_DT
    alpha = s / T
    if direction > 0:
        tgt_b, dtn = True, T - s
    else:
        tgt_b, dtn = False, s
    if abs(alpha) < 1e-12:
        return "a", 0.0, tgt_b, dtn
    if abs(alpha - 1.0) < 1e-12:
        return "b", 0.0, tgt_b, dtn
    return "lerp", alpha, tgt_b, dtn


def run_case(name, direction, stepper):
    """Drive ``stepper(state, fields, order, alphas, step_i, kick)`` (a
    callable returning the new state dict) over the case's steps; returns
    {key: (step, particle) array}."""
    order, snap_seeds, kick = CASES[name]
    mesh = dual_mesh()
    snaps = [field(mesh, s) for s in snap_seeds]
    steady = len(snaps) == 1
    a, b = snaps[0], snaps[-1]
    lon, lat, z = seeds()
    st = dict(lon=lon, lat=lat, z=z, status=np.ones(lon.size, np.int32),
              nn=np.full(lon.size, -1, np.int32), fr=np.zeros(lon.size, np.int8))
    rows = {k: [] for k in ("lon", "lat", "z", "status", "failure_reason", "u", "v", "w")}
    c = {1: (0.0,), 2: (0.0, 0.5), 4: (0.0, 0.5, 0.5, 1.0)}[order]
    for i in range(NODE_STEPS):
        if steady:
            fields = dict(u_a=a[0], v_a=a[1], w_a=a[2], u_t=a[0], v_t=a[1], w_t=a[2], alpha0=0.0)
            st = stepper(mesh, st, fields, order, [0.0] * len(c), i, kick)
            for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("status", "status"),
                           ("failure_reason", "fr"), ("u", "u"), ("v", "v"), ("w", "w")):
                rows[k].append(np.array(st[src], copy=True))
            continue
        kind, alpha0, tgt_b, dtn = payload(i, direction)
        base = {"a": a, "b": b, "lerp": a}[kind]
        tgt = b if tgt_b else a
        fields = dict(u_a=base[0], v_a=base[1], w_a=base[2], u_t=tgt[0], v_t=tgt[1], w_t=tgt[2],
                      alpha0=alpha0)
        if kind == "lerp":
            fields.update(u_b=b[0], v_b=b[1], w_b=b[2])
        alphas = [ck * NODE_DT / dtn for ck in c]
        st = stepper(mesh, st, fields, order, alphas, i, kick)
        for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("status", "status"),
                       ("failure_reason", "fr"), ("u", "u"), ("v", "v"), ("w", "w")):
            rows[k].append(np.array(st[src], copy=True))
    return {k: np.stack(v) for k, v in rows.items()}
