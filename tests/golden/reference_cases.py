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


def run_pchip_case(direction, stepper):
    """tests/_gen_golden_3d.py:305 (fesom2_node_rk2_fused_pchip): four node
    snapshots at 0, T, 2T, 3T (T = the run's span), the run inside the
    central interval; every RK2 stage reads the monotone-cubic field at
    tq = q + direction c dt (temporal_reconstruct.py:298).  Drives
    ``stepper(mesh, state, snaps, times, tqs, step_i)``."""
    mesh = dual_mesh()
    snaps = [field(mesh, s) for s in (11, 12, 13, 14)]
    T = NODE_STEPS * NODE_DT
    times = [0.0, T, 2.0 * T, 3.0 * T]
    t0 = T if direction > 0 else 2.0 * T
    lon, lat, z = seeds()
    st = dict(lon=lon, lat=lat, z=z, status=np.ones(lon.size, np.int32),
              nn=np.full(lon.size, -1, np.int32), fr=np.zeros(lon.size, np.int8))
    rows = {k: [] for k in ("lon", "lat", "z", "status", "failure_reason", "u", "v", "w")}
    for i in range(NODE_STEPS):
        q = t0 + direction * i * NODE_DT
        tqs = [float(q) + float(direction) * c * NODE_DT for c in (0.0, 0.5)]
        st = stepper(mesh, st, snaps, times, tqs, i)
        for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("status", "status"),
                       ("failure_reason", "fr"), ("u", "u"), ("v", "v"), ("w", "w")):
            rows[k].append(np.array(st[src], copy=True))
    return {k: np.stack(v) for k, v in rows.items()}


# ---------------------------------------------------------------------------
# element-layout cases: leastsq u/v at element centres, leastsq W
# (tests/_gen_golden_3d.py:154-219 with tests/test_3d_analytical.py:39-101)
# ---------------------------------------------------------------------------

ELEM_LEVELS = np.array([0.0, 10.0, 25.0, 50.0, 100.0, 200.0])
ELEM_DT = 3600.0
ELEM_STEPS = 8

#: case -> (integrator order, snapshot shifts)
ELEM_CASES = {
    "fesom2_elem_euler_fused": (1, (0.0,)),
    "fesom2_elem_rk2_unfused_lerp": (2, (0.0, 1.0)),
}


def elem_mesh():
    """A 5 x 5 node lattice over 4-6 E, 4-6 N, two triangles per cell
    ((n0, n1, n2), (n1, n3, n2)), element centres the plain vertex means,
    bathymetry 60 m (west) .. 200 m (east) with a 35 m centre node."""
    from kelp_amd.mesh import make_mesh
    xs = np.linspace(4.0, 6.0, 5)
    ys = np.linspace(4.0, 6.0, 5)
    gx, gy = np.meshgrid(xs, ys)
    lon, lat = gx.ravel(), gy.ravel()
    elems = []
    for j in range(4):
        for i in range(4):
            n0 = j * 5 + i
            elems += [[n0, n0 + 1, n0 + 5], [n0 + 1, n0 + 6, n0 + 5]]
    en = np.array(elems, dtype=np.int32)
    depth = 60.0 + 70.0 * (lon - 4.0)
    depth[12] = 35.0
    mesh = make_mesh("test_local", lon, lat, en, depth=depth, z_levels=ELEM_LEVELS)
    mesh.elem_center_lon = np.array([lon[e].mean() for e in en])
    mesh.elem_center_lat = np.array([lat[e].mean() for e in en])
    return mesh


def elem_field(mesh, s):
    """Element u/v (K-1, E) varying with layer and position, node w (K, N)
    positive down, zero at the surface; ``s`` shifts the snapshot."""
    k = np.arange(mesh.sigma.size - 1, dtype=np.float64)[:, None]
    x = (mesh.elem_center_lon - 5.0)[None, :]
    y = (mesh.elem_center_lat - 5.0)[None, :]
    u = (0.35 - 0.04 * k) * (1.0 + 0.4 * x) + 0.15 * y + 0.05 * s
    v = 0.12 * (1.0 + 0.2 * k) * np.sin(2.0 * x) - 0.08 * y * (1.0 + s)
    zk = (mesh.sigma / mesh.sigma[-1])[:, None]
    xn = (mesh.nodes_lon - 5.0)[None, :]
    w = 1.5e-3 * np.sin(np.pi * zk) * (1.0 + 0.5 * xn) * (1.0 - 0.3 * s)
    return (np.ascontiguousarray(u, np.float64), np.ascontiguousarray(v, np.float64),
            np.ascontiguousarray(w, np.float64))


def elem_seeds():
    i = np.arange(32, dtype=np.float64)
    lon = 4.1 + 1.85 * ((i * 0.618034) % 1.0)
    lat = 4.1 + 1.8 * ((i * 0.381966 + 0.13) % 1.0)
    lon[-4:] = [5.96, 5.975, 5.99, 5.998]
    depths = np.array([0.0, 5.0, 10.0, 17.5, 25.0, 37.0, 50.0, 75.0, 100.0, 150.0, 199.0, 260.0])
    z = depths[np.arange(32) % depths.size]
    return lon, lat, z


def run_elem_case(name, direction, stepper):
    """Drive ``stepper(mesh, state, fields, order, alphas, step_i)`` over
    the element case's steps; ``fields`` carries uv_layout="element"."""
    order, shifts = ELEM_CASES[name]
    mesh = elem_mesh()
    snaps = [elem_field(mesh, s) for s in shifts]
    a, b = snaps[0], snaps[-1]
    lon, lat, z = elem_seeds()
    st = dict(lon=lon, lat=lat, z=z, status=np.ones(lon.size, np.int32),
              nn=np.full(lon.size, -1, np.int32), fr=np.zeros(lon.size, np.int8))
    rows = {k: [] for k in ("lon", "lat", "z", "status", "failure_reason", "u", "v", "w")}
    c = {1: (0.0,), 2: (0.0, 0.5), 4: (0.0, 0.5, 0.5, 1.0)}[order]
    T = ELEM_STEPS * ELEM_DT
    t0 = 0.0 if direction > 0 else T
    for i in range(ELEM_STEPS):
        if len(snaps) == 1:
            fields = dict(u_a=a[0], v_a=a[1], w_a=a[2], alpha0=0.0)
            alphas = [0.0] * len(c)
        else:
            s = t0 + direction * i * ELEM_DT
            alpha = s / T
            if abs(alpha) < 1e-12:
                base, alpha0, lerp = a, 0.0, False
            elif abs(alpha - 1.0) < 1e-12:
                base, alpha0, lerp = b, 0.0, False
            else:
                base, alpha0, lerp = a, alpha, True
            tgt, dtn = (b, T - s) if direction > 0 else (a, s)
            fields = dict(u_a=base[0], v_a=base[1], w_a=base[2], u_t=tgt[0], v_t=tgt[1],
                          w_t=tgt[2], alpha0=alpha0)
            if lerp:
                fields.update(u_b=b[0], v_b=b[1], w_b=b[2])
            alphas = [ck * ELEM_DT / dtn for ck in c]
        fields["uv_layout"] = "element"
        st = stepper(mesh, st, fields, order, alphas, i)
        for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("status", "status"),
                       ("failure_reason", "fr"), ("u", "u"), ("v", "v"), ("w", "w")):
            rows[k].append(np.array(st[src], copy=True))
    return {k: np.stack(v) for k, v in rows.items()}


# ---------------------------------------------------------------------------
# run-level cases: run_single_realization on an in-memory FESOM2 node dataset
# (tests/_gen_golden_3d.py:368-470)
# ---------------------------------------------------------------------------

RUN_LEVELS = np.array([0.0, 10.0, 25.0, 50.0, 100.0])
RUN_FILE_DT = 3600.0
RUN_DT = 1800.0


def run_mesh(depth_fn):
    """The 5 x 5 lattice of the element cases with levels 0-100 m and the
    given bathymetry."""
    from kelp_amd.mesh import make_mesh
    xs = np.linspace(4.0, 6.0, 5)
    gx, gy = np.meshgrid(xs, xs)
    lon, lat = gx.ravel(), gy.ravel()
    elems = []
    for j in range(4):
        for i in range(4):
            n0 = j * 5 + i
            elems += [[n0, n0 + 1, n0 + 5], [n0 + 1, n0 + 6, n0 + 5]]
    en = np.array(elems, dtype=np.int32)
    mesh = make_mesh("test_local", lon, lat, en, depth=depth_fn(lon, lat), z_levels=RUN_LEVELS)
    mesh.elem_center_lon = np.array([lon[e].mean() for e in en])
    mesh.elem_center_lat = np.array([lat[e].mean() for e in en])
    return mesh


def fesom2_node_ds(mesh, n_time, eta, scalars=False):
    """FESOM2-shaped node dataset: unod/vnod (time, nz1, nod2), w (time, nz,
    nod2; file convention positive up), ssh (time, nod2)."""
    from kelp_amd.readers import ArrayDataset
    K, N = mesh.sigma.size, mesh.n_nodes
    t = np.arange(n_time, dtype=np.float64)[:, None, None]
    k = np.arange(K - 1, dtype=np.float64)[None, :, None]
    x = (mesh.nodes_lon - 5.0)[None, None, :]
    y = (mesh.nodes_lat - 5.0)[None, None, :]
    zk = (mesh.sigma / mesh.sigma[-1])[None, :, None]
    unod = (0.25 - 0.03 * k) * (1.0 + 0.3 * x) + 0.05 * y + 0.02 * np.sin(t)
    vnod = 0.1 * (1.0 + 0.2 * k) * np.cos(2.0 * x) - 0.03 * y * (1.0 + 0.1 * t)
    w = 1e-3 * np.sin(np.pi * zk) * (1.0 + 0.4 * x) * np.cos(0.5 * t)
    shp_u, shp_w = (n_time, K - 1, N), (n_time, K, N)
    c = lambda a: np.ascontiguousarray(np.array(a, dtype=np.float64))
    dvars = {
        "unod": (("time", "nz1", "nod2"), c(np.broadcast_to(unod, shp_u))),
        "vnod": (("time", "nz1", "nod2"), c(np.broadcast_to(vnod, shp_u))),
        "w": (("time", "nz", "nod2"), c(np.broadcast_to(w, shp_w))),
        "ssh": (("time", "nod2"), c(np.broadcast_to(np.asarray(eta)[:, None], (n_time, N)))),
    }
    if scalars:
        z_ifc = mesh.sigma[None, :, None]
        z_mid = (0.5 * (mesh.sigma[:-1] + mesh.sigma[1:]))[None, :, None]
        s_lev = 10.0 + 0.1 * z_ifc + 2.0 * x - y + 0.5 * t
        s_mid = 35.0 - 0.02 * z_mid + x * y + 0.1 * t
        dvars["s_lev"] = (("time", "nz", "nod2"), c(np.broadcast_to(s_lev, shp_w)))
        dvars["s_mid"] = (("time", "nz1", "nod2"), c(np.broadcast_to(s_mid, shp_u)))
    return ArrayDataset(dvars, time=np.arange(n_time) * RUN_FILE_DT)


def beaching_case(direction):
    """tests/_gen_golden_3d.py:433: a shallow strip falls dry on a falling
    tide and rewets; depth + stochastic beaching with use_eta,
    auto-or-stochastic resuspension, node RK2 with a time blend."""
    from kelp_amd.config import BeachingConfig, ModeConfig, NumericsConfig
    from kelp_amd.particles import make_state

    def depth(lon, lat):
        d = np.full(lon.shape, 90.0)
        d[(lon == 5.5) & (lat >= 4.5) & (lat <= 5.5)] = 1.0
        return d
    mesh = run_mesh(depth)
    eta = [0.0, -0.45, -0.85, -0.85, -0.45, 0.0, 0.3, 0.0]
    ds = fesom2_node_ds(mesh, len(eta), eta)
    lon = np.concatenate([np.repeat([5.22, 5.3, 5.45, 5.6], 3),
                          np.repeat([4.3, 4.7], 4), [5.97, 5.985] * 2])
    lat = np.concatenate([np.tile([4.6, 5.0, 5.4], 4),
                          np.tile([4.3, 4.8, 5.2, 5.7], 2), [4.4, 4.4, 5.6, 5.6]])
    z = np.concatenate([np.tile([0.2, 0.5, 0.8], 4),
                        np.tile([5.0, 30.0, 60.0, 85.0], 2), [20.0] * 4])
    beach = BeachingConfig(enabled=True, criteria=["depth", "stochastic"], depth_min=0.3,
                           use_eta=True, t_beach=10800.0, resuspension="auto_or_stochastic",
                           t_resuspend=7200.0, min_wet_steps=1, seed=4242)
    numerics = NumericsConfig(dt=RUN_DT, integrator="rk2", uv_interp="barycentric",
                              w_interp="barycentric", temporal_interp=True, beaching=beach)
    mode = ModeConfig(Fdrift2D=False, Fdrift3D=True, time_dir=direction, var_u="unod",
                      var_v="vnod", var_w="w", var_eta="ssh")
    n_time = len(eta)
    return dict(mesh=mesh, ds=ds, init=make_state(lon, lat, z), mode=mode, numerics=numerics,
                start_time_idx=0 if direction > 0 else n_time - 1, n_steps=14,
                n_time_total=n_time, file_times=np.arange(n_time) * RUN_FILE_DT,
                file_dt=RUN_FILE_DT)


def sampling_case(direction):
    """tests/_gen_golden_3d.py:463: record-point sampling of an interface
    (K, N) and a mid-layer (K-1, N) scalar along node RK2 trajectories
    (leastsq W: the unfused barycentric path)."""
    from kelp_amd.config import ModeConfig, NumericsConfig
    from kelp_amd.particles import make_state
    mesh = run_mesh(lambda lon, lat: 60.0 + 20.0 * (lon - 4.0))
    ds = fesom2_node_ds(mesh, 5, np.zeros(5), scalars=True)
    i = np.arange(24, dtype=np.float64)
    lon = 4.15 + 1.7 * ((i * 0.618034) % 1.0)
    lat = 4.15 + 1.7 * ((i * 0.381966 + 0.29) % 1.0)
    z = np.array([0.0, 4.0, 10.0, 17.5, 25.0, 33.0, 50.0, 64.0, 79.0, 99.0, 12.5, 40.0] * 2)
    numerics = NumericsConfig(dt=RUN_DT, integrator="rk2", uv_interp="barycentric",
                              w_interp="leastsq", temporal_interp=True)
    mode = ModeConfig(Fdrift2D=False, Fdrift3D=True, time_dir=direction, var_u="unod",
                      var_v="vnod", var_w="w", var_eta="ssh")
    return dict(mesh=mesh, ds=ds, init=make_state(lon, lat, z), mode=mode, numerics=numerics,
                start_time_idx=0 if direction > 0 else 4, n_steps=8, n_time_total=5,
                file_times=np.arange(5) * RUN_FILE_DT, file_dt=RUN_FILE_DT,
                sample_vars=("s_lev", "s_mid"))


def sigma_case(direction):
    """tests/_gen_golden_3d.py:527 (fesomc_sigma_rk2): FESOM-C sigma
    coordinates on the 3 x 3 node grid of tests/test_fesomc_w_contract.py:49
    -- stretched sigma 1 - linspace(0, 1, 11)^2, a sloping bottom, a tide, a
    dia-sigma flux and layered element u/v; RK2 with a linear time blend,
    W and the column height by least squares."""
    from kelp_amd.config import ModeConfig, NumericsConfig
    from kelp_amd.particles import make_state
    from kelp_amd.readers import ArrayDataset, FesomCArrayReader
    xs, ys = np.array([4.0, 5.0, 6.0]), np.array([-1.0, 0.0, 1.0])
    gx, gy = np.meshgrid(xs, ys)
    lon_n, lat_n = gx.ravel(), gy.ravel()
    tris = []
    for j in range(2):
        for i in range(2):
            n0 = j * 3 + i
            tris += [[n0, n0 + 1, n0 + 3], [n0 + 1, n0 + 4, n0 + 3]]
    tris = np.array(tris)
    nv = np.column_stack([tris + 1, tris[:, :1] + 1]).astype(np.int32)
    sig = 1.0 - np.linspace(0.0, 1.0, 11) ** 2
    K, E = sig.size, tris.shape[0]
    n_time = 7
    times = np.arange(n_time) * 900.0
    tide = 0.8 * np.sin(2.0 * np.pi * times / 5400.0)
    eta = tide[:, None] * (1.0 + 0.1 * (lon_n - 5.0))[None, :]
    depth = 12.0 + 6.0 * (lat_n + 1.0) + 2.0 * (lon_n - 4.0)
    H = depth[None, :] + eta
    wvel = 2e-5 * H[:, None, :] * np.sin(np.pi * (1.0 - sig))[None, :, None]
    wvel[:, 0, :] = 0.0
    wvel[:, -1, :] = 0.0
    kk = ((np.arange(K - 1) + 0.5) / (K - 1))[None, None, :]
    lonc = lon_n[tris].mean(1)[None, :, None]
    latc = lat_n[tris].mean(1)[None, :, None]
    ph = 2.0 * np.pi * times[:, None, None] / 5400.0
    u_e = 0.25 * (1.0 - 0.5 * kk) * (1.0 + 0.3 * (lonc - 5.0)) + 0.05 * np.cos(ph)
    v_e = 0.15 * (1.0 + 0.4 * kk) * (1.0 + 0.2 * latc) - 0.04 * np.sin(ph)
    shp = (n_time, E, K - 1)
    ds = ArrayDataset({
        "lon": (["node"], lon_n), "lat": (["node"], lat_n), "depth": (["node"], depth),
        "nv": (["nele", "four"], nv),
        "lon_elem": (["nele"], lon_n[tris].mean(1)), "lat_elem": (["nele"], lat_n[tris].mean(1)),
        "sigma_lev": (["sigma"], sig),
        "u": (["time", "nele", "sigmam1"], np.ascontiguousarray(np.broadcast_to(u_e, shp))),
        "v": (["time", "nele", "sigmam1"], np.ascontiguousarray(np.broadcast_to(v_e, shp))),
        "w": (["time", "node", "sigma"], np.ascontiguousarray(np.transpose(wvel, (0, 2, 1)))),
        "eta": (["time", "node"], eta),
    }, time=times)
    mesh = FesomCArrayReader.load_mesh(ds)
    lon = np.concatenate([np.repeat([4.6, 5.0, 5.4], 3), [5.996, 5.998]])
    lat = np.concatenate([np.tile([-0.4, 0.0, 0.4], 3), [0.3, -0.4]])
    frac = np.concatenate([np.tile([0.1, 0.4, 0.8], 3), [0.5, 0.95]])
    h0 = 12.0 + 6.0 * (lat + 1.0) + 2.0 * (lon - 4.0)
    numerics = NumericsConfig(dt=450.0, integrator="rk2", output_freq=1,
                              temporal_reconstruction="linear", w_interp="leastsq",
                              temporal_interp=True)
    mode = ModeConfig(Fdrift2D=False, Fdrift3D=True, time_dir=direction)
    return dict(mesh=mesh, ds=ds, init=make_state(lon, lat, frac * h0), mode=mode,
                numerics=numerics, start_time_idx=0 if direction > 0 else n_time - 1, n_steps=12,
                n_time_total=n_time, file_times=times, file_dt=900.0)


HP_LEVELS = [0.0, 10.0, 30.0, 70.0]      # tests/test_healpix_reader.py:28 LEVELS_M
HP_NTIME, HP_NLEV = 3, 4                 # :27 NTIME, :26 NLEV


def healpix_store(nside=8):
    """tests/test_healpix_reader.py:45 _make_uv_store(nside=8, with_eta,
    with_w, n_finite_levels as tests/_gen_golden_3d.py:636-637, w_value
    1e-4) with u/v rewritten per level, cell and record
    (tests/_gen_golden_3d.py:594 _rewrite_healpix_uv), as arrays: the
    northern cap above 60 N is land, wet columns hold 3 finite layers south
    of the equator, 2 north of it and 1 at every fifth pixel; the deepest
    level is NaN padding for u/v; W sits on the interfaces (one more finite
    entry than u); eta 0.25 on the wet cells; daily records."""
    from kelp_amd import healpix
    from kelp_amd.readers import ArrayDataset
    npix = 12 * nside * nside
    plon, plat = healpix.pix2ang_nest(nside, np.arange(npix), lonlat=True)
    plon = np.mod(plon, 360.0)
    wet = plat < 60.0
    nfin = np.where(wet, np.where(plat < 0, 3, 2), 0)
    nfin[::5] = np.where(wet[::5], 1, 0)
    T, L = HP_NTIME, HP_NLEV
    t = np.arange(T, dtype=np.float64)[:, None, None]
    k = np.arange(L, dtype=np.float64)[None, :, None]
    x, y = (plon / 360.0)[None, None, :], (plat / 60.0)[None, None, :]
    alive = nfin[None, None, :] > k
    u = np.where(alive, 0.4 - 0.12 * k + 0.03 * y + 0.01 * t, np.nan)
    v = np.where(alive, 0.05 - 0.05 * k + 0.02 * x - 0.005 * t, np.nan)
    on_ifc = (nfin[None, None, :] >= k) & (nfin[None, None, :] > 0)
    w = np.where(on_ifc, 1e-4, np.nan) + np.zeros((T, 1, 1))
    eta = np.repeat(np.where(wet, 0.25, np.nan)[None, :], T, axis=0)
    c = lambda a: np.ascontiguousarray(np.broadcast_to(a, (T, L, npix)), dtype=np.float64)
    ds = ArrayDataset({"avg_uoe": (("time", "level", "value"), c(u)),
                       "avg_von": (("time", "level", "value"), c(v)),
                       "avg_wo": (("time", "level", "value"), c(w)),
                       "avg_zos": (("time", "value"), np.ascontiguousarray(eta))},
                      time=np.arange(T) * 86400.0,
                      attrs={"crs": {"grid_mapping_name": "healpix", "healpix_nside": nside,
                                     "healpix_order": "nest"}})
    return ds


def healpix_case(direction):
    """tests/_gen_golden_3d.py:625 healpix_3d_rk2: the native HEALPix
    reader's dual mesh and records, float32 velocities, node RK2 with
    barycentric u/v and W and a linear time blend, 6-hour steps over the
    three daily records."""
    from kelp_amd.config import ModeConfig, NumericsConfig
    from kelp_amd.particles import make_state
    from kelp_amd.readers import get_reader
    ds = healpix_store(8)
    mode = ModeConfig(Fdrift2D=False, Fdrift3D=True, time_dir=direction)
    mesh = get_reader("healpix_arrays").load_mesh(ds, levels=HP_LEVELS, mode_cfg=mode)
    lon = np.repeat([-150.0, -60.0, 0.0, 45.0, 120.0, 170.0], 4)
    lat = np.tile([-8.0, -2.0, 3.0, 51.0], 6)
    z = np.array([6.3, 15.0, 25.0, 29.5, 35.0, 47.6, 65.0, 69.0] * 3)
    numerics = NumericsConfig(dt=21600.0, integrator="rk2", uv_interp="barycentric",
                              w_interp="barycentric", temporal_interp=True, velocity_float32=True)
    return dict(mesh=mesh, ds=ds, init=make_state(lon, lat, z), mode=mode, numerics=numerics,
                start_time_idx=0 if direction > 0 else HP_NTIME - 1, n_steps=8,
                n_time_total=HP_NTIME, file_times=np.arange(HP_NTIME) * 86400.0, file_dt=86400.0)


MY_YEARS = {1958: 365, 1959: 2}              # tests/_gen_golden_3d.py:750
MY_FACTOR = {1958: 1.0, 1959: 1.25}          # :689 MULTIYEAR_YEAR
MY_W_UP = np.array([1.5e-5, 2.25e-5, 3.0e-5])  # :690 MULTIYEAR_W_UP


def multiyear_case():
    """tests/_gen_golden_3d.py:740 _multiyear_config: run_tracking on the
    FESOM2 unit-square mesh of tests/_fesom2_synth.py:16-24 (4 nodes, 2
    triangles, interfaces 0/10/30 m, bed 30 m, the FESOM2 Earth radius) over
    daily records stamped at 23:15 in the files of 1958 (365) and 1959 (2),
    read as one axis; unod/vnod/temp per layer and node, w on the
    interfaces (positive up), every record of a year alike (:693
    _multiyear_field), ssh 0.1, all stored float32; daily means shifted to
    the window midpoints; RK2, 6-hour steps from record 362 across the year
    boundary, temp sampled, failure reasons tracked, prefetch 2."""
    from kelp_amd.config import (AppConfig, DataConfig, MeshConfig, ModeConfig, NumericsConfig,
                                 TimeConfig)
    from kelp_amd.mesh import make_mesh
    from kelp_amd.particles import make_state
    from kelp_amd.readers import ArrayDataset
    lon_n, lat_n = np.array([0.0, 1.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0, 1.0])
    mesh = make_mesh("fesom2", lon_n, lat_n, np.array([[0, 1, 2], [0, 2, 3]], np.int32),
                     depth=np.full(4, 30.0), z_levels=np.array([0.0, 10.0, 30.0]),
                     r_earth=6_367_500.0)
    x, y = lon_n[None, :], lat_n[None, :]

    def fld(var, year):
        f = MY_FACTOR[year]
        if var == "unod":
            return (np.array([0.06, 0.02])[:, None] + 0.01 * x - 0.008 * y) * f
        if var == "vnod":
            return (np.array([0.03, -0.02])[:, None] + 0.006 * x + 0.004 * y) * f
        if var == "temp":
            return np.array([12.0, 6.0])[:, None] + 0.5 * x - 0.3 * y + 0.5 * (year - 1958)
        return np.broadcast_to(MY_W_UP[:, None] * f, (3, 4))

    data = {}
    for var, dim in (("unod", "nz1"), ("vnod", "nz1"), ("temp", "nz1"), ("w", "nz")):
        parts = [np.broadcast_to(fld(var, yr).astype(np.float32)[None], (n,) + fld(var, yr).shape)
                 for yr, n in MY_YEARS.items()]
        data[var] = (("time", dim, "nod2"), np.ascontiguousarray(np.concatenate(parts)))
    nt = sum(MY_YEARS.values())
    data["ssh"] = (("time", "nod2"), np.full((nt, 4), 0.1, np.float32))
    # one record a day, from the first, stamped at 23:15 (tests/_fesom2_synth.py:304)
    ds = ArrayDataset(data, time=np.arange(nt) * 86400.0, attrs={"time_origin": "1958-01-01 23:15:00"})
    seeds = np.array([(0.2, 0.3, 12.3), (0.5, 0.5, 17.6), (0.7, 0.2, 24.1),
                      (0.3, 0.8, 27.4), (0.72, 0.55, 8.7), (0.98, 0.97, 14.2)])
    cfg = AppConfig(
        mode=ModeConfig(Fdrift2D=False, Fdrift3D=True, var_u="unod", var_v="vnod", var_w="w",
                        var_eta="ssh", sample_vars=["temp"], track_failure_reason=True),
        numerics=NumericsConfig(dt=21600.0, integrator="rk2", uv_interp="barycentric",
                                w_interp="barycentric", temporal_interp=True, prefetch_size=2),
        mesh=MeshConfig(type="fesom2"),
        data=DataConfig(time_averaging="mean"),
        time=TimeConfig(start_time_idx=362, n_steps=12))
    return dict(mesh=mesh, ds=ds, config=cfg, init=make_state(seeds[:, 0], seeds[:, 1], seeds[:, 2]))
