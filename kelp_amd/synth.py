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


# ---------------------------------------------------------------------------
# 3-D global workload: Fibonacci-sphere Delaunay mesh with a land mask
# ---------------------------------------------------------------------------

C2_SEED = 2024
R_EARTH = 6_371_000.0
SNAP_DT = 21600.0           # six-hourly snapshots


def fibonacci_sphere(n):
    """Golden-angle spiral: n nodes quasi-uniform on the sphere (degrees)."""
    i = np.arange(n, dtype=np.float64) + 0.5
    lat = np.degrees(np.arcsin(1.0 - 2.0 * i / n))
    lon = (np.degrees(i * np.pi * (3.0 - np.sqrt(5.0))) % 360.0) - 180.0
    return lon, lat


def unit_xyz(lon, lat):
    lr, la = np.radians(lon), np.radians(lat)
    return np.stack([np.cos(la) * np.cos(lr), np.cos(la) * np.sin(lr), np.sin(la)], axis=1)


def harmonic_field(lon, lat, lmax=8, seed=C2_SEED):
    """Smooth random field: real spherical harmonics up to degree ``lmax``
    with N(0, 1)/(l+1) coefficients, scaled to unit standard deviation over
    the points."""
    from scipy.special import sph_harm_y
    rng = np.random.default_rng(seed)
    theta = np.radians(90.0 - np.asarray(lat))       # polar angle
    phi = np.radians(np.asarray(lon))                # azimuth
    f = np.zeros(theta.shape)
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            y = sph_harm_y(l, m, theta, phi)
            f += rng.normal() / (l + 1) * y.real
            if m > 0:
                f += rng.normal() / (l + 1) * y.imag
    return (f - f.mean()) / f.std()


def sphere_mesh(n_nodes, land_threshold=0.7, seed=C2_SEED, z_levels=None, name="fibonacci"):
    """Spherical Delaunay triangulation (convex hull) of ``n_nodes`` Fibonacci
    nodes; nodes where the harmonic field exceeds ``land_threshold`` are land
    and every triangle touching one is dropped (orphan nodes removed).  The
    bottom depth deepens away from the coast, 50 m at the land edge to 6000 m
    in the open ocean."""
    from scipy.spatial import ConvexHull
    lon, lat = fibonacci_sphere(n_nodes)
    tri = ConvexHull(unit_xyz(lon, lat)).simplices.astype(np.int64)
    f = harmonic_field(lon, lat, seed=seed)
    wet = f <= land_threshold
    tri = tri[wet[tri].all(axis=1)]
    used = np.zeros(n_nodes, dtype=bool)
    used[tri.ravel()] = True
    remap = np.full(n_nodes, -1, dtype=np.int64)
    remap[used] = np.arange(int(used.sum()))
    tri = remap[tri]
    # counter-clockwise seen from outside
    xyz = unit_xyz(lon[used], lat[used])
    a, b, c = xyz[tri[:, 0]], xyz[tri[:, 1]], xyz[tri[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a) < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    fw = f[used]
    depth = 50.0 + 5950.0 * np.clip((land_threshold - fw) / 2.0, 0.0, 1.0) ** 0.7
    if z_levels is None:
        z_levels = np.concatenate([[0.0], np.geomspace(5.0, 6000.0, 47)])
    mesh = make_mesh(name, lon[used], lat[used], tri.astype(np.int32), depth=depth,
                     z_levels=z_levels)
    from .mesh import build_boundary_info
    build_boundary_info(mesh)
    return mesh


def jet_rossby_snapshot(mesh, t, dtype=np.float32, amp=0.3, H=800.0):
    """u, v on the K-1 layers (mid-layer depth), W on the K interfaces
    (positive down), T and S on the layers, at time ``t`` seconds.

    Horizontal flow = g(z) * (zonal jet + Rossby waves from a streamfunction +
    a small divergent part grad(chi)), g(z) = exp(-z/H).  chi is a degree-2
    harmonic (Laplacian -6 chi / R^2), so continuity gives
    W(z) = 6 chi H (1 - exp(-z/H)) / R^2 exactly."""
    lam = np.radians(mesh.nodes_lon)[None, :]
    phi = np.radians(mesh.nodes_lat)[None, :]
    zl = mesh.sigma
    zm = 0.5 * (zl[:-1] + zl[1:])[:, None]
    g = np.exp(-zm / H)
    cphi, sphi = np.cos(phi), np.sin(phi)
    # zonal jets
    latd = mesh.nodes_lat[None, :]
    u = amp * (np.exp(-((latd - 35.0) / 8.0) ** 2) - 0.7 * np.exp(-((latd + 45.0) / 9.0) ** 2))
    v = np.zeros_like(u)
    # Rossby modes psi = A cos^2(phi) sin(2 phi) cos(m (lam - c t))
    for m, A, cdeg in ((3, 0.45 * amp * R_EARTH, -4.0), (5, 0.25 * amp * R_EARTH, -2.5)):
        ph = m * (lam - np.radians(cdeg) * t / 86400.0)
        dpsi_dphi = A * np.cos(ph) * (-2.0 * cphi * sphi * np.sin(2 * phi) + 2.0 * cphi ** 2 * np.cos(2 * phi))
        dpsi_dlam = -A * m * cphi ** 2 * np.sin(2 * phi) * np.sin(ph)
        u = u - dpsi_dphi / R_EARTH
        v = v + dpsi_dlam / (R_EARTH * np.maximum(cphi, 1e-6))
    # divergent part chi = B cos^2(phi) cos(2 lam - w t)
    B = 0.05 * amp * R_EARTH
    ph2 = 2.0 * lam - 2.0 * np.pi * t / (5 * 86400.0)
    chi = B * cphi ** 2 * np.cos(ph2)
    u = u + (-2.0 * B * cphi * np.sin(ph2)) / R_EARTH          # (1/(R cos)) dchi/dlam
    v = v + (-2.0 * B * cphi * sphi * np.cos(ph2)) / R_EARTH   # (1/R) dchi/dphi
    U = (g * u).astype(dtype)
    V = (g * v).astype(dtype)
    W = (6.0 * chi * H * (1.0 - np.exp(-zl[:, None] / H)) / R_EARTH ** 2).astype(dtype)
    T = (2.0 + 26.0 * np.exp(-zm / 700.0) * (0.6 + 0.4 * cphi)).astype(dtype)
    S = (34.0 + 1.5 * (0.5 + 0.5 * sphi * np.exp(-zm / 1500.0))).astype(dtype)
    return U, V, W, T, S


def wet_area_seeds(mesh, n, zmax=1000.0, seed=C2_SEED + 1):
    """``n`` seeds uniform over the wet area (area-weighted triangles,
    uniform barycentric sampling, projected to the sphere) with depths
    uniform in [0, zmax]."""
    rng = np.random.default_rng(seed)
    xyz = unit_xyz(mesh.nodes_lon, mesh.nodes_lat)
    en = mesh.elem_nodes
    a, b, c = xyz[en[:, 0]], xyz[en[:, 1]], xyz[en[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    e = rng.choice(en.shape[0], size=n, p=area / area.sum())
    r1, r2 = rng.random(n), rng.random(n)
    s = np.sqrt(r1)
    p = (1 - s)[:, None] * a[e] + (s * (1 - r2))[:, None] * b[e] + (s * r2)[:, None] * c[e]
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    lon = np.degrees(np.arctan2(p[:, 1], p[:, 0]))
    lat = np.degrees(np.arcsin(np.clip(p[:, 2], -1, 1)))
    z = rng.uniform(0.0, zmax, n)
    return lon, lat, z


@dataclass
class Workload3D:
    mesh: Mesh
    snaps: list                 # [(U, V, W, T, S)] per snapshot
    snap_dt: float
    lon: np.ndarray
    lat: np.ndarray
    z: np.ndarray
    order: int
    dt: float
    n_steps: int
    kh: float
    output_every: int


def global_3d_workload(n_nodes=1_000_000, n_particles=1_000_000, n_snaps=8,
                       seed=C2_SEED) -> Workload3D:
    """The 3-D global workload: 1M-node Fibonacci Delaunay mesh, 47 layers,
    8 six-hourly float32 snapshots of u, v, w, T, S; RK4, dt 1800 s, 336
    steps, time interpolation, Kh = 10 m^2/s random walk, beaching."""
    mesh = sphere_mesh(n_nodes, seed=seed)
    snaps = [jet_rossby_snapshot(mesh, s * SNAP_DT) for s in range(n_snaps)]
    lon, lat, z = wet_area_seeds(mesh, n_particles, seed=seed + 1)
    return Workload3D(mesh, snaps, SNAP_DT, lon, lat, z, order=4, dt=1800.0, n_steps=336,
                      kh=10.0, output_every=12)


def snapshot_bracket(t, snap_dt, n_snaps, direction=1.0):
    """Bracket (idx_a, idx_b, alpha0, dt_to_next) of time ``t`` over cycled
    snapshots (``time.on_exhausted: cycle``), as the payload of
    tracking._build_vel_payload: alpha0 snaps to 0/1 within 1e-12 and the
    stage target is b forward / a backward."""
    k = int(np.floor(t / snap_dt))
    ta = k * snap_dt
    tb = ta + snap_dt
    if direction < 0 and t == ta:
        k -= 1
        ta, tb = ta - snap_dt, ta
    ia, ib = k % n_snaps, (k + 1) % n_snaps
    alpha = (t - ta) / (tb - ta)
    if direction > 0:
        dt_to_next, it = tb - t, ib
    else:
        dt_to_next, it = t - ta, ia
    if abs(alpha) < 1e-12:
        return ia, -1, 0.0, it, dt_to_next
    if abs(alpha - 1.0) < 1e-12:
        return ib, -1, 0.0, it, dt_to_next
    return ia, ib, alpha, it, dt_to_next


# ---------------------------------------------------------------------------
# The batching workload: a mixed triangle/quad lattice with sigma layers and
# 256 independent release clouds.
# ---------------------------------------------------------------------------

C4_SEED = 40404
DAY = 86400.0


def mixed_lattice_mesh(lon0=0.0, lon1=80.0, lat0=20.0, lat1=70.0, step=0.25, jitter=0.05, n_sigma=21,
                       seed=C4_SEED, name="mixed_lattice") -> Mesh:
    """Jittered lon-lat lattice (321 x 201 nodes over 0-80E x 20-70N at the
    default step) whose quads are kept whole where the jitter's convexity
    test passes and split into two triangles otherwise.  The test: a quad's
    sharpest corner (the smallest sine of its four corner angles, all
    counter-clockwise) is at least the median over all quads, so half the
    quads stay whole.  Sigma coordinates with ``n_sigma`` interfaces (1 at
    the surface .. 0 at the bed), bathymetry a seeded smooth field in
    50-4000 m."""
    nlon = int(round((lon1 - lon0) / step)) + 1
    nlat = int(round((lat1 - lat0) / step)) + 1
    LON, LAT = np.meshgrid(lon0 + step * np.arange(nlon), lat0 + step * np.arange(nlat))
    rng = np.random.default_rng(seed)
    LON = LON + rng.uniform(-jitter, jitter, LON.shape)
    LAT = LAT + rng.uniform(-jitter, jitter, LAT.shape)
    lon, lat = LON.ravel(), LAT.ravel()
    nid = np.arange(nlat * nlon, dtype=np.int64).reshape(nlat, nlon)
    q = np.stack([nid[:-1, :-1].ravel(), nid[:-1, 1:].ravel(), nid[1:, 1:].ravel(), nid[1:, :-1].ravel()], 1)
    x = lon[q] * np.cos(np.radians(lat[q]))
    y = lat[q]
    sines = []
    for k in range(4):
        p, c, n = (k - 1) % 4, k, (k + 1) % 4
        ax, ay = x[:, c] - x[:, p], y[:, c] - y[:, p]
        bx, by = x[:, n] - x[:, c], y[:, n] - y[:, c]
        sines.append((ax * by - ay * bx) / (np.hypot(ax, ay) * np.hypot(bx, by)))
    sharp = np.min(np.stack(sines, 1), axis=1)
    keep = (sharp >= np.median(sharp)) & (sharp > 0.0)
    quads = q[keep]
    split = q[~keep]
    tris = np.concatenate([split[:, [0, 1, 2]], split[:, [0, 2, 3]]])
    en = np.full((quads.shape[0] + tris.shape[0], 4), -1, dtype=np.int32)
    en[:quads.shape[0]] = quads
    en[quads.shape[0]:, :3] = tris
    rng2 = np.random.default_rng(seed + 1)
    phase = rng2.uniform(0, 2 * np.pi, 6)
    xs, ys = np.radians(lon), np.radians(lat)
    f = (np.sin(3 * xs + phase[0]) * np.cos(4 * ys + phase[1]) + 0.6 * np.sin(7 * xs + phase[2])
         * np.sin(5 * ys + phase[3]) + 0.4 * np.cos(11 * xs + phase[4] + 9 * ys + phase[5]))
    f = (f - f.min()) / (f.max() - f.min())
    depth = 50.0 + 3950.0 * f ** 1.5
    mesh = make_mesh(name, lon, lat, en, depth=depth)
    mesh.sigma = 1.0 - np.linspace(0.0, 1.0, n_sigma)
    mesh.vertical_coord = "sigma"
    from .mesh import build_boundary_info
    build_boundary_info(mesh)
    return mesh


def mixed_lattice_fields(mesh, U=0.4, dtype=np.float32, lon0=0.0, lon1=80.0, lat0=20.0, lat1=70.0):
    """Steady fields of the batching workload: a double gyre over the domain
    with a depth-decaying sigma profile on the K-1 layers, a small dia-sigma
    flux W on the K interfaces (0 at the surface and the bed) and a surface
    elevation; and the 2-D (surface layer) u, v."""
    K = mesh.sigma.size
    x = (mesh.nodes_lon - lon0) / (lon1 - lon0)
    y = (mesh.nodes_lat - lat0) / (lat1 - lat0)
    u0 = -U * np.sin(np.pi * x) * np.cos(np.pi * y)
    v0 = U * np.cos(np.pi * x) * np.sin(np.pi * y)
    smid = 0.5 * (mesh.sigma[:-1] + mesh.sigma[1:])[:, None]
    prof = 0.3 + 0.7 * smid ** 2
    u = (prof * u0[None, :]).astype(dtype)
    v = (prof * v0[None, :]).astype(dtype)
    s = mesh.sigma[:, None]
    w = (2e-6 * np.sin(np.pi * s) * np.sin(2 * np.pi * x)[None, :] * np.cos(np.pi * y)[None, :]).astype(dtype)
    w[0] = 0.0
    w[-1] = 0.0
    eta = 0.3 * np.sin(2 * np.pi * x) * np.sin(np.pi * y)
    return u, v, w, eta, u[0].copy(), v[0].copy()


def batch_clouds(mesh, n_clouds=256, seed=C4_SEED + 2, min_size=100, max_size=200_000, min_days=10,
                 max_days=365, kh=10.0, size_scale=1.0):
    """The batching workload's release clouds: sizes log-uniform in
    [min_size, max_size] (times ``size_scale``), durations uniform in
    [min_days, max_days] daily steps, half forward and half backward, 2-D
    and 3-D 1:1, the random walk on for a random half; each cloud uniform
    in a disc of radius 0.5-2 degrees around a random interior point, 3-D
    depths uniform in 0-40 m.  Returns a list of dicts (lon, lat, z,
    n_steps, direction, is3d, kh, seed)."""
    rng = np.random.default_rng(seed)
    lon_min, lon_max = mesh.nodes_lon.min() + 3.0, mesh.nodes_lon.max() - 3.0
    lat_min, lat_max = mesh.nodes_lat.min() + 3.0, mesh.nodes_lat.max() - 3.0
    out = []
    dirs = np.array([1.0, -1.0] * (n_clouds // 2) + [1.0] * (n_clouds % 2))
    rng.shuffle(dirs)
    kinds = np.array([True, False] * (n_clouds // 2) + [True] * (n_clouds % 2))
    rng.shuffle(kinds)
    for k in range(n_clouds):
        size = int(round(np.exp(rng.uniform(np.log(min_size), np.log(max_size))) * size_scale))
        size = max(size, 1)
        days = int(rng.integers(min_days, max_days + 1))
        c_lon, c_lat = rng.uniform(lon_min, lon_max), rng.uniform(lat_min, lat_max)
        rad = rng.uniform(0.5, 2.0)
        r = rad * np.sqrt(rng.random(size))
        th = rng.uniform(0, 2 * np.pi, size)
        lat = c_lat + r * np.sin(th)
        lon = c_lon + r * np.cos(th) / np.cos(np.radians(lat))
        z = rng.uniform(0.0, 40.0, size)
        out.append(dict(lon=lon, lat=lat, z=z, n_steps=days, direction=float(dirs[k]), is3d=bool(kinds[k]),
                        kh=kh if rng.random() < 0.5 else 0.0, seed=int(rng.integers(1, 2 ** 62))))
    return out


# ---------------------------------------------------------------------------
# The precision workload: a NEST HEALPix dual mesh with 69 z-levels.
# ---------------------------------------------------------------------------

C3_LEVELS = np.concatenate([[0.0], np.geomspace(5.0, 6000.0, 68)])     # 69 interfaces


def real_harmonics_field(lon, lat, lmax=8, seed=C2_SEED, chunk=2_000_000):
    """:func:`harmonic_field` evaluated by the three-term recurrence of the
    fully normalised associated Legendre functions (the normalisation and
    Condon-Shortley phase of scipy's sph_harm_y, the same coefficients in
    the same order), in chunks: fast enough for the 12.6 M pixel centres of
    a HEALPix grid at nside 1024.  Equal to :func:`harmonic_field` to
    rounding."""
    lon = np.asarray(lon, dtype=np.float64)
    lat = np.asarray(lat, dtype=np.float64)
    rng = np.random.default_rng(seed)
    coef = []
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            coef.append((l, m, rng.normal() / (l + 1), rng.normal() / (l + 1) if m > 0 else 0.0))
    f = np.empty(lon.shape)
    for s in range(0, lon.size, chunk):
        th = np.radians(90.0 - lat[s:s + chunk])
        ph = np.radians(lon[s:s + chunk])
        x, sx = np.cos(th), np.sin(th)
        acc = np.zeros(th.shape)
        for m in range(0, lmax + 1):
            # P_mm, then upwards in l
            norm = (2 * m + 1) / (4 * np.pi) * np.prod([(2 * k - 1) / (2 * k) for k in range(1, m + 1)])
            pmm = (-1.0) ** m * np.sqrt(norm) * sx ** m
            p = {m: pmm}
            if m + 1 <= lmax:
                p[m + 1] = np.sqrt(2 * m + 3.0) * x * pmm
            for l in range(m + 2, lmax + 1):
                a = np.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
                b = np.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
                p[l] = a * (x * p[l - 1] - b * p[l - 2])
            cm, sm = np.cos(m * ph), np.sin(m * ph)
            for l, mm, cre, cim in coef:
                if mm == m:
                    acc += cre * p[l] * cm
                    if m > 0:
                        acc += cim * p[l] * sm
        f[s:s + chunk] = acc
    return (f - f.mean()) / f.std()


def healpix_c3_mesh(nside=1024, land_threshold=0.7, seed=C2_SEED, z_levels=None):
    """NEST HEALPix dual mesh (pixel centres = nodes) over the wet pixels of
    the seeded l <= 8 harmonic mask of the 3-D workload (land where the
    field exceeds ``land_threshold``), 69 z-level interfaces from 0 to
    6000 m, bottom depth from the field as on the Fibonacci mesh."""
    from .healpix import dual_mesh, pix2ang_nest
    npix = 12 * nside * nside
    lon, lat = pix2ang_nest(nside, np.arange(npix, dtype=np.int64), lonlat=True)
    f = real_harmonics_field(lon, lat, seed=seed)
    wet = f <= land_threshold
    depth = 50.0 + 5950.0 * np.clip((land_threshold - f) / 2.0, 0.0, 1.0) ** 0.7
    mesh = dual_mesh(nside, wet_mask=wet, depth=depth,
                     z_levels=C3_LEVELS if z_levels is None else z_levels, name=f"healpix_nest_{nside}")
    from .mesh import build_boundary_info
    build_boundary_info(mesh)
    return mesh


def jet_rossby_snapshot_torch(mesh, t, torch, device, dtype, amp=0.3, H=800.0):
    """:func:`jet_rossby_snapshot` (u, v, W) computed with torch on
    ``device`` in float64, returned as numpy arrays of ``dtype``; for meshes
    of millions of nodes and tens of layers."""
    f64 = torch.float64
    lam = torch.as_tensor(np.radians(mesh.nodes_lon), dtype=f64, device=device)[None, :]
    phi = torch.as_tensor(np.radians(mesh.nodes_lat), dtype=f64, device=device)[None, :]
    latd = torch.as_tensor(mesh.nodes_lat, dtype=f64, device=device)[None, :]
    zl = torch.as_tensor(mesh.sigma, dtype=f64, device=device)
    zm = (0.5 * (zl[:-1] + zl[1:]))[:, None]
    g = torch.exp(-zm / H)
    cphi, sphi = torch.cos(phi), torch.sin(phi)
    u = amp * (torch.exp(-((latd - 35.0) / 8.0) ** 2) - 0.7 * torch.exp(-((latd + 45.0) / 9.0) ** 2))
    v = torch.zeros_like(u)
    for m, A, cdeg in ((3, 0.45 * amp * R_EARTH, -4.0), (5, 0.25 * amp * R_EARTH, -2.5)):
        ph = m * (lam - np.radians(cdeg) * t / 86400.0)
        dpsi_dphi = A * torch.cos(ph) * (-2.0 * cphi * sphi * torch.sin(2 * phi) + 2.0 * cphi ** 2 * torch.cos(2 * phi))
        dpsi_dlam = -A * m * cphi ** 2 * torch.sin(2 * phi) * torch.sin(ph)
        u = u - dpsi_dphi / R_EARTH
        v = v + dpsi_dlam / (R_EARTH * torch.clamp(cphi, min=1e-6))
    B = 0.05 * amp * R_EARTH
    ph2 = 2.0 * lam - 2.0 * np.pi * t / (5 * 86400.0)
    chi = B * cphi ** 2 * torch.cos(ph2)
    u = u + (-2.0 * B * cphi * torch.sin(ph2)) / R_EARTH
    v = v + (-2.0 * B * cphi * sphi * torch.cos(ph2)) / R_EARTH
    tdt = torch.float32 if dtype == np.float32 else f64
    U = (g * u).to(tdt).cpu().numpy()
    V = (g * v).to(tdt).cpu().numpy()
    W = (6.0 * chi * H * (1.0 - torch.exp(-zl[:, None] / H)) / R_EARTH ** 2).to(tdt).cpu().numpy()
    return U, V, W
