"""Pins of the oracle beyond the reference's golden vectors (CPU).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, a transposed operand, a wrong tie rule) fails it:
published SplitMix64 outputs, Python's own `%`, brute-force nearest node and
containment, the linear precision of barycentric and Wachspress weights,
closed forms of the steppers on uniform flow, the first two moments of the
random walk and of the stochastic beaching, and the device math against
glibc.
"""
import ctypes
import math

import numpy as np
import pytest

from kelp_amd import healpix, synth
from oracle.oracle import OracleMesh, lib, math_probe, step_args

from .helpers import brute_nearest, chord2, global_latlon_mesh

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def sm64_py(x):
    """SplitMix64 round (Steele, Lea & Flood 2014; Vigna's splitmix64.c):
    x += gamma, then the two xor-shift-multiply mixes and a final xor-shift."""
    z = (x + GAMMA) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_splitmix64_published_sequence_and_key_chain():
    # the first outputs of splitmix64 seeded with 0 (Vigna's reference code)
    assert sm64_py(0) == 0xE220A8397B1DCDAF
    assert sm64_py(GAMMA) == 0x6E789E6AA1B965F4
    assert sm64_py((2 * GAMMA) & M64) == 0x06C45D188009454F
    L = lib("glibc")
    # a chain whose xor keys walk the generator's state: outputs 1..4
    h1 = sm64_py(0)
    k2 = h1 ^ GAMMA
    h2 = sm64_py(GAMMA)
    k3 = h2 ^ ((2 * GAMMA) & M64)
    h3 = sm64_py((2 * GAMMA) & M64)
    k4 = h3 ^ ((3 * GAMMA) & M64)
    assert L.or_hash5(0, 0, k2, k3, k4) == sm64_py((3 * GAMMA) & M64)
    rng = np.random.default_rng(7)
    for _ in range(200):
        k = [int(x) for x in rng.integers(0, 2**63, 5, dtype=np.int64)]
        h = k[0]
        for kk in k[1:]:
            h = sm64_py(h ^ kk)
        assert L.or_hash5(*k) == h
        # beaching.py:46: (h >> 11) / 2^53 exactly
        assert L.or_u01(*k) == (h >> 11) / 9007199254740992.0


def test_box_muller_pair_moments():
    L = lib("glibc")
    g1 = np.empty(40000)
    g2 = np.empty(40000)
    a, b = ctypes.c_double(), ctypes.c_double()
    for i in range(g1.size):
        L.or_normal_pair(11, 0, i, 3, 0, ctypes.byref(a), ctypes.byref(b))
        g1[i], g2[i] = a.value, b.value
    for g in (g1, g2):
        assert abs(g.mean()) < 0.02 and abs(g.var() - 1.0) < 0.03
    assert abs(np.corrcoef(g1, g2)[0, 1]) < 0.02


@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
def test_python_modulo_semantics(flavour):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-2000, 2000, 5000), [-0.0, 0.0, -360.0, 360.0, 720.0,
                                                         -1e-300, 1e-300, -180.0, 179.99999999999997]])
    got = math_probe(7, x, flavour=flavour)
    want = np.array([float(v) % 360.0 for v in x])
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    d = np.exp(rng.uniform(-3, 8, 5000))
    x = rng.choice([-1.0, 1.0], d.size) * np.exp(rng.uniform(-5, 45, d.size))
    got = math_probe(9, x, d, flavour=flavour)
    want = np.array([float(a) % float(b) for a, b in zip(x, d)])
    assert np.array_equal(got.view(np.int64), want.view(np.int64))


@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
def test_reflect_depth_is_a_specular_fold(flavour):
    # numerics/diffusion.py:334 on dyadic values, where every step is exact:
    # inside the column it is the identity; an overshoot of d past either
    # end comes back by d; the fold has period 2L
    L = 64.0
    z = np.arange(-3 * 64, 5 * 64, 0.25)
    got = math_probe(10, z, np.full(z.size, L), flavour=flavour)
    want = np.abs(((z + 2 * L) % (2 * L)))
    want = np.where(want > L, 2 * L - want, want)
    assert np.array_equal(got, want)
    inside = (z >= 0) & (z <= L)
    assert np.array_equal(got[inside], z[inside])
    assert np.all((got >= 0) & (got <= L))


@pytest.mark.parametrize("fn,lo,hi", [(0, -7, 7), (1, -7, 7), (2, -1, 1), (3, -1, 1),
                                      (5, 1e-300, 1e300), (6, -700, 700)])
def test_device_math_within_two_ulp_of_glibc(fn, lo, hi):
    rng = np.random.default_rng(fn)
    if fn == 5:
        x = np.exp(rng.uniform(math.log(lo), math.log(hi), 50000))
    else:
        x = rng.uniform(lo, hi, 50000)
    a = math_probe(fn, x, flavour="kmath")
    b = math_probe(fn, x, flavour="glibc")
    ulp = np.spacing(np.abs(b))
    assert np.max(np.abs(a - b) / ulp) <= 2.0


def _box_muller_args(n, seed):
    """The uniforms of the counter RNG's Box-Muller pair, (hi, lo 32 bits)
    2^-32 + 2^-33 (numerics/diffusion.py:219), and 2 pi u2."""
    rng = np.random.default_rng(seed)
    u1 = rng.integers(0, 2 ** 32, n) * 2.0 ** -32 + 2.0 ** -33
    u2 = rng.integers(0, 2 ** 32, n) * 2.0 ** -32 + 2.0 ** -33
    return u1, 2.0 * np.pi * u2


@pytest.mark.parametrize("fn", [5, 0, 1])
def test_draw_math_correctly_rounded(fn):
    """log, sin and cos of the random walk's Box-Muller arguments (and cos of
    latitudes) are correctly rounded: equal to the 160-bit mpmath value
    rounded to double on every sample.  The same samples measure how often
    glibc (the reference's libm under Numba) rounds otherwise; those are the
    only places where device draws can differ from the reference's."""
    import mpmath as mp
    mp.mp.prec = 160
    u1, a = _box_muller_args(4000, 11 + fn)
    x = u1 if fn == 5 else np.concatenate([a, np.radians(np.linspace(-90.0, 90.0, 1000))])
    f = {5: mp.log, 0: mp.sin, 1: mp.cos}[fn]
    ref = np.array([float(f(mp.mpf(float(v)))) for v in x])
    k = math_probe(fn, x, flavour="kmath")
    g = math_probe(fn, x, flavour="glibc")
    assert np.array_equal(k, ref)
    assert np.mean(g != ref) < 0.01


@pytest.mark.parametrize("fn", [2, 3, 4, 6, 13])
def test_inverse_trig_exp_pow_correctly_rounded(fn):
    """asin and atan2 of the random walk's displaced unit vectors
    (numerics/diffusion.py:539 sphere_displace), acos near 1 (the exact
    search's cap angles), exp of the beaching probabilities' arguments and
    pow(x, 0.2) of the rk45 controller: equal to the 160-bit mpmath value
    rounded to double on every sample; glibc rounds otherwise on a small
    fraction, where the device can differ from the reference."""
    import mpmath as mp
    mp.mp.prec = 160
    rng = np.random.default_rng(40 + fn)
    n = 3000
    y = None
    if fn == 2:
        x = np.concatenate([rng.uniform(-1, 1, n), np.sin(np.radians(rng.uniform(-90, 90, n)))])
        f = lambda a, b: mp.asin(a)
    elif fn == 3:
        x = np.concatenate([rng.uniform(-1, 1, n), 1 - 10 ** rng.uniform(-16, -1, n)])
        f = lambda a, b: mp.acos(a)
    elif fn == 4:
        x, y = rng.normal(size=2 * n), rng.normal(size=2 * n)
        f = lambda a, b: mp.atan2(a, b)
    elif fn == 6:
        x = np.concatenate([rng.uniform(-30, 30, n), -rng.uniform(0, 0.1, n)])
        f = lambda a, b: mp.exp(a)
    else:
        x, y = 10 ** rng.uniform(-6, 6, 2 * n), np.full(2 * n, 0.2)
        f = lambda a, b: mp.power(a, b)
    yy = x if y is None else y
    ref = np.array([float(f(mp.mpf(float(a)), mp.mpf(float(b)))) for a, b in zip(x, yy)])
    k = math_probe(fn, x, y, flavour="kmath")
    g = math_probe(fn, x, y, flavour="glibc")
    assert np.array_equal(k, ref)
    assert np.mean(g != ref) < 0.01


def test_atan2_within_two_ulp_of_glibc():
    rng = np.random.default_rng(4)
    y = rng.uniform(-3, 3, 50000)
    x = rng.uniform(-3, 3, 50000)
    a = math_probe(4, y, x, flavour="kmath")
    b = math_probe(4, y, x, flavour="glibc")
    assert np.max(np.abs(a - b) / np.spacing(np.abs(b))) <= 2.0


# ---------------------------------------------------------------------------
# localisation
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def meshes():
    lat = synth.lattice_mesh(0.0, 10.0, 30.0, 40.0, 0.25, jitter=0.05, seed=9)
    glob = global_latlon_mesh(step=6.0, jitter=1.0, seed=2)
    sph = synth.sphere_mesh(4000)
    return {"lattice": lat, "global": glob, "sphere": sph}


@pytest.mark.parametrize("name", ["lattice", "global", "sphere"])
def test_exact_nearest_is_brute_force_argmin(meshes, name):
    mesh = meshes[name]
    om = OracleMesh(mesh)
    rng = np.random.default_rng(1)
    if name == "lattice":
        lon, lat = rng.uniform(-1, 11, 3000), rng.uniform(29, 41, 3000)
    else:
        lon = rng.uniform(-180, 180, 3000)
        lat = np.degrees(np.arcsin(rng.uniform(-1, 1, 3000)))
    got = om.nearest(lon, lat, exact=True)
    want, d2 = brute_nearest(om.node_xyz, lon, lat)
    same = got == want
    # a different node must be an exact tie in chord length
    assert np.all(same | (chord2(om.node_xyz, got, lon, lat) == d2))
    assert same.mean() > 0.999


def test_containment_verdict_matches_brute_force(meshes):
    mesh = meshes["lattice"]
    om = OracleMesh(mesh)
    rng = np.random.default_rng(2)
    lon, lat = rng.uniform(0.2, 9.8, 1500), rng.uniform(30.2, 39.8, 1500)
    nn = om.nearest(lon, lat, exact=True)
    e = om.containing(lon, lat, nn)
    for i in range(lon.size):
        hits = [k for k in range(mesh.n_elements) if om.element_contains(k, lon[i], lat[i])]
        assert (e[i] >= 0) == bool(hits)
        if e[i] >= 0:
            assert e[i] in hits


def test_outside_points_are_not_contained(meshes):
    mesh = meshes["lattice"]
    om = OracleMesh(mesh)
    lon = np.array([-0.5, 10.6, 5.0, 5.0])
    lat = np.array([35.0, 35.0, 29.5, 40.6])
    nn = om.nearest(lon, lat, exact=True)
    assert np.all(om.containing(lon, lat, nn) < 0)


# ---------------------------------------------------------------------------
# weights: linear precision on triangles and Wachspress quads
# ---------------------------------------------------------------------------

def test_interpolation_reproduces_linear_fields():
    """A field linear in the element's local metric plane is reproduced by
    barycentric (triangles) and Wachspress (quads) weights; the plane of
    interpolate.py:1376 is affine in (lon, lat) per element, so a field
    linear in lon and lat is reproduced to rounding."""
    mesh = healpix.dual_mesh(8)
    assert mesh.max_v == 4
    om = OracleMesh(mesh)
    rng = np.random.default_rng(6)
    lon = rng.uniform(-170, 170, 800)
    lat = rng.uniform(-60, 60, 800)
    nn = om.nearest(lon, lat)
    f = lambda lo, la: 0.3 + 0.01 * lo - 0.02 * la
    g = lambda lo, la: -1.0 + 0.005 * la
    U, V, ok, el = om.interp2d(lon, lat, nn, f(mesh.nodes_lon, mesh.nodes_lat),
                               g(mesh.nodes_lon, mesh.nodes_lat))
    assert ok.mean() > 0.99
    quads = (mesh.elem_nodes[el[ok], 3] >= 0).sum()
    assert quads > 100
    np.testing.assert_allclose(U[ok], f(lon[ok], lat[ok]), rtol=0, atol=1e-11)
    np.testing.assert_allclose(V[ok], g(lon[ok], lat[ok]), rtol=0, atol=1e-11)
    # partition of unity: a constant field comes back exactly to rounding
    U1, _, ok1, _ = om.interp2d(lon, lat, nn, np.full(mesh.n_nodes, 2.5), np.zeros(mesh.n_nodes))
    np.testing.assert_allclose(U1[ok1], 2.5, rtol=0, atol=1e-13)


# ---------------------------------------------------------------------------
# steppers: closed forms on uniform flow, reversibility
# ---------------------------------------------------------------------------

def _uniform_run(mesh, order, u0, v0, dt, n, time_dir=1.0, lon=None, lat=None):
    om = OracleMesh(mesh)
    u = np.full(mesh.n_nodes, u0)
    v = np.full(mesh.n_nodes, v0)
    st = dict(lon=lon.copy(), lat=lat.copy(), status=np.ones(lon.size, np.int32),
              nn=np.full(lon.size, -1, np.int32))
    for _ in range(n):
        st = om.step2d(st, order, u, v, dt, time_dir)
    return st


def test_uniform_zonal_flow_closed_form(meshes):
    # CLAUDE.md's constant-flow identity: dlon = U dt / (cos(lat) 111120)
    mesh = meshes["lattice"]
    lon = np.linspace(2, 6, 50)
    lat = np.linspace(31, 39, 50)
    outs = [_uniform_run(mesh, o, 0.4, 0.0, 900.0, 6, lon=lon, lat=lat) for o in (1, 2, 4)]
    want = lon + 6 * 900.0 * 0.4 / (np.cos(np.radians(lat)) * 111120.0)
    for o in outs:
        assert np.all(o["status"] == 1)
        np.testing.assert_allclose(o["lon"], want, rtol=0, atol=1e-11)
        np.testing.assert_array_equal(o["lat"], lat)
    # on a latitude-independent metric the three orders agree to rounding
    np.testing.assert_allclose(outs[1]["lon"], outs[0]["lon"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(outs[2]["lon"], outs[0]["lon"], rtol=0, atol=1e-12)


def test_uniform_meridional_flow_orders(meshes):
    # v only: dlat = V dt / 111120 exactly per step for every order
    mesh = meshes["lattice"]
    lon = np.linspace(2, 8, 40)
    lat = np.linspace(31, 35, 40)
    for o in (1, 2, 4):
        out = _uniform_run(mesh, o, 0.0, -0.3, 600.0, 5, lon=lon, lat=lat)
        np.testing.assert_allclose(out["lat"], lat - 5 * 600.0 * 0.3 / 111120.0, rtol=0, atol=1e-11)
        # lon moves only by the rounding of the ((lon + 180) % 360) - 180 fold
        np.testing.assert_allclose(out["lon"], lon, rtol=0, atol=1e-13)


def test_backward_run_retraces_forward_run(meshes):
    w = synth.double_gyre_workload(n_particles=300)
    om = OracleMesh(w.mesh)
    st = dict(lon=w.lon.copy(), lat=w.lat.copy(), status=np.ones(w.lon.size, np.int32),
              nn=np.full(w.lon.size, -1, np.int32))
    for _ in range(30):
        st = om.step2d(st, 4, w.u, w.v, 600.0, 1.0)
    fwd = st
    for _ in range(30):
        st = om.step2d(st, 4, w.u, w.v, 600.0, -1.0)
    ok = (fwd["status"] == 1) & (st["status"] == 1)
    assert ok.mean() > 0.95
    assert np.max(np.abs(st["lon"][ok] - w.lon[ok])) < 1e-5
    assert np.max(np.abs(st["lat"][ok] - w.lat[ok])) < 1e-5


# ---------------------------------------------------------------------------
# random walk and beaching moments
# ---------------------------------------------------------------------------

def _still_state(mesh, n, rng):
    lon = rng.uniform(3.0, 7.0, n)
    lat = rng.uniform(33.0, 37.0, n)
    return dict(lon=lon, lat=lat, z=np.zeros(n), status=np.ones(n, np.int32),
                nn=np.full(n, -1, np.int32), pid=np.arange(n, dtype=np.int64))


def test_kick_variance_is_2_K_dt(meshes):
    """One split kick on still water: the displacement in metres has mean 0
    and variance 2 K dt per axis (diffusion.py:261); anisotropic K keeps the
    two axes apart."""
    mesh = meshes["lattice"]
    om = OracleMesh(mesh)
    n = 20000
    rng = np.random.default_rng(8)
    st = _still_state(mesh, n, rng)
    zero = np.zeros(mesh.n_nodes)
    f = dict(u_a=zero, v_a=zero, alpha0=0.0)
    kx, ky, dt = 40.0, 10.0, 3600.0
    out = om.step(st, f, step_args(1, dt, 1.0, is3d=False, exact=1,
                                   kick=dict(kh_x=kx, kh_y=ky, seed=5)))
    assert np.all(out["status"] == 1)
    dy = (out["lat"] - st["lat"]) * 111120.0
    dx = (out["lon"] - st["lon"]) * 111120.0 * np.cos(np.radians(st["lat"]))
    assert abs(dx.var() / (2 * kx * dt) - 1.0) < 0.04
    assert abs(dy.var() / (2 * ky * dt) - 1.0) < 0.04
    assert abs(dx.mean()) < 4 * math.sqrt(2 * kx * dt / n)
    # (kx - ky) tan(lat) / R |dt| northward drift of the anisotropic metric
    drift = (kx - ky) * np.tan(np.radians(st["lat"])) / (111120.0 / (np.pi / 180.0)) * dt
    assert abs(dy.mean() - drift.mean()) < 4 * math.sqrt(2 * ky * dt / n)


def test_stochastic_beaching_probability(meshes):
    """In the stochastic zone an alive particle beaches with probability
    1 - exp(-|dt| / t_beach) per step (beaching.py:140)."""
    mesh = meshes["lattice"]
    mesh_shallow = synth.lattice_mesh(0.0, 10.0, 30.0, 40.0, 0.25, jitter=0.05, seed=9)
    mesh_shallow.depth[:] = 0.4      # every node inside zone_depth = 2 * depth_min
    om = OracleMesh(mesh_shallow)
    n = 20000
    rng = np.random.default_rng(12)
    st = _still_state(mesh_shallow, n, rng)
    zero = np.zeros(mesh_shallow.n_nodes)
    f = dict(u_a=zero, v_a=zero, alpha0=0.0)
    dt, tb = 3600.0, 7200.0
    beach = dict(criteria=["stochastic"], depth_min=0.3, t_beach=tb, seed=17)
    out = om.step(st, f, step_args(1, dt, 1.0, is3d=False, exact=1, beach=beach))
    p = 1.0 - math.exp(-dt / tb)
    frac = (out["status"] == 2).mean()
    assert abs(frac - p) < 4 * math.sqrt(p * (1 - p) / n)
    # depth criterion: deterministic
    beach = dict(criteria=["depth"], depth_min=0.5)
    out = om.step(st, f, step_args(1, dt, 1.0, is3d=False, exact=1, beach=beach))
    assert np.all(out["status"] == 2)
    beach = dict(criteria=["depth"], depth_min=0.3)
    out = om.step(st, f, step_args(1, dt, 1.0, is3d=False, exact=1, beach=beach))
    assert np.all(out["status"] == 1)


def test_periodic_fold_is_python_modulo(meshes):
    """tracking.py:1089: a particle leaving the window [x0, x1) comes back
    at x0 + (x - x0) % (x1 - x0); in-window particles are untouched."""
    mesh = meshes["lattice"]
    om = OracleMesh(mesh)
    lon = np.linspace(8.6, 8.99, 30)
    lat = np.linspace(31, 39, 30)
    st = dict(lon=lon, lat=lat, z=np.zeros(30), status=np.ones(30, np.int32),
              nn=np.full(30, -1, np.int32), pid=np.arange(30))
    u = np.full(mesh.n_nodes, 3.0)
    v = np.zeros(mesh.n_nodes)
    f = dict(u_a=u, v_a=v, alpha0=0.0)
    free = om.step(st, f, step_args(1, 3600.0, 1.0, exact=1))
    wrapped = om.step(st, f, step_args(1, 3600.0, 1.0, exact=1, periodic_lon=(1.0, 9.0)))
    out = free["lon"] >= 9.0
    assert out.any() and (~out).any()
    want = np.where(out, [1.0 + (float(x) - 1.0) % 8.0 for x in free["lon"]], free["lon"])
    np.testing.assert_array_equal(wrapped["lon"], want)
    # the folded particles carry the nearest node of their new position
    np.testing.assert_array_equal(wrapped["nn"][out], om.nearest(wrapped["lon"][out], wrapped["lat"][out]))


# ---------------------------------------------------------------------------
# edge walk (elem_search: walk)
# ---------------------------------------------------------------------------

def test_element_neighbours_are_symmetric_and_bound_the_mesh(meshes):
    """core/mesh.py:2824: the neighbour across edge k of e has e across the
    same edge; -1 marks exactly the edges owned by one element."""
    from kelp_amd.mesh import build_elem_neighbors
    mesh = meshes["lattice"]
    nb = build_elem_neighbors(mesh)
    en = mesh.elem_nodes
    edges = {}
    for e in range(mesh.n_elements):
        v = [x for x in en[e] if x >= 0]
        for k in range(len(v)):
            a, b = v[k], v[(k + 1) % len(v)]
            edges.setdefault((min(a, b), max(a, b)), []).append((e, k))
    for owners in edges.values():
        if len(owners) == 1:
            e, k = owners[0]
            assert nb[e, k] == -1
        else:
            (e1, k1), (e2, k2) = owners
            assert nb[e1, k1] == e2 and nb[e2, k2] == e1


@pytest.mark.parametrize("name", ["lattice", "sphere"])
def test_walk_verdict_matches_brute_force(meshes, name):
    """The walk (or its fallback) finds a containing element exactly when
    one exists, from any starting element (far, near, invalid)."""
    from kelp_amd.mesh import build_elem_neighbors
    mesh = meshes[name]
    build_elem_neighbors(mesh)
    om = OracleMesh(mesh)
    rng = np.random.default_rng(4)
    if name == "lattice":
        lon, lat = rng.uniform(-0.3, 10.3, 600), rng.uniform(29.7, 40.3, 600)
    else:
        lon = rng.uniform(-180, 180, 600)
        lat = np.degrees(np.arcsin(rng.uniform(-1, 1, 600)))
    start = rng.integers(-1, mesh.n_elements, lon.size)
    e = om.walk_locate(lon, lat, start)
    for i in range(lon.size):
        hits = [k for k in range(mesh.n_elements) if om.element_contains(k, lon[i], lat[i])]
        assert (e[i] >= 0) == bool(hits), i
        if e[i] >= 0:
            assert e[i] in hits


@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
def test_pchip_is_scipy_pchip_on_the_central_interval(flavour):
    """temporal_reconstruct.py:184 on [t1, t2] against scipy's
    PchipInterpolator (the same Fritsch-Carlson interior slopes; its end
    slopes do not enter the central interval), on uneven record spacing,
    monotone and non-monotone data, flat and sign-changing differences;
    and the interpolation conditions at t1 and t2 exactly."""
    from scipy.interpolate import PchipInterpolator
    from oracle.oracle import pchip4
    rng = np.random.default_rng(3)
    times = np.array([0.0, 2100.0, 3600.0, 8000.0])
    f = rng.standard_normal((4, 400))
    f[:, :50] = np.cumsum(np.abs(f[:, :50]), axis=0)      # monotone
    f[1, 50:60] = f[0, 50:60]                             # flat left difference
    f[2, 60:70] = f[1, 60:70]                             # flat central difference
    for tq in (2100.0, 2500.0, 2850.5, 3599.0, 3600.0):
        got = pchip4(f[0], f[1], f[2], f[3], times, tq, flavour)
        want = PchipInterpolator(times, f, axis=0)(tq)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(pchip4(f[0], f[1], f[2], f[3], times, 2100.0, flavour), f[1])
    np.testing.assert_array_equal(pchip4(f[0], f[1], f[2], f[3], times, 3600.0, flavour), f[2])
    # monotone data stay within the bracket values
    mid = pchip4(f[0, :50], f[1, :50], f[2, :50], f[3, :50], times, 2900.0, flavour)
    assert np.all((mid >= f[1, :50]) & (mid <= f[2, :50]))


def test_bf16_round_is_round_to_nearest_even():
    """The bfloat16 rounding the snapshot upload applies, against a value
    restatement: of the two bfloat16 neighbours of x (its float32 pattern
    truncated to 16 bits, and the next one up in magnitude) the nearer in
    exact float64 arithmetic, a tie going to the even pattern; overflow past
    the largest bfloat16 to infinity."""
    from oracle.oracle import bf16_round
    rng = np.random.default_rng(8)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000),
                        [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0, -0.0,
                         3.3e38, 1e-39]]).astype(np.float32)
    got = bf16_round(x)
    bits = x.view(np.uint32)
    lo = (bits & 0xFFFF0000).view(np.float32).astype(np.float64)
    hi_bits = (bits & 0xFFFF0000) + 0x10000
    hi = hi_bits.view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dl, dh = np.abs(xd - lo), np.abs(hi - xd)
    even_lo = ((bits >> 16) & 1) == 0
    want = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where(even_lo, lo, hi)))
    np.testing.assert_array_equal(got.astype(np.float64), want)
    assert np.array_equal(bf16_round(np.array([np.inf, -np.inf], np.float32)), [np.inf, -np.inf])
    assert np.isnan(bf16_round(np.array([np.nan], np.float32)))[0]
