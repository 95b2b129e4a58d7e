"""The vertical kick with an Az field (integrators.py:842 with vel.az,
numerics/diffusion.py:378 az_column_value_gradient, :427
_loop_az_kick_zlevel), pinned in the oracle by what the mathematics fixes.

No reference golden vector holds an Az-field run.  The pins: on a profile
linear in depth the column value is the line and the gradient its slope,
so one step moves z by exactly s|dt| + sqrt(2 (a + s z)|dt|) xi with the
oracle's own normal xi; a negative value switches the whole kick off; a
field constant in depth is the constant source; and the well-mixed
condition (Visser 1997): a uniform column stays uniform under a strongly
depth-dependent Az -- with the Ito drift dropped the particles pile up where
Az is small (a 2x excess near the surface in the same setting), with it
every fifth of the column keeps its share within 25 %.
"""
import ctypes

import numpy as np
import pytest

from kelp_amd.mesh import build_grid_hash, make_mesh
from oracle.oracle import OracleMesh, lib, step_args

H = 20.0
ZL = np.linspace(0.0, H, 11)


def _column_mesh():
    xs = np.arange(0.0, 1.0 + 1e-9, 0.1)
    gx, gy = np.meshgrid(xs, 40.0 + xs)
    nx = xs.size
    elems = []
    for j in range(nx - 1):
        for i in range(nx - 1):
            n0 = j * nx + i
            elems += [[n0, n0 + 1, n0 + nx + 1], [n0, n0 + nx + 1, n0 + nx]]
    mesh = make_mesh("az_column", gx.ravel(), gy.ravel(), np.array(elems, np.int32),
                     depth=np.full(gx.size, H))
    mesh.sigma = ZL.copy()
    build_grid_hash(mesh)
    return mesh


@pytest.fixture(scope="module")
def column():
    mesh = _column_mesh()
    K, N = ZL.size, mesh.n_nodes
    zero = dict(u_a=np.zeros((K - 1, N)), v_a=np.zeros((K - 1, N)), w_a=np.zeros((K, N)), alpha0=0.0)
    return mesh, zero


def _state(P, z, seed=0):
    rng = np.random.default_rng(seed)
    lon, lat = rng.uniform(0.15, 0.85, P), rng.uniform(40.15, 40.85, P)
    return dict(lon=lon, lat=lat, z=np.asarray(z, np.float64), status=np.ones(P, np.int32),
                nn=np.full(P, -1, np.int32), pid=np.arange(P, dtype=np.int64) + 77)


def _xi(L, seed, pid, step_i):
    a, b = ctypes.c_double(), ctypes.c_double()
    L.or_normal_pair(seed, 0, int(pid), step_i, 1, ctypes.byref(a), ctypes.byref(b))
    return a.value


@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_linear_profile_closed_form(column, flavour, direction):
    mesh, zero = column
    om = OracleMesh(mesh, flavour)
    a0, s = 2e-4, 5e-4
    az = np.repeat((a0 + s * ZL)[:, None], mesh.n_nodes, axis=1)
    P, dt, seed = 400, 30.0, 11
    z0 = np.random.default_rng(1).uniform(4.0, 16.0, P)
    st = _state(P, z0)
    args = step_args(1, dt, direction, is3d=True, kick=dict(kh_x=0.0, kh_y=0.0, seed=seed, az_field=az),
                     step_i=3)
    out = om.step(st, zero, args)
    L = lib(flavour)
    xi = np.array([_xi(L, seed, p, 3) for p in st["pid"]])
    want = z0 + s * dt + np.sqrt(2.0 * (a0 + s * z0) * dt) * xi
    inside = (want > 0.5) & (want < H - 0.5)
    assert inside.mean() > 0.9
    np.testing.assert_allclose(out["z"][inside], want[inside], rtol=0, atol=1e-12)
    # no horizontal motion: the longitude only passes the (lon + 180) % 360 fold
    np.testing.assert_allclose(out["lon"], st["lon"], rtol=0, atol=1e-12)


def test_negative_value_switches_the_kick_off(column):
    mesh, zero = column
    om = OracleMesh(mesh, "glibc")
    az = np.full((ZL.size, mesh.n_nodes), -1e-3)
    z0 = np.linspace(1.0, 19.0, 50)
    st = _state(50, z0)
    out = om.step(st, zero, step_args(1, 60.0, 1.0, is3d=True,
                                      kick=dict(kh_x=0.0, kh_y=0.0, seed=3, az_field=az)))
    np.testing.assert_array_equal(out["z"], z0)


def test_depth_constant_field_is_the_constant_source(column):
    mesh, zero = column
    om = OracleMesh(mesh, "glibc")
    c = 3e-3
    z0 = np.random.default_rng(2).uniform(0.0, H, 300)
    st = _state(300, z0)
    kick = dict(kh_x=0.0, kh_y=0.0, seed=9)
    a = om.step(st, zero, step_args(1, 60.0, 1.0, is3d=True, step_i=5,
                                    kick=dict(kick, az_field=np.full((ZL.size, mesh.n_nodes), c))))
    b = om.step(st, zero, step_args(1, 60.0, 1.0, is3d=True, step_i=5, kick=dict(kick, az_const=c)))
    np.testing.assert_allclose(a["z"], b["z"], rtol=0, atol=1e-12)
    assert not np.array_equal(b["z"], z0)


def test_well_mixed_condition(column):
    mesh, zero = column
    om = OracleMesh(mesh, "glibc")
    az = np.repeat((1e-4 + 0.01 * ZL / H)[:, None], mesh.n_nodes, axis=1)
    P = 2000
    st = _state(P, np.random.default_rng(4).uniform(0.0, H, P), seed=4)
    for k in range(600):
        st = om.step(st, zero, step_args(1, 50.0, 1.0, is3d=True, step_i=k,
                                         kick=dict(kh_x=0.0, kh_y=0.0, seed=21, az_field=az)))
    assert (st["status"] == 1).all()
    h, _ = np.histogram(st["z"], bins=5, range=(0.0, H))
    assert np.abs(h / (P / 5) - 1.0).max() < 0.25, h
