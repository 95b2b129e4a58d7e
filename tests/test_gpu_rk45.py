"""The adaptive Dormand-Prince 5(4) step (order 5, integrators.py:6448
rk45_step) on the device against the oracle (GPU): 2-D u/v at element
centres, a rotating flow fast enough to push particles across the lattice's
edges (beached there by on_failure), steady and blended through a bracket towards a
target snapshot, forward and backward, guard on and off, with the Kh walk
and beaching after the step.  Bar: bit-identical to the oracle built with
the device's math (kmath) in every output.
"""
import numpy as np
import pytest

from kelp_amd import synth
from oracle.oracle import OracleMesh, step_args

pytestmark = pytest.mark.gpu

DT = 10800.0


@pytest.fixture(scope="module")
def world():
    mesh = synth.lattice_mesh(0.0, 8.0, 40.0, 46.0, 0.25, jitter=0.05, seed=3)
    x = (mesh.elem_center_lon - 4.0) * np.cos(np.radians(43.0))
    y = mesh.elem_center_lat - 43.0
    snaps = [(-(1.0 + 0.4 * s) * y + 0.05 * s, (1.0 + 0.4 * s) * x) for s in range(3)]
    rng = np.random.default_rng(9)
    P = 4000
    lon, lat = rng.uniform(0.3, 7.7, P), rng.uniform(40.3, 45.7, P)
    pid = np.arange(P, dtype=np.int64) + 500
    return mesh, snaps, lon, lat, pid


@pytest.mark.parametrize("exact", [1, 2])
@pytest.mark.parametrize("blend,direction", [(False, 1.0), (True, 1.0), (True, -1.0)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rk45_bitwise(gpu_ctx, world, exact, blend, direction, dtype):
    mesh, snaps, lon, lat, pid = world
    gpu_ctx.upload_mesh(mesh)
    for s, (u, v) in enumerate(snaps):
        gpu_ctx.upload_field(s, u.astype(dtype), v.astype(dtype), uv_layout="element")
    om = OracleMesh(mesh, "kmath")
    kick = dict(kh_x=30.0, kh_y=10.0, seed=77)
    beach = dict(criteria=["on_failure"])
    p = gpu_ctx.particles(lon, lat, pid=pid)
    st = dict(lon=lon, lat=lat, z=np.zeros(lon.size), status=np.ones(lon.size, np.int32),
              nn=np.full(lon.size, -1, np.int32), pid=pid)
    c = lambda a: a.astype(dtype)
    for k in range(4):
        if blend:
            a0, inv = 0.3 + 0.1 * k, 1.0 / (2.0 * DT)
            f = dict(u_a=c(snaps[0][0]), v_a=c(snaps[0][1]), u_b=c(snaps[1][0]), v_b=c(snaps[1][1]),
                     u_t=c(snaps[2][0]), v_t=c(snaps[2][1]), alpha0=a0, uv_layout="element")
            sb, stt = 1, 2
        else:
            a0, inv = 0.0, 0.0
            f = dict(u_a=c(snaps[0][0]), v_a=c(snaps[0][1]), alpha0=0.0, uv_layout="element")
            sb, stt = -1, -1
        rk = dict(tol_m=0.5, inv_dtn=inv)
        st = om.step(st, f, step_args(5, DT, direction, exact=exact, kick=kick, beach=beach, step_i=k, rk45=rk))
        gpu_ctx.step(p, 5, 0, DT, direction, slot_b=sb, slot_t=stt, alpha0=a0, exact_nearest=exact,
                     kick=kick, beach=beach, step_i=k, rk45=rk)
    gpu_ctx.synchronize()
    h = p.to_host()
    # some particles crossed the lattice's edge (beached there by on_failure)
    assert 0.5 < (st["status"] == 1).mean() < 0.99
    for key in ("status", "fr", "nn"):
        np.testing.assert_array_equal(h[key], st[key], err_msg=key)
    for key in ("lon", "lat", "u", "v"):
        np.testing.assert_array_equal(h[key].view(np.int64), st[key].view(np.int64), err_msg=key)


def test_rk45_refuses_what_it_does_not_take(gpu_ctx, world):
    from kelp_amd._native import KelpError
    mesh, snaps, lon, lat, pid = world
    gpu_ctx.upload_mesh(mesh)
    un = np.zeros(mesh.n_nodes)
    gpu_ctx.upload_field(1, un, un)
    p = gpu_ctx.particles(lon[:10], lat[:10], pid=pid[:10])
    with pytest.raises(KelpError, match="rk45"):
        gpu_ctx.step(p, 5, 1, DT, 1.0)       # node-layout u/v
