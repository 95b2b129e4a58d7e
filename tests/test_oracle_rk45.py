"""The adaptive Dormand-Prince 5(4) step (integrators.py:6448 rk45_step,
:6233 _rk45_kernel_2d) in the oracle, pinned by what the mathematics fixes.

No reference golden vector holds an rk45 run.  The pins: on a uniform
zonal flow the right-hand side is constant along the path, so one substep
is exact and the step equals the closed form; on a smooth rotating flow,
steady and time-blended, rk45 at a tight tolerance agrees with RK4 on 64
substeps of the same field (a wrong tableau entry, stage time or blend
fraction breaks the agreement by orders of magnitude); loosening the
tolerance moves the end points by about the tolerance and no more.
"""
import numpy as np
import pytest

from kelp_amd import synth
from oracle.oracle import OracleMesh, step_args

DT = 10800.0


@pytest.fixture(scope="module")
def mesh():
    return synth.lattice_mesh(0.0, 8.0, 40.0, 46.0, 0.25, jitter=0.05, seed=3)


def _state(P, seed=1):
    rng = np.random.default_rng(seed)
    lon, lat = rng.uniform(2.0, 6.0, P), rng.uniform(42.0, 44.0, P)
    return dict(lon=lon, lat=lat, z=np.zeros(P), status=np.ones(P, np.int32),
                nn=np.full(P, -1, np.int32), pid=np.arange(P, dtype=np.int64))


def _rotating(mesh, phase=0.0):
    """Solid-body-like rotation about (4 E, 43 N) at the element centres."""
    x = (mesh.elem_center_lon - 4.0) * np.cos(np.radians(43.0))
    y = mesh.elem_center_lat - 43.0
    om = 1.0 * (1.0 + phase)
    return -om * y, om * x


@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
def test_uniform_zonal_flow_is_one_exact_substep(mesh, flavour):
    om = OracleMesh(mesh, flavour)
    U = 0.4
    f = dict(u_a=np.full(mesh.n_elements, U), v_a=np.zeros(mesh.n_elements), alpha0=0.0, uv_layout="element")
    st = _state(200)
    for direction in (1.0, -1.0):
        out = om.step(st, f, step_args(5, DT, direction, exact=2))
        assert (out["status"] == 1).all()
        z0 = np.maximum(np.cos(st["lat"] * 0.017453292519943295) * 111120.0, 1.0)
        np.testing.assert_allclose(out["lon"], st["lon"] + direction * U * DT / z0, rtol=0, atol=1e-11)
        np.testing.assert_array_equal(out["lat"], st["lat"])
        np.testing.assert_allclose(out["u"], direction * U, rtol=0, atol=1e-12)


def _rk4_fine(om, st, f, n, direction, inv_dtn):
    h = DT / n
    for k in range(n):
        t = k * h
        al = [min((t + c * h) * inv_dtn, 1.0) for c in (0.0, 0.5, 0.5, 1.0)]
        st = om.step(st, f, step_args(4, h, direction, exact=2, alphas=al))
    return st


@pytest.mark.parametrize("tinterp", [False, True])
@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_tight_tolerance_agrees_with_fine_rk4(mesh, tinterp, direction):
    om = OracleMesh(mesh, "glibc")
    ua, va = _rotating(mesh)
    f = dict(u_a=ua, v_a=va, alpha0=0.0, uv_layout="element")
    inv = 0.0
    if tinterp:
        ut, vt = _rotating(mesh, phase=0.5)
        f.update(u_t=ut, v_t=vt)
        inv = 1.0 / (2.0 * DT)
    st = _state(300)
    a = om.step(st, f, step_args(5, DT, direction, exact=2, rk45=dict(tol_m=1e-4, inv_dtn=inv)))
    b = _rk4_fine(om, st, f, 64, direction, inv)
    ok = (a["status"] == 1) & (b["status"] == 1)
    assert ok.mean() > 0.95
    dm = np.hypot((a["lon"][ok] - b["lon"][ok]) * 111120.0 * np.cos(np.radians(a["lat"][ok])),
                  (a["lat"][ok] - b["lat"][ok]) * 111120.0)
    assert dm.max() < 0.05, dm.max()
    # and the flow did move them by kilometres
    assert np.median(np.abs(a["lon"][ok] - st["lon"][ok])) * 111120.0 * 0.73 > 5000.0


def test_tolerance_bounds_the_error(mesh):
    om = OracleMesh(mesh, "glibc")
    ua, va = _rotating(mesh)
    f = dict(u_a=ua, v_a=va, alpha0=0.0, uv_layout="element")
    st = _state(300, seed=5)
    ref = _rk4_fine(om, st, f, 64, 1.0, 0.0)
    errs = []
    for tol in (10.0, 1.0, 0.01):
        a = om.step(st, f, step_args(5, DT, 1.0, exact=2, rk45=dict(tol_m=tol)))
        ok = (a["status"] == 1) & (ref["status"] == 1)
        d = np.hypot((a["lon"][ok] - ref["lon"][ok]) * 111120.0 * np.cos(np.radians(a["lat"][ok])),
                     (a["lat"][ok] - ref["lat"][ok]) * 111120.0)
        errs.append(d.max())
        assert d.max() < 3.0 * tol + 0.05, (tol, d.max())
    assert errs[2] <= errs[0]
