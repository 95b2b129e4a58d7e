"""The reference-shaped integrator registry (``rk2_step(state, vel, mesh,
dt, time_dir, ...)``) on the device against the oracle (GPU).

Bar: bit-identical to the oracle built with the device's math (kmath):
positions, depth, status, nearest node, failure code, velocities.
"""
import numpy as np
import pytest

from kelp_amd import integrators as I
from kelp_amd import synth
from kelp_amd.diffusion import DiffusionStepArgs
from kelp_amd.fields import LazyLerpField, VelocityField
from kelp_amd.mesh import set_exact_nearest
from kelp_amd.particles import ParticleState, make_state
from oracle.oracle import OracleMesh, step_args

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float64).view(np.int64)


def _assert_same(st: ParticleState, ref: dict, is3d: bool):
    keys = ("lon", "lat", "u", "v") + (("z", "w") if is3d else ())
    for k in keys:
        assert np.array_equal(_bits(getattr(st, k)), _bits(ref[k])), k
    assert np.array_equal(st.status, ref["status"])
    assert np.array_equal(st.nearest_node, ref["nn"])
    assert np.array_equal(st.failure_reason, ref["fr"])


@pytest.mark.parametrize("exact", [True, False])
def test_rk2_2d_registry_bitwise(exact):
    w = synth.double_gyre_workload(n_particles=3000)
    mesh = w.mesh
    set_exact_nearest(mesh, exact)
    vel = VelocityField(u=w.u, v=w.v, uv_layout="node")
    step = I.get_integrator("rk2")
    st = make_state(w.lon, w.lat)
    om = OracleMesh(mesh, "kmath")
    ref = dict(lon=w.lon.copy(), lat=w.lat.copy(), status=np.ones(w.lon.size, np.int32),
               nn=np.full(w.lon.size, -1, np.int32))
    for _ in range(20):
        st = step(st, vel, mesh, w.dt, 1.0, uv_interp="barycentric")
        ref = om.step2d(ref, 2, w.u, w.v, w.dt, 1.0, exact=1 if exact else 2)
    _assert_same(st, ref, False)


def test_rk4_3d_tinterp_kick_registry_bitwise():
    mesh = synth.sphere_mesh(20000)
    snaps = [synth.jet_rossby_snapshot(mesh, s * synth.SNAP_DT) for s in range(3)]
    vels = [VelocityField(u=s[0], v=s[1], w=s[2], uv_layout="node") for s in snaps]
    lon, lat, z = synth.wet_area_seeds(mesh, 4000)
    st = make_state(lon, lat, z)
    om = OracleMesh(mesh, "kmath")
    ref = dict(lon=lon, lat=lat, z=z, status=np.ones(lon.size, np.int32),
               nn=np.full(lon.size, -1, np.int32), pid=np.arange(lon.size, dtype=np.int64) + 77)
    dt = 1800.0
    kick = dict(kh_x=10.0, kh_y=10.0, seed=99)
    for k in range(12):
        ia, ib, a0, it, dtn = synth.snapshot_bracket(k * dt, synth.SNAP_DT, len(snaps))
        vel = vels[ia] if ib < 0 else LazyLerpField(vels[ia], vels[ib], a0)
        d = DiffusionStepArgs("uniform", 10.0, 10.0, 99, 0, k, pid_offset=77)
        st = I.rk4_step(st, vel, mesh, dt, 1.0, uv_interp="barycentric", w_interp="barycentric",
                        vel_next=vels[it], dt_to_next=dtn, diffusion=d)
        A, T = snaps[ia], snaps[it]
        f = dict(u_a=A[0], v_a=A[1], w_a=A[2], u_t=T[0], v_t=T[1], w_t=T[2], alpha0=a0)
        if ib >= 0:
            f.update(u_b=snaps[ib][0], v_b=snaps[ib][1], w_b=snaps[ib][2])
        al = I.stage_alphas(4, dt, dtn)
        ref = om.step(ref, f, step_args(4, dt, 1.0, is3d=True, w_method=2, exact=2, alphas=al,
                                        kick=kick, step_i=k))
    _assert_same(st, ref, True)


def test_registry_element_layout_matches_reference_golden():
    """rk2_step with element-centre u/v (uv_interp="leastsq", the reference
    default) and leastsq W, driven like the reference's golden generator
    (tests/_gen_golden_3d.py:200-219), against its stored output."""
    import os
    from .golden import reference_cases as rc
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_golden_3d",
                                "fesom2_elem_rk2_unfused_lerp.npz"), allow_pickle=False)
    mesh = rc.elem_mesh()
    a, b = (VelocityField(u=f[0], v=f[1], w=f[2], uv_layout="element")
            for f in (rc.elem_field(mesh, 0.0), rc.elem_field(mesh, 1.0)))
    lon, lat, z = rc.elem_seeds()
    st = make_state(lon, lat, z)
    T = rc.ELEM_STEPS * rc.ELEM_DT
    rows = []
    for i in range(rc.ELEM_STEPS):
        s = i * rc.ELEM_DT
        alpha = s / T
        vel = a if abs(alpha) < 1e-12 else LazyLerpField(a, b, alpha)
        st = I.rk2_step(st, vel, mesh, rc.ELEM_DT, 1.0, uv_interp="leastsq", w_interp="leastsq",
                        vel_next=b, dt_to_next=T - s)
        rows.append(st)
    for k in ("lon", "lat", "z", "u", "v", "w"):
        got = np.stack([getattr(r, k) for r in rows])
        np.testing.assert_allclose(got, gold[k], rtol=1e-9, atol=1e-9, err_msg=k)
    np.testing.assert_array_equal(np.stack([r.status for r in rows]), gold["status"])
    np.testing.assert_array_equal(np.stack([r.failure_reason for r in rows]), gold["failure_reason"])
