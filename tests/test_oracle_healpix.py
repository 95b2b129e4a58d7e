"""The native HEALPix reader path (readers/healpix.py:1295 load_mesh, :1872
read_velocities) against the reference's run-level golden
``healpix_3d_rk2`` (tests/_gen_golden_3d.py:625), on the CPU: the store
restated as arrays (finite-prefix columns of 1-3 layers, NaN land and
padding, W on the interfaces, float32 velocities), the dual mesh built from
its wet mask, and the tracking loop's clock with the oracle doing every
step.  Bar: the reference's cross-environment rule, integers exact, floats
within rtol = atol = 1e-9.
"""
import os

import numpy as np
import pytest

from kelp_amd.mesh import exact_nearest_on
from kelp_amd.readers import get_reader
from kelp_amd.tracking import (_C, mask_record, resolve_time_bracket, snapshot_band,
                               step_end_on_record, vel_payload)
from oracle.oracle import OracleMesh, step_args

from .golden import reference_cases as rc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_golden_3d")


def run_zlevel_oracle(case, reader_name, flavour="glibc"):
    """The case's tracking loop (z-level, time-interpolated) with the oracle
    doing each step; the masked records ``(particle, record)``."""
    mesh, ds, mode, numerics = case["mesh"], case["ds"], case["mode"], case["numerics"]
    order = {"euler": 1, "rk2": 2, "rk4": 4}[numerics.integrator]
    w_method = {"nearest": 0, "leastsq": 1, "barycentric": 2}[numerics.w_interp]
    reader = get_reader(reader_name)
    file_times = np.asarray(case["file_times"], dtype=np.float64)
    n = case["n_time_total"]
    fields = {i: reader.read_velocities(ds, i, mesh, mode, velocity_float32=numerics.velocity_float32)
              for i in range(n)}
    om = OracleMesh(mesh, flavour)
    direction = 1 if mode.time_dir >= 0 else -1
    dt = float(numerics.dt)
    snap = snapshot_band(dt) if order > 1 else 5e-10 * max(dt, 1.0)
    init = case["init"]
    P = init.n_particles
    state = dict(lon=init.lon.copy(), lat=init.lat.copy(), z=init.z.copy(),
                 status=np.ones(P, np.int32), nn=np.full(P, -1, np.int32), fr=np.zeros(P, np.int8))
    sim_time = file_times[case["start_time_idx"]]
    exact = 1 if exact_nearest_on(mesh) else 2
    rows = []
    for _ in range(case["n_steps"]):
        br = resolve_time_bracket(sim_time, file_times, n, "stop", direction, snap)
        if br is None:
            break
        pay = vel_payload(br, direction)
        step_len, sim_time_end, dtn = dt, None, pay.dt_to_next
        if order > 1 and dtn is not None and abs(dt - dtn) <= snap:
            step_len = dtn
            sim_time_end = step_end_on_record(sim_time, br[4], dtn, direction, file_times, n, "stop")
        elif dtn is not None and dtn < dt <= dtn + snap:
            dtn = dt
        if dtn is not None and step_len > dtn:
            pay.it, dtn = None, None
        alphas = [0.0] * 4
        if dtn is not None:
            alphas = [c * step_len / dtn for c in _C[order]] + [0.0] * (4 - order)
        fa = fields[pay.ia]
        flds = dict(u_a=fa.u, v_a=fa.v, w_a=fa.w, alpha0=pay.alpha0, uv_layout=fa.uv_layout)
        if pay.ib is not None:
            fb = fields[pay.ib]
            flds.update(u_b=fb.u, v_b=fb.v, w_b=fb.w)
        if pay.it is not None:
            ft = fields[pay.it]
            flds.update(u_t=ft.u, v_t=ft.v, w_t=ft.w)
        args = step_args(order, step_len, float(direction), is3d=True, w_method=w_method, exact=exact,
                         alphas=alphas)
        prev = state["status"].copy()
        state = om.step(state, flds, args)
        rows.append(mask_record(state, prev, direction, True))
        sim_time = sim_time_end if sim_time_end is not None else sim_time + dt * direction
    return {k: np.stack([r[k] for r in rows], axis=1) for k in ("lon", "lat", "z", "status", "fr",
                                                               "u", "v", "w")}


def test_store_reads_like_the_reference_reader():
    """The reader's contract on the golden's store: wet = finite surface
    u/v, depth = the interface below the last finite layer, W negated (the
    store is positive up) and zero below each column, u/v zero below their
    prefix, float32."""
    case = rc.healpix_case(1.0)
    mesh, ds = case["mesh"], case["ds"]
    vf = get_reader("healpix_arrays").read_velocities(ds, 1, mesh, case["mode"], velocity_float32=True)
    assert vf.u.dtype == np.float32 and vf.u.shape == (3, mesh.n_nodes) and vf.w.shape == (4, mesh.n_nodes)
    lev = np.asarray(rc.HP_LEVELS)
    nfin = np.searchsorted(lev, mesh.depth)           # finite layers per node column
    assert set(np.unique(mesh.depth)) == {10.0, 30.0, 70.0}
    for k in range(3):
        assert np.all((vf.u[k] != 0.0) == (nfin > k))
    for k in range(4):
        np.testing.assert_array_equal(vf.w[k], np.where(nfin >= k, np.float32(-1e-4), 0.0))
    np.testing.assert_array_equal(vf.eta, 0.25)


@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_healpix_oracle_matches_reference_golden(direction, flavour):
    gold = np.load(os.path.join(GOLD, "healpix_3d_rk2" + ("_bwd" if direction < 0 else "") + ".npz"),
                   allow_pickle=False)
    rec = run_zlevel_oracle(rc.healpix_case(direction), "healpix_arrays", flavour)
    np.testing.assert_array_equal(rec["status"], gold["status"])
    np.testing.assert_array_equal(rec["fr"], gold["failure_reason"])
    for k in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_allclose(rec[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)
