"""``run_tracking`` end to end against the reference's golden
``run_tracking_fesom2_multiyear_node`` (tests/_gen_golden_3d.py:740): daily
means in two yearly files read as one axis, shifted to the window
midpoints, RK2 across the year boundary on the FESOM2 unit square.  On the
CPU the record times come from ``run_tracking``'s own resolution and the
oracle does every step; the trajectory variables meet the reference's bar
(integers exact, floats within rtol = atol = 1e-9).  The writer's
attributes and the NetCDF layout are not compared (no NetCDF here).
"""
import os

import numpy as np
import pytest

from kelp_amd.mesh import set_exact_nearest
from kelp_amd.tracking import resolve_record_times

from .golden import reference_cases as rc
from .test_oracle_healpix import run_zlevel_oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_golden_3d")


def test_record_times_are_the_window_midpoints():
    case = rc.multiyear_case()
    gold = np.load(os.path.join(GOLD, "run_tracking_fesom2_multiyear_node.npz"), allow_pickle=False)
    ft, fdt, dt, avg = resolve_record_times(case["config"], case["ds"])
    assert avg == "mean" and fdt == 86400.0 and dt == 21600.0
    np.testing.assert_array_equal(ft[1:], np.arange(1, ft.size) * 86400.0 - 43200.0)
    np.testing.assert_array_equal(gold["release_time"], ft[362])
    np.testing.assert_array_equal(gold["time"], ft[362] + np.arange(1, 13) * 21600.0)


@pytest.mark.parametrize("flavour", ["glibc", "kmath"])
def test_multiyear_oracle_matches_reference_golden(flavour):
    case = rc.multiyear_case()
    cfg = case["config"]
    gold = np.load(os.path.join(GOLD, "run_tracking_fesom2_multiyear_node.npz"), allow_pickle=False)
    ft, _, _, _ = resolve_record_times(cfg, case["ds"])
    set_exact_nearest(case["mesh"], True)
    run = dict(mesh=case["mesh"], ds=case["ds"], mode=cfg.mode, numerics=cfg.numerics, init=case["init"],
               file_times=ft, n_time_total=ft.size, start_time_idx=cfg.time.start_time_idx,
               n_steps=cfg.time.n_steps)
    rec = run_zlevel_oracle(run, "fesom2_arrays", flavour)
    assert (gold["status"] != 1).any() and (gold["failure_reason"] != 0).any()
    np.testing.assert_array_equal(rec["status"], gold["status"])
    np.testing.assert_array_equal(rec["fr"], gold["failure_reason"])
    for k, g in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("u", "vel_u"), ("v", "vel_v"),
                 ("w", "vel_w")):
        np.testing.assert_allclose(rec[k], gold[g], rtol=1e-9, atol=1e-9, err_msg=k)


def test_output_file_matches_the_reference_file(tmp_path):
    """Trajectory output format 2 written to a file (classic NetCDF here)
    against the reference's own output file of this run, as the golden
    holds it: every variable's values (trajectories from the oracle, the
    sampled temp and release data from the golden's run), dimensions and
    attributes, and the run's global attributes; only the build identity
    (kelp_version, history, w_contract) differs by design."""
    from kelp_amd.writers import read_trajectories, run_attributes, write_trajectories
    from kelp_amd.tracking import mask_record  # noqa: F401  (records come masked)
    from scipy.io import netcdf_file
    case = rc.multiyear_case()
    cfg = case["config"]
    gold = np.load(os.path.join(GOLD, "run_tracking_fesom2_multiyear_node.npz"), allow_pickle=False)
    ft, _, _, avg = resolve_record_times(cfg, case["ds"])
    set_exact_nearest(case["mesh"], True)
    run = dict(mesh=case["mesh"], ds=case["ds"], mode=cfg.mode, numerics=cfg.numerics, init=case["init"],
               file_times=ft, n_time_total=ft.size, start_time_idx=cfg.time.start_time_idx,
               n_steps=cfg.time.n_steps)
    rec = run_zlevel_oracle(run, "fesom2_arrays", "glibc")
    arrays = dict(lon=rec["lon"], lat=rec["lat"], z=rec["z"], status=rec["status"],
                  failure_reason=rec["fr"], vel_u=rec["u"], vel_v=rec["v"], vel_w=rec["w"],
                  sample__temp=gold["temp"], time=ft[362] + np.arange(1, 13) * 21600.0,
                  release_lon=case["init"].lon, release_lat=case["init"].lat, release_z=case["init"].z,
                  release_time=np.full(6, ft[362]))
    path = str(tmp_path / "traj.nc")
    write_trajectories(path, arrays, time_units=str(gold["attr__time__units"]), calendar="standard",
                       attrs=run_attributes(cfg, 362, 12, avg, case["mesh"]))
    got = read_trajectories(path)
    assert set(got) == {k for k in gold.files if not k.startswith(("attr__", "dims__"))}
    for k in got:
        if got[k].dtype.kind == "f":
            np.testing.assert_allclose(got[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)
        else:
            np.testing.assert_array_equal(got[k], gold[k], err_msg=k)
    with netcdf_file(path, "r", mmap=False) as nc:
        for name, var in nc.variables.items():
            assert " ".join(var.dimensions) == str(gold["dims__" + name]), name
            want = {k.split("__", 2)[2]: str(gold[k]) for k in gold.files
                    if k.startswith(f"attr__{name}__")}
            have = {k: (str(np.asarray(v)) if not isinstance(v, bytes) else v.decode())
                    for k, v in var._attributes.items()}
            assert have == want, (name, have, want)
        build_identity = {"kelp_version", "history", "w_contract"}
        want = {k[len("attr__"):]: str(gold[k]) for k in gold.files
                if k.startswith("attr__") and k.count("__") == 1}
        have = {k: (v.decode() if isinstance(v, bytes) else str(np.asarray(v).item() if np.ndim(v) == 0 or
                                                                 np.size(v) == 1 else np.asarray(v)))
                for k, v in nc._attributes.items()}
        for k in set(want) - build_identity:
            assert k in have and have[k] == want[k], (k, have.get(k), want[k])
