"""The device tracking loop (run_single_realization) against the
reference's run-level golden vectors (GPU).

``beaching_resuspension_rk2`` is the reference's run of its own tracking
loop (run_single_realization with its NetCDF trajectory writer) on a node
dataset whose tide dries and rewets a shallow strip: time-blended 3-D RK2,
depth + stochastic beaching on depth + eta, auto-or-stochastic
resuspension, records with output format 2's velocity masks.  Bar: the
reference's cross-environment rule, integers exact, floats within
rtol = atol = 1e-9 (NaN where the record masks them).
"""
import os

import numpy as np
import pytest

from kelp_amd.readers import get_reader
from kelp_amd.tracking import run_single_realization

from .golden import reference_cases as rc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_golden_3d")


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_beaching_run_matches_reference_golden(gpu_ctx, direction):
    case = rc.beaching_case(direction)
    gold = np.load(os.path.join(GOLD, "beaching_resuspension_rk2" + ("_bwd" if direction < 0 else "")
                                + ".npz"), allow_pickle=False)
    rec = run_single_realization(
        case["mesh"], case["ds"], get_reader("fesom2_arrays"), case["init"], case["mode"],
        case["numerics"], start_time_idx=case["start_time_idx"], n_steps=case["n_steps"],
        n_time_total=case["n_time_total"], file_times=case["file_times"],
        file_dt=case["file_dt"], ctx=gpu_ctx).as_arrays()
    assert (gold["status"] == 2).any() and ((gold["status"][:, 1:] == 1) & (gold["status"][:, :-1] == 2)).any()
    np.testing.assert_array_equal(rec["status"], gold["status"])
    np.testing.assert_array_equal(rec["failure_reason"], gold["failure_reason"])
    for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("u", "vel_u"), ("v", "vel_v"),
                   ("w", "vel_w")):
        np.testing.assert_allclose(rec[src], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_sampling_run_matches_reference_golden(gpu_ctx, direction):
    """Record-point sampling of an interface and a mid-layer scalar
    (leastsq), re-blended to each record's instant."""
    case = rc.sampling_case(direction)
    gold = np.load(os.path.join(GOLD, "scalar_sampling_rk2" + ("_bwd" if direction < 0 else "")
                                + ".npz"), allow_pickle=False)
    rec = run_single_realization(
        case["mesh"], case["ds"], get_reader("fesom2_arrays"), case["init"], case["mode"],
        case["numerics"], start_time_idx=case["start_time_idx"], n_steps=case["n_steps"],
        n_time_total=case["n_time_total"], file_times=case["file_times"],
        file_dt=case["file_dt"], sample_vars=case["sample_vars"], ctx=gpu_ctx).as_arrays()
    np.testing.assert_array_equal(rec["status"], gold["status"])
    for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("u", "vel_u"), ("w", "vel_w"),
                   ("sample__s_lev", "sample__s_lev"), ("sample__s_mid", "sample__s_mid")):
        np.testing.assert_allclose(rec[src], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)
