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


def test_activate_kernel_matches_release_rule(gpu_ctx):
    """kelp_activate against core/release.py:112 (activate_released), with
    the particles in shuffled pid order on the device."""
    from kelp_amd import synth
    from kelp_amd.tracking import activate_released
    mesh = synth.sphere_mesh(2000)
    gpu_ctx.upload_mesh(mesh)
    rng = np.random.default_rng(3)
    P, off = 5000, 700
    rel = rng.integers(0, 20, P).astype(np.int64)
    status = rng.choice([1, 2, 3, 4, 4, 4], P).astype(np.int32)
    perm = rng.permutation(P)
    lon, lat = np.zeros(P), np.zeros(P)
    p = gpu_ctx.particles(lon, lat, None, status=status[perm], pid=(off + perm).astype(np.int64))
    torch = gpu_ctx.torch
    rel_dev = torch.as_tensor(rel, device=p.lon.device)
    prev = torch.empty_like(p.status)
    want = status.copy()
    for g in (0, 5, 5, 19, 30):
        gpu_ctx.activate(p, rel_dev, off, g, prev)
        gpu_ctx.synchronize()
        want = activate_released(want, rel, g)
        np.testing.assert_array_equal(p.status.cpu().numpy(), want[perm])
        np.testing.assert_array_equal(prev.cpu().numpy(), want[perm])


@pytest.mark.parametrize("direction,is3d", [(1, True), (-1, True), (1, False)])
def test_record_staging_equals_host_masking(gpu_ctx, direction, is3d):
    """kelp_record_stage against the host record (core/release.py:132 and
    :170 masks, the backward flip, particle-id order): every variable bit
    for bit, NaN signs included, with the particles shuffled on the device,
    statuses of every kind, sampled scalars, and a block smaller than the
    number of records (three flushes)."""
    from kelp_amd import synth
    from kelp_amd.tracking import RecordStager, TrajectoryRecord, mask_record
    mesh = synth.sphere_mesh(2000)
    gpu_ctx.upload_mesh(mesh)
    rng = np.random.default_rng(5)
    P, off = 3000, 40
    perm = rng.permutation(P)
    torch = gpu_ctx.torch
    rec = TrajectoryRecord()
    want = []
    stager = None
    for r in range(7):
        status = rng.choice([1, 1, 1, 2, 3, 4], P).astype(np.int32)
        prev = rng.choice([1, 1, 1, 2, 4], P).astype(np.int32)
        vals = {k: rng.normal(size=P) for k in ("lon", "lat", "z", "u", "v", "w")}
        fr = rng.integers(0, 5, P).astype(np.int8)
        p = gpu_ctx.particles(vals["lon"][perm], vals["lat"][perm], vals["z"][perm] if is3d else None,
                              status=status[perm], fr=fr[perm], pid=(off + perm).astype(np.int64))
        p.u.copy_(torch.as_tensor(vals["u"][perm]))
        p.v.copy_(torch.as_tensor(vals["v"][perm]))
        if p.w is not None:
            p.w.copy_(torch.as_tensor(vals["w"][perm]))
        if stager is None:
            stager = RecordStager(gpu_ctx, p, off, rec, ["t", "s"], max_records=3)
            assert stager.R == 3
        samp = torch.as_tensor(rng.normal(size=(2, P))[:, perm].copy(), device=p.lon.device)
        prev_dev = torch.as_tensor(prev[perm], device=p.lon.device)
        torch.cuda.current_stream().synchronize()     # the fills above ran on torch's stream
        stager.append(p, prev_dev, direction, samp, float(r))
        h = p.to_host()
        order = np.argsort(h["pid"], kind="stable")
        h = {k: v[order] for k, v in h.items()}
        m = mask_record(h, prev, direction, is3d)
        row = dict(lon=m["lon"], lat=m["lat"], z=m["z"], status=m["status"], failure_reason=m["fr"],
                   vel_u=m["u"], vel_v=m["v"], vel_w=m["w"], sample__t=samp.cpu().numpy()[0][order],
                   sample__s=samp.cpu().numpy()[1][order])
        want.append(row)
    stager.flush()
    got = rec.as_arrays()
    assert got["time"].tolist() == [float(r) for r in range(7)]
    for k in want[0]:
        if k == "vel_w" and "vel_w" not in got:
            continue
        w = np.stack([row[k] for row in want], axis=1)
        g = got[k]
        assert g.dtype == w.dtype or g.dtype.kind == w.dtype.kind, k
        if g.dtype.kind == "f":
            np.testing.assert_array_equal(g.view(np.int64), w.view(np.int64), err_msg=k)
        else:
            np.testing.assert_array_equal(g, w, err_msg=k)


def test_staggered_release_run(gpu_ctx):
    """A release schedule in the device loop: dormant particles are masked
    and do not move, particles released at step 0 follow the run without a
    schedule bit for bit (the RNG is keyed by pid), later ones start from
    their seed position at their release step."""
    case = rc.beaching_case(1.0)
    init = case["init"]
    P = init.n_particles
    rel = np.zeros(P, np.int64)
    rel[1::3] = 4
    rel[2::3] = 9

    def run(state):
        return run_single_realization(
            case["mesh"], case["ds"], get_reader("fesom2_arrays"), state, case["mode"],
            case["numerics"], start_time_idx=case["start_time_idx"], n_steps=case["n_steps"],
            n_time_total=case["n_time_total"], file_times=case["file_times"],
            file_dt=case["file_dt"], ctx=gpu_ctx).as_arrays()

    status0 = init.status.copy()
    status0[rel > 0] = 4
    base = run(init)
    st = run(init.replace(status=status0, release_step=rel))
    now = rel == 0
    for k in ("lon", "lat", "z", "status", "vel_u", "vel_w"):
        np.testing.assert_array_equal(st[k][now], base[k][now], err_msg=k)
    for r in (4, 9):
        sel = rel == r
        # record j holds the state after global step j (records every step)
        assert (st["status"][sel, :r] == 4).all()
        assert np.isnan(st["lon"][sel, :r]).all() and np.isnan(st["vel_u"][sel, :r]).all()
        assert (st["status"][sel, r] != 4).all()
        moved = st["lon"][sel, r] != init.lon[sel]
        assert moved[st["status"][sel, r] == 1].all()


def test_resort_every_step_leaves_records_unchanged(gpu_ctx):
    """The tracking loop's periodic re-sort (resort_every=1) with staggered
    releases, beaching and resuspension: every record -- positions, status,
    failure reasons and the velocities the record masks keep or blank from
    the step-start status -- is bit-identical to the run that never sorts."""
    case = rc.beaching_case(1.0)
    init = case["init"]
    P = init.n_particles
    rel = np.zeros(P, np.int64)
    rel[1::3] = 3
    rel[2::3] = 7
    status0 = init.status.copy()
    status0[rel > 0] = 4
    state = init.replace(status=status0, release_step=rel)

    def run(every):
        return run_single_realization(
            case["mesh"], case["ds"], get_reader("fesom2_arrays"), state, case["mode"],
            case["numerics"], start_time_idx=case["start_time_idx"], n_steps=case["n_steps"],
            n_time_total=case["n_time_total"], file_times=case["file_times"],
            file_dt=case["file_dt"], ctx=gpu_ctx, resort_every=every).as_arrays()

    base, srt = run(0), run(1)
    assert (base["status"] == 2).any(), "the case should beach particles"
    for k in ("status", "failure_reason"):
        np.testing.assert_array_equal(srt[k], base[k], err_msg=k)
    for k in ("lon", "lat", "z", "vel_u", "vel_v", "vel_w"):
        a, b = np.asarray(base[k], np.float64), np.asarray(srt[k], np.float64)
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=k + " mask")
        np.testing.assert_array_equal(a.view(np.int64), b.view(np.int64), err_msg=k)


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_sigma_run_matches_reference_golden(gpu_ctx, direction):
    """FESOM-C sigma coordinates (tests/_gen_golden_3d.py:527): stretched
    sigma, sloping bottom, a tide and a dia-sigma flux, element u/v by least
    squares, W and the column height by least squares, RK2 with a linear
    time blend; particles leave through the east edge."""
    case = rc.sigma_case(direction)
    gold = np.load(os.path.join(GOLD, "fesomc_sigma_rk2" + ("_bwd" if direction < 0 else "")
                                + ".npz"), allow_pickle=False)
    rec = run_single_realization(
        case["mesh"], case["ds"], get_reader("fesomc_arrays"), case["init"], case["mode"],
        case["numerics"], start_time_idx=case["start_time_idx"], n_steps=case["n_steps"],
        n_time_total=case["n_time_total"], file_times=case["file_times"],
        file_dt=case["file_dt"], ctx=gpu_ctx).as_arrays()
    np.testing.assert_array_equal(rec["status"], gold["status"])
    np.testing.assert_array_equal(rec["failure_reason"], gold["failure_reason"])
    for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("u", "vel_u"), ("v", "vel_v"),
                   ("w", "vel_w")):
        np.testing.assert_allclose(rec[src], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)


@pytest.mark.parametrize("split,tinterp", [(6, True), (9, True), (5, False)])
def test_checkpoint_resume_is_bitwise(gpu_ctx, tmp_path, split, tinterp):
    """A run checkpointed after ``split`` steps and resumed from the file
    (tracking.py:2785 checkpoints, writers.py:1541 load_restart) records
    exactly what the uninterrupted run records: positions, status, failure
    codes, velocities -- through stochastic beaching and resuspension, whose
    draws and wet streaks continue on the global step."""
    from kelp_amd.tracking import load_restart
    case = rc.beaching_case(1.0)
    if not tinterp:
        # discrete: one record per step
        case["numerics"].temporal_interp = False
        case["numerics"].dt = float(case["file_dt"])
        case["n_steps"] = case["n_time_total"] - 1
    args = dict(n_time_total=case["n_time_total"], file_times=case["file_times"],
                file_dt=case["file_dt"], ctx=gpu_ctx)
    reader = get_reader("fesom2_arrays")
    full = run_single_realization(case["mesh"], case["ds"], reader, case["init"], case["mode"],
                                  case["numerics"], start_time_idx=case["start_time_idx"],
                                  n_steps=case["n_steps"], **args).as_arrays()
    ck = str(tmp_path / "restart.nc")
    first = run_single_realization(case["mesh"], case["ds"], reader, case["init"], case["mode"],
                                   case["numerics"], start_time_idx=case["start_time_idx"],
                                   n_steps=split, checkpoint_freq=split, checkpoint_path=ck,
                                   **args).as_arrays()
    r = load_restart(ck)
    rest = run_single_realization(case["mesh"], case["ds"], reader, r["state"], case["mode"],
                                  case["numerics"], start_time_idx=r["start_time_idx"],
                                  sim_time_start=r["sim_time_start"],
                                  start_step_i=r["start_step_i"],
                                  start_wet_streak=r["start_wet_streak"],
                                  particle_offset=r["particle_offset"],
                                  n_steps=case["n_steps"] - split, **args).as_arrays()
    assert r["start_step_i"] == split
    for k in ("lon", "lat", "z", "status", "failure_reason", "vel_u", "vel_v", "vel_w", "time"):
        both = np.concatenate([first[k], rest[k]], axis=-1)
        np.testing.assert_array_equal(both, full[k], err_msg=k)


def test_beaching_golden_through_a_three_slot_ring(gpu_ctx, monkeypatch):
    """The tracking loop with only three device slots: snapshots are evicted
    and the read-ahead refills them asynchronously behind the running steps
    (eta included); the reference's run-level golden still holds."""
    import kelp_amd.tracking as T
    monkeypatch.setattr(T, "N_SLOTS", 3)
    case = rc.beaching_case(1.0)
    gold = np.load(os.path.join(GOLD, "beaching_resuspension_rk2.npz"), allow_pickle=False)
    rec = run_single_realization(
        case["mesh"], case["ds"], get_reader("fesom2_arrays"), case["init"], case["mode"],
        case["numerics"], start_time_idx=case["start_time_idx"], n_steps=case["n_steps"],
        n_time_total=case["n_time_total"], file_times=case["file_times"],
        file_dt=case["file_dt"], ctx=gpu_ctx).as_arrays()
    np.testing.assert_array_equal(rec["status"], gold["status"])
    for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("u", "vel_u"), ("w", "vel_w")):
        np.testing.assert_allclose(rec[src], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)


def test_beaching_golden_with_elem_search_walk(gpu_ctx):
    """elem_search: walk (carried elements seeded by one global search,
    tracking.py:2236) only changes which element holds a point on a shared
    edge: the reference's local-search golden holds at its own bar."""
    case = rc.beaching_case(1.0)
    case["numerics"].elem_search = "walk"
    gold = np.load(os.path.join(GOLD, "beaching_resuspension_rk2.npz"), allow_pickle=False)
    rec = run_single_realization(
        case["mesh"], case["ds"], get_reader("fesom2_arrays"), case["init"], case["mode"],
        case["numerics"], start_time_idx=case["start_time_idx"], n_steps=case["n_steps"],
        n_time_total=case["n_time_total"], file_times=case["file_times"],
        file_dt=case["file_dt"], ctx=gpu_ctx).as_arrays()
    np.testing.assert_array_equal(rec["status"], gold["status"])
    for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("u", "vel_u"), ("w", "vel_w")):
        np.testing.assert_allclose(rec[src], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)


def _pchip_runs(gpu_ctx, direction, linear_data):
    """The sampling workload run twice, with temporal_reconstruction pchip
    and linear; with ``linear_data`` every variable is replaced by its
    record-0 value plus a per-record increment (linear in time)."""
    import copy
    import warnings
    from kelp_amd.readers import ArrayDataset
    case = rc.sampling_case(direction)
    if linear_data:
        ds = case["ds"]
        nt = ds.time.size
        t = np.arange(nt, dtype=np.float64)
        rng = np.random.default_rng(5)
        dv = {}
        for name, (dims, arr) in ds.vars.items():
            if dims and dims[0] == "time":
                step = 0.05 * np.abs(arr[0]).max() * rng.standard_normal(arr.shape[1:])
                arr = arr[0][None] + t.reshape((nt,) + (1,) * (arr.ndim - 1)) * step[None]
            dv[name] = (dims, np.ascontiguousarray(arr))
        case["ds"] = ArrayDataset(dv, time=ds.time, attrs=ds.attrs)
    out = {}
    for recon in ("pchip", "linear"):
        num = copy.deepcopy(case["numerics"])
        num.temporal_reconstruction = recon
        with warnings.catch_warnings(record=True) as caught:
            warnings.simplefilter("always")
            out[recon] = run_single_realization(
                case["mesh"], case["ds"], get_reader("fesom2_arrays"), case["init"], case["mode"],
                num, start_time_idx=case["start_time_idx"], n_steps=case["n_steps"],
                n_time_total=case["n_time_total"], file_times=case["file_times"],
                file_dt=case["file_dt"], sample_vars=case["sample_vars"], ctx=gpu_ctx).as_arrays()
        out[recon + "_warned"] = any("fewer than 4" in str(w.message) for w in caught)
    return out


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_pchip_run_reduces_to_linear_on_linear_data(gpu_ctx, direction):
    """On data linear in time the monotone cubic through four records is the
    straight line, so every stage field the loop reconstructs (its four
    snapshots, their times, the stage times q + direction c dt) must equal
    the linear blend's: the PCHIP run equals the linear run to rounding."""
    r = _pchip_runs(gpu_ctx, direction, linear_data=True)
    np.testing.assert_array_equal(r["pchip"]["status"], r["linear"]["status"])
    for k in ("lon", "lat", "z", "vel_u", "vel_v", "vel_w"):
        np.testing.assert_allclose(r["pchip"][k], r["linear"][k], rtol=0, atol=1e-10, err_msg=k)


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_pchip_run_reconstructs_interior_steps(gpu_ctx, direction):
    """On the sampling workload (u and w vary as sin t, cos t/2) the steps of
    the first record interval have no four-snapshot stencil and blend
    linearly (bit-identical to the linear run, one warning); the steps of
    the two interior intervals move differently under PCHIP."""
    r = _pchip_runs(gpu_ctx, direction, linear_data=False)
    assert r["pchip_warned"] and not r["linear_warned"]
    a, b = r["pchip"], r["linear"]
    for k in ("lon", "lat", "z", "vel_u", "vel_w"):
        np.testing.assert_array_equal(a[k][:, :2], b[k][:, :2], err_msg=k)
    assert np.nanmax(np.abs(a["lon"][:, 2:6] - b["lon"][:, 2:6])) > 1e-7
    assert np.nanmax(np.abs(a["vel_u"][:, 2:6] - b["vel_u"][:, 2:6])) > 1e-6


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_az_field_run_follows_the_payload_bracket(gpu_ctx, direction):
    """diffusion.var_az through the tracking loop: the reader's Az record
    rides in the snapshot slots and each step's kick takes the payload's
    bracket blend (integrators.py:873 vel.az).  With no flow, Az uniform in
    space and depth but different in every record, and particles far from
    the column ends, step g moves z by exactly sqrt(2 c_g dt) xi_g: c_g the
    blend of the records around the step's start, xi_g the counter-RNG
    normal of (seed, 0, pid, g, 1)."""
    import ctypes
    from kelp_amd.config import DiffusionConfig, ModeConfig, NumericsConfig
    from kelp_amd.particles import make_state
    from kelp_amd.readers import ArrayDataset
    from oracle.oracle import lib
    mesh = rc.run_mesh(lambda lon, lat: np.full(lon.shape, 90.0))
    K, N, T = mesh.sigma.size, mesh.n_nodes, 5
    cs = np.array([4e-4, 1.1e-3, 2.5e-4, 7e-4, 1.6e-3])
    z3 = lambda k: np.zeros((T, k, N))
    ds = ArrayDataset({"unod": (("time", "nz1", "nod2"), z3(K - 1)), "vnod": (("time", "nz1", "nod2"), z3(K - 1)),
                       "w": (("time", "nz", "nod2"), z3(K)), "ssh": (("time", "nod2"), np.zeros((T, N))),
                       "Av": (("time", "nz", "nod2"), np.ascontiguousarray(
                           np.broadcast_to(cs[:, None, None], (T, K, N))))},
                      time=np.arange(T) * rc.RUN_FILE_DT)
    i = np.arange(16, dtype=np.float64)
    lon, lat = 4.3 + 1.4 * ((i * 0.618034) % 1.0), 4.3 + 1.4 * ((i * 0.381966) % 1.0)
    z0 = 30.0 + 30.0 * ((i * 0.7548) % 1.0)
    seed = 1234
    numerics = NumericsConfig(dt=rc.RUN_DT, integrator="rk2", uv_interp="barycentric", w_interp="barycentric",
                              temporal_interp=True,
                              diffusion=DiffusionConfig(enabled=True, seed=seed, az_enabled=True, var_az="Av"))
    mode = ModeConfig(Fdrift2D=False, Fdrift3D=True, time_dir=direction, var_u="unod", var_v="vnod", var_w="w")
    n_steps, start = 8, (0 if direction > 0 else T - 1)
    rec = run_single_realization(mesh, ds, get_reader("fesom2_arrays"), make_state(lon, lat, z0), mode, numerics,
                                 start_time_idx=start, n_steps=n_steps, n_time_total=T,
                                 file_times=np.arange(T) * rc.RUN_FILE_DT, file_dt=rc.RUN_FILE_DT,
                                 ctx=gpu_ctx).as_arrays()
    L = lib("glibc")
    a, b = ctypes.c_double(), ctypes.c_double()
    z = z0.copy()
    for g in range(n_steps):
        t = start * rc.RUN_FILE_DT + direction * g * rc.RUN_DT
        ia = int(np.floor(t / rc.RUN_FILE_DT + 1e-9))
        if direction < 0 and abs(t - ia * rc.RUN_FILE_DT) < 1e-6:
            ia -= 1          # backward on a record: the interval below it
        al = (t - ia * rc.RUN_FILE_DT) / rc.RUN_FILE_DT
        c = cs[ia] if abs(al) < 1e-12 else (1.0 - al) * cs[ia] + al * cs[ia + 1]
        for j in range(z.size):
            L.or_normal_pair(seed, 0, j, g, 1, ctypes.byref(a), ctypes.byref(b))
            z[j] = z[j] + np.sqrt(2.0 * c * rc.RUN_DT) * a.value
        np.testing.assert_allclose(rec["z"][:, g], z, rtol=0, atol=1e-9, err_msg=f"step {g}")
    assert (rec["status"] == 1).all()


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_healpix_run_matches_reference_golden(gpu_ctx, direction):
    """The native HEALPix reader's store (finite-prefix columns of 1-3
    layers, NaN land and padding, W on the interfaces, float32) on its wet
    dual mesh, node RK2 with a time blend: the reference's run-level golden
    ``healpix_3d_rk2`` through the device tracking loop."""
    case = rc.healpix_case(direction)
    gold = np.load(os.path.join(GOLD, "healpix_3d_rk2" + ("_bwd" if direction < 0 else "") + ".npz"),
                   allow_pickle=False)
    rec = run_single_realization(
        case["mesh"], case["ds"], get_reader("healpix_arrays"), case["init"], case["mode"],
        case["numerics"], start_time_idx=case["start_time_idx"], n_steps=case["n_steps"],
        n_time_total=case["n_time_total"], file_times=case["file_times"],
        file_dt=case["file_dt"], ctx=gpu_ctx).as_arrays()
    assert (gold["status"] == 3).any() and (gold["status"] == 1).any()
    np.testing.assert_array_equal(rec["status"], gold["status"])
    np.testing.assert_array_equal(rec["failure_reason"], gold["failure_reason"])
    for k, src in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("u", "vel_u"), ("v", "vel_v"),
                   ("w", "vel_w")):
        np.testing.assert_allclose(rec[src], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)


def test_run_tracking_matches_reference_multiyear_golden(gpu_ctx, tmp_path):
    """run_tracking(config) end to end (tests/_gen_golden_3d.py:740): two
    yearly files of daily means read as one axis and shifted to the window
    midpoints, RK2 across the year boundary, temp sampled at every record,
    failure reasons tracked, read-ahead of two; every trajectory variable of
    the reference's output file, its time axis and release metadata."""
    from kelp_amd.tracking import run_tracking
    from kelp_amd.writers import read_trajectories
    case = rc.multiyear_case()
    gold = np.load(os.path.join(GOLD, "run_tracking_fesom2_multiyear_node.npz"), allow_pickle=False)
    case["config"].data.output_path = str(tmp_path / "traj.nc")
    (rec,) = run_tracking(case["config"], mesh=case["mesh"], dataset=case["ds"], reader="fesom2_arrays",
                          init_state=case["init"], ctx=gpu_ctx)
    # the output file: every variable the reference's file holds
    f = read_trajectories(case["config"].data.output_path)
    assert set(f) == {k for k in gold.files if not k.startswith(("attr__", "dims__"))}
    for k in f:
        if f[k].dtype.kind == "f":
            np.testing.assert_allclose(f[k], gold[k], rtol=1e-9, atol=1e-9, err_msg=k)
        else:
            np.testing.assert_array_equal(f[k], gold[k], err_msg=k)
    out = rec.as_arrays()
    np.testing.assert_array_equal(out["status"], gold["status"])
    np.testing.assert_array_equal(out["failure_reason"], gold["failure_reason"])
    np.testing.assert_array_equal(out["time"], gold["time"])
    for k in ("release_lon", "release_lat", "release_z", "release_time"):
        np.testing.assert_array_equal(out[k], gold[k], err_msg=k)
    assert rec.meta["time_averaging"] == str(gold["attr__time_averaging"])
    for k, g in (("lon", "lon"), ("lat", "lat"), ("z", "z"), ("vel_u", "vel_u"), ("vel_v", "vel_v"),
                 ("vel_w", "vel_w"), ("sample__temp", "temp")):
        np.testing.assert_allclose(out[k], gold[g], rtol=1e-9, atol=1e-9, err_msg=k)


def test_run_tracking_output_keeps_the_last_checkpoint(gpu_ctx, tmp_path):
    """With numerics.checkpoint_freq > 0 the output file also holds the
    checkpoint (tracking.py:2785-2862 through the writer): after the run it
    is the last step's state, in the reference's restart variables."""
    import copy
    from kelp_amd.tracking import load_restart, run_tracking
    case = rc.multiyear_case()
    cfg = copy.deepcopy(case["config"])
    cfg.data.output_path = str(tmp_path / "traj.nc")
    cfg.numerics.checkpoint_freq = 3
    (rec,) = run_tracking(cfg, mesh=case["mesh"], dataset=case["ds"], reader="fesom2_arrays",
                          init_state=case["init"], ctx=gpu_ctx)
    out = rec.as_arrays()
    r = load_restart(cfg.data.output_path)
    n_rec = out["time"].size
    assert r["step_i"] == n_rec - 1 and r["start_step_i"] == n_rec
    np.testing.assert_array_equal(r["state"].status, out["status"][:, -1])
    alive = out["status"][:, -1] == 1
    for k in ("lon", "lat", "z"):
        np.testing.assert_array_equal(getattr(r["state"], k)[alive], out[k][alive, -1], err_msg=k)
    np.testing.assert_array_equal(r["state"].release_lon, out["release_lon"])


def test_rk45_run_agrees_with_fine_rk4_run(gpu_ctx):
    """integrator rk45 through the tracking loop (2-D, element layout, a
    flow rotating faster from record to record): at a 1 cm tolerance its
    records agree with an RK4 run of 15-minute steps to well under a metre,
    so the loop hands rk45 the right bracket, blend fractions (t over the
    time to the next record) and step lengths."""
    from kelp_amd import synth
    from kelp_amd.config import ModeConfig, NumericsConfig
    from kelp_amd.particles import make_state
    from kelp_amd.readers import ArrayDataset
    mesh = synth.lattice_mesh(0.0, 8.0, 40.0, 46.0, 0.25, jitter=0.05, seed=3)
    x = (mesh.elem_center_lon - 4.0) * np.cos(np.radians(43.0))
    y = mesh.elem_center_lat - 43.0
    T = 4
    u = np.stack([-(0.5 + 0.3 * s) * y for s in range(T)])
    v = np.stack([(0.5 + 0.3 * s) * x for s in range(T)])
    ds = ArrayDataset({"u": (("time", "elem"), u), "v": (("time", "elem"), v)}, time=np.arange(T) * 21600.0)
    rng = np.random.default_rng(2)
    lon, lat = rng.uniform(2.5, 5.5, 200), rng.uniform(42.0, 44.0, 200)
    mode = ModeConfig(Fdrift2D=True, Fdrift3D=False, var_u="u", var_v="v")
    runs = {}
    for integ, dt in (("rk45", 10800.0), ("rk4", 900.0)):
        num = NumericsConfig(dt=dt, integrator=integ, uv_interp="leastsq", temporal_interp=True,
                             rk45_tol_m=0.01, output_freq=int(10800.0 / dt))
        runs[integ] = run_single_realization(
            mesh, ds, get_reader("fesom2_arrays"), make_state(lon, lat, np.zeros(lon.size)), mode, num,
            start_time_idx=0, n_steps=int(3 * 21600.0 / dt), n_time_total=T,
            file_times=np.arange(T) * 21600.0, file_dt=21600.0, ctx=gpu_ctx).as_arrays()
    a, b = runs["rk45"], runs["rk4"]
    np.testing.assert_array_equal(a["time"], b["time"])
    ok = (a["status"] == 1) & (b["status"] == 1)
    assert ok.mean() > 0.95
    d = np.hypot((a["lon"] - b["lon"]) * 111120.0 * np.cos(np.radians(a["lat"])), (a["lat"] - b["lat"]) * 111120.0)
    assert np.nanmax(np.where(ok, d, 0.0)) < 0.5
    moved = np.hypot(a["lon"][:, -1] - lon, a["lat"][:, -1] - lat) * 111120.0
    assert np.median(moved) > 10000.0
