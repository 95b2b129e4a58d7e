"""The device path at the full sizes of the two headline configurations,
against the oracle built with glibc's libm (the reference's own libm).

* The 2-D double gyre at full size: the jittered 0.25-degree lattice
  (19,481 nodes), 10,000 particles, RK2, dt 600 s, 1000 steps, compared at
  every output (every 10 steps).
* A 2,000-particle subsample of the 3-D global workload at its full mesh
  size (769,830 wet nodes of the 1,000,000-node Fibonacci sphere, 47
  layers, 8 six-hourly float32 snapshots): RK4, dt 1800 s, 336 steps, time
  blend, Kh = 10 m^2/s random walk, beaching (on_failure + stochastic), T
  and S sampled by least squares at every output (every 12 steps).  The
  device runs exactly as the benchmark drives it: anchors reused from the
  previous step, the state re-sorted by hash cell every 8 steps.

Bar: integer outputs (status, nearest node, failure code, wet streak)
identical; positions within 1e-9 degrees and 1e-6 m; sampled scalars and
velocities within 1e-9 relative (the reference's cross-environment rule,
tests/test_golden_3d_runs.py:11-18).
"""
import numpy as np
import pytest

from kelp_amd import synth
from oracle.oracle import OracleMesh, step_args

pytestmark = pytest.mark.gpu


def _compare(h, st, keys_f, tol_deg=1e-9, tol_m=1e-6):
    for k in ("status", "nn", "fr"):
        if k in st:
            np.testing.assert_array_equal(h[k], st[k], err_msg=k)
    if "ws" in st:
        np.testing.assert_array_equal(h["ws"], st["ws"], err_msg="ws")
    for k in ("lon", "lat"):
        d = np.abs(h[k] - st[k])
        # a longitude difference across the seam is a multiple of 360
        if k == "lon":
            d = np.minimum(d, np.abs(d - 360.0))
        assert np.max(d) < tol_deg, (k, float(np.max(d)))
    if "z" in keys_f:
        assert np.max(np.abs(h["z"] - st["z"])) < tol_m, ("z", float(np.max(np.abs(h["z"] - st["z"]))))
    for k in keys_f:
        if k in ("lon", "lat", "z"):
            continue
        np.testing.assert_allclose(h[k], st[k], rtol=1e-9, atol=1e-12, err_msg=k)


def test_double_gyre_full_size_vs_glibc(gpu_ctx):
    """10,000 particles x 1000 RK2 steps on the full lattice."""
    w = synth.double_gyre_workload(n_particles=10_000)
    assert w.mesh.n_nodes == 161 * 121 and w.n_steps == 1000
    gpu_ctx.upload_mesh(w.mesh)
    gpu_ctx.upload_field(0, w.u, w.v)
    p = gpu_ctx.particles(w.lon, w.lat)
    om = OracleMesh(w.mesh, "glibc")
    st = dict(lon=w.lon.copy(), lat=w.lat.copy(), status=np.ones(w.lon.size, np.int32),
              nn=np.full(w.lon.size, -1, np.int32))
    checked = 0
    for k in range(w.n_steps):
        gpu_ctx.step(p, 2, 0, w.dt, 1.0, exact_nearest=1, reuse_anchor=True)
        st = om.step2d(st, 2, w.u, w.v, w.dt, 1.0)
        if (k + 1) % w.output_every == 0:
            h = p.to_host()
            _compare(h, st, ("lon", "lat", "u", "v"))
            checked += 1
    assert checked == 100
    # particles left the basin through the open edges over 1000 steps
    assert (st["status"] == 1).sum() > 0.5 * w.lon.size


@pytest.fixture(scope="module")
def global3d():
    w = synth.global_3d_workload(n_nodes=1_000_000, n_particles=1_000_000)
    return w


def test_global_3d_full_mesh_subsample_vs_glibc(gpu_ctx, global3d):
    """2,000 of the benchmark's 1,000,000 seeds through all 336 RK4 steps
    on the full 769,830-node mesh."""
    w = global3d
    assert w.mesh.n_nodes == 769_830 and w.n_steps == 336
    sel = np.arange(0, w.lon.size, w.lon.size // 2000)[:2000]
    lon, lat, z = w.lon[sel], w.lat[sel], w.z[sel]
    pid = sel.astype(np.int64)
    gpu_ctx.upload_mesh(w.mesh)
    for s, (U, V, W, T, S) in enumerate(w.snaps):
        gpu_ctx.upload_field(s, U, V, W)
        gpu_ctx.upload_scalar(s, 0, T)
        gpu_ctx.upload_scalar(s, 1, S)
    kick = dict(kh_x=w.kh, kh_y=w.kh, seed=20240501)
    beach = dict(criteria=["on_failure", "stochastic"], t_beach=86400.0, depth_min=0.3,
                 seed=20240501)
    p = gpu_ctx.particles(lon, lat, z, pid=pid)
    om = OracleMesh(w.mesh, "glibc")
    st = dict(lon=lon.copy(), lat=lat.copy(), z=z.copy(), status=np.ones(lon.size, np.int32),
              nn=np.full(lon.size, -1, np.int32), pid=pid.copy())
    n_snap = len(w.snaps)
    checked = 0
    for k in range(w.n_steps):
        t = k * w.dt
        ia, ib, a0, it, dtn = synth.snapshot_bracket(t, w.snap_dt, n_snap)
        al = [c * w.dt / dtn for c in (0.0, 0.5, 0.5, 1.0)]
        A, T_ = w.snaps[ia], w.snaps[it]
        f = dict(u_a=A[0], v_a=A[1], w_a=A[2], u_t=T_[0], v_t=T_[1], w_t=T_[2], alpha0=a0)
        if ib >= 0:
            B = w.snaps[ib]
            f.update(u_b=B[0], v_b=B[1], w_b=B[2])
        st = om.step(st, f, step_args(4, w.dt, 1.0, is3d=True, w_method=2, exact=1, alphas=al,
                                      kick=kick, beach=beach, step_i=k))
        if k % 8 == 0:
            gpu_ctx.sort(p)
        gpu_ctx.step(p, 4, ia, w.dt, 1.0, slot_b=ib, slot_t=it, alpha0=a0, alphas=al,
                     exact_nearest=1, w_method=2, kick=kick, beach=beach, step_i=k,
                     reuse_anchor=True)
        if (k + 1) % w.output_every == 0:
            ra, rb, ra0, _, _ = synth.snapshot_bracket(t + w.dt, w.snap_dt, n_snap)
            rows = gpu_ctx.sample_vars(p, (0, 1), ra, rb, ra0, "leastsq")
            gpu_ctx.synchronize()
            rows = rows.cpu().numpy()
            h = p.to_host()
            order = np.argsort(h["pid"], kind="stable")
            h = {kk: v[order] for kk, v in h.items()}
            for j, var in enumerate((3, 4)):
                h[f"s{var}"] = rows[j][order]
                st[f"s{var}"] = om.sample(w.snaps[ra][var], st["lon"], st["lat"], st["z"],
                                          st["status"], st["nn"],
                                          values_b=None if rb < 0 else w.snaps[rb][var], alpha=ra0)
            _compare(h, st, ("lon", "lat", "z", "u", "v", "w", "s3", "s4"))
            checked += 1
    assert checked == 28
    assert (st["status"] == 1).sum() > 1800


def test_healpix_precision_mesh_reduced_nside_vs_glibc(gpu_ctx):
    """The precision workload at a reduced nside: the NEST HEALPix dual mesh
    at nside 32 over the harmonic wet mask (quads and the valence-3
    triangles), 69 z-levels, four six-hourly float32 snapshots of the 3-D
    generators cycled, RK2 with dt 6 h and the Kh = 10 m^2/s random walk, 120
    steps (30 days) of 3,000 particles, re-sorted every 8 steps; compared with
    the glibc oracle every 20 steps."""
    mesh = synth.healpix_c3_mesh(32)
    assert mesh.sigma.size == 69 and mesh.max_v == 4
    snaps = [synth.jet_rossby_snapshot(mesh, k * synth.SNAP_DT) for k in range(4)]
    lon, lat, z = synth.wet_area_seeds(mesh, 3000, seed=77)
    pid = np.arange(lon.size, dtype=np.int64)
    gpu_ctx.upload_mesh(mesh)
    for s, (U, V, W, _, _) in enumerate(snaps):
        gpu_ctx.upload_field(s, U, V, W)
    kick = dict(kh_x=10.0, kh_y=10.0, seed=20241231)
    p = gpu_ctx.particles(lon, lat, z, pid=pid)
    om = OracleMesh(mesh, "glibc")
    st = dict(lon=lon.copy(), lat=lat.copy(), z=z.copy(), status=np.ones(lon.size, np.int32),
              nn=np.full(lon.size, -1, np.int32), pid=pid.copy())
    dt = 21600.0
    for k in range(120):
        ia, ib, a0, it, dtn = synth.snapshot_bracket(k * dt, synth.SNAP_DT, 4)
        al = [c * dt / dtn for c in (0.0, 0.5)] + [0.0, 0.0]
        A, T_ = snaps[ia], snaps[it]
        f = dict(u_a=A[0], v_a=A[1], w_a=A[2], u_t=T_[0], v_t=T_[1], w_t=T_[2], alpha0=a0)
        if ib >= 0:
            B = snaps[ib]
            f.update(u_b=B[0], v_b=B[1], w_b=B[2])
        st = om.step(st, f, step_args(2, dt, 1.0, is3d=True, w_method=2, exact=1, alphas=al, kick=kick,
                                      step_i=k))
        if k % 8 == 0:
            gpu_ctx.sort(p)
        gpu_ctx.step(p, 2, ia, dt, 1.0, slot_b=ib, slot_t=it, alpha0=a0, alphas=al, exact_nearest=1,
                     w_method=2, kick=kick, step_i=k, reuse_anchor=True)
        if (k + 1) % 20 == 0:
            h = p.to_host()
            order = np.argsort(h["pid"], kind="stable")
            h = {kk: v[order] for kk, v in h.items()}
            _compare(h, st, ("lon", "lat", "z", "u", "v", "w"))
    assert (st["status"] == 1).mean() > 0.8
