"""Batched clouds (kelp_batch_*) against each cloud run alone (GPU).

Every cloud of a batch -- its own direction, start step, duration, time on
the snapshot axis, 2-D or 3-D kind, random walk and seed -- must come out
bit-identical to the same cloud advanced alone by the plain step
(``kelp_step`` with the bracket of the cloud's time, its seed and its
cloud-local step and particle ids): positions, depth, status and failure
code.  Also: the result does not depend on the launch shape (workgroup
size, re-sort interval, steps per graph launch), a cloud admitted while
others run equals its run alone, and a sigma-mesh batch (the mixed
triangle/quad lattice) matches the sigma step.
"""
import numpy as np
import pytest

from kelp_amd import synth
from kelp_amd.batching import Batch, CloudSpec, TuneChoice, autotune
from kelp_amd.device import DeviceContext

pytestmark = pytest.mark.gpu

C = {1: (0.0,), 2: (0.0, 0.5), 4: (0.0, 0.5, 0.5, 1.0)}
BEACH = dict(criteria=["on_failure", "stochastic"], t_beach=86400.0, depth_min=0.3)


def _alone(ctx, spec, axis3, axis2, snap_dt, order, dt, beach, sigma=None):
    """The cloud run alone through kelp_step, bracket by bracket."""
    p = ctx.particles(spec.lon, spec.lat, spec.z if spec.is3d else None, pid=np.arange(spec.n))
    slot0, n = axis3 if spec.is3d else axis2
    for k in range(spec.n_steps):
        if n <= 1:
            sa, sb, st, a0, al = slot0, -1, -1, 0.0, [0.0] * 4
        else:
            t = spec.t0 + spec.direction * (k * dt)
            ia, ib, a0, it, dtn = synth.snapshot_bracket(t, snap_dt, n, spec.direction)
            sa, sb = slot0 + ia, (slot0 + ib if ib >= 0 else -1)
            if dt > dtn:
                st, al = -1, [0.0] * 4
            else:
                st, al = slot0 + it, [c * dt / dtn for c in C[order]] + [0.0] * (4 - order)
        kick = dict(kh_x=spec.kh, kh_y=spec.kh, seed=spec.seed) if spec.kh > 0 else None
        b = None if beach is None else dict(beach, seed=spec.seed)
        ctx.step(p, order, sa, dt, spec.direction, slot_b=sb, slot_t=st, alpha0=a0, alphas=al,
                 exact_nearest=1, w_method=2, kick=kick, beach=b, step_i=k, real_id=spec.real_id,
                 reuse_anchor=True, sigma=sigma if spec.is3d else None)
    return p.to_host()


def _assert_same(got, want, tag):
    for k in ("status", "fr"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=f"{k} {tag}")
    for k in ("lon", "lat", "z"):
        np.testing.assert_array_equal(got[k].view(np.int64), want[k].view(np.int64), err_msg=f"{k} {tag}")


@pytest.fixture(scope="module")
def sphere_world():
    mesh = synth.sphere_mesh(20000)
    snaps = [synth.jet_rossby_snapshot(mesh, s * synth.SNAP_DT) for s in range(3)]
    lon, lat, z = synth.wet_area_seeds(mesh, 6000)
    return mesh, snaps, lon, lat, z


def _specs(lon, lat, z):
    # (offset, size, n_steps, direction, is3d, kh, seed, t0, start)
    rows = [(0, 1500, 14, 1.0, True, 10.0, 11, 0.0, 0), (1500, 900, 9, -1.0, True, 0.0, 12, 50000.0, 2),
            (2400, 700, 12, 1.0, False, 25.0, 13, 3600.0, 1), (3100, 1200, 7, -1.0, False, 0.0, 14, 0.0, 5),
            (4300, 500, 16, 1.0, True, 40.0, 15, 20000.0, 0), (4800, 1200, 10, -1.0, True, 10.0, 16, 80000.0, 3)]
    return [CloudSpec(lon=lon[o:o + n], lat=lat[o:o + n], z=z[o:o + n], n_steps=ns, direction=d, is3d=d3,
                      kh=kh, seed=sd, t0=t0, start_step=st) for o, n, ns, d, d3, kh, sd, t0, st in rows]


def _setup_zlevel(ctx, mesh, snaps):
    ctx.upload_mesh(mesh)
    for s, sn in enumerate(snaps):
        ctx.upload_field(s, sn[0], sn[1], sn[2])
    ctx.upload_field(3, snaps[0][0][0].copy(), snaps[0][1][0].copy())     # steady 2-D: the top layer


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("choice", [TuneChoice(256, 0, 1), TuneChoice(64, 3, 4)])
def test_batch_equals_each_cloud_alone(gpu_ctx, sphere_world, order, choice):
    mesh, snaps, lon, lat, z = sphere_world
    _setup_zlevel(gpu_ctx, mesh, snaps)
    dt = 1800.0
    specs = _specs(lon, lat, z)
    b = Batch(gpu_ctx, 8000, axis3=(0, 3), axis2=(3, 1), snap_dt=synth.SNAP_DT, max_clouds=16,
              order=order, dt=dt, exact_nearest=1, w_method=2, beach=dict(BEACH, seed=0))
    ids = [b.add(s) for s in specs]
    total = max(s.start_step + s.n_steps for s in specs)
    b.steps(total, choice)
    assert sorted(b.finished()) == sorted(ids)
    for cid, s in zip(ids, specs):
        got = b.collect(cid)
        want = _alone(gpu_ctx, s, (0, 3), (3, 1), synth.SNAP_DT, order, dt, BEACH)
        assert (want["status"] == 1).mean() > 0.5
        _assert_same(got, want, (cid, order, choice.block, choice.resort_every, choice.graph_steps))


def test_admitted_mid_run_and_freed_slots_reused(gpu_ctx, sphere_world):
    """Continuous batching: clouds admitted at later batch steps, into slots
    freed by finished clouds after a re-sort, equal their runs alone."""
    mesh, snaps, lon, lat, z = sphere_world
    _setup_zlevel(gpu_ctx, mesh, snaps)
    dt = 1800.0
    specs = _specs(lon, lat, z)
    b = Batch(gpu_ctx, 4000, axis3=(0, 3), axis2=(3, 1), snap_dt=synth.SNAP_DT, max_clouds=4,
              order=2, dt=dt, exact_nearest=1, w_method=2)
    ch = TuneChoice(128, 4, 2)
    pending = list(specs)
    results = {}
    live = {}
    while pending or live:
        # admit what fits now; a cloud starts at the current batch step
        while pending and pending[0].n <= b.room() and len(live) < 4:
            s = pending.pop(0)
            s = CloudSpec(**{**s.__dict__, "start_step": b.step_index})
            live[b.add(s)] = s
        b.steps(1, ch)
        for cid in b.finished():
            s = live.pop(cid)
            results[id(s)] = (s, b.collect(cid))
        if b.room() < (pending[0].n if pending else 0):
            b.sort()
    assert len(results) == len(specs)
    for s, got in results.values():
        want = _alone(gpu_ctx, s, (0, 3), (3, 1), synth.SNAP_DT, 2, dt, None)
        _assert_same(got, want, s.seed)


def test_autotune_keeps_results(gpu_ctx, sphere_world):
    mesh, snaps, lon, lat, z = sphere_world
    _setup_zlevel(gpu_ctx, mesh, snaps)
    specs = _specs(lon, lat, z)
    b = Batch(gpu_ctx, 8000, axis3=(0, 3), axis2=(3, 1), snap_dt=synth.SNAP_DT, max_clouds=8,
              order=2, dt=1800.0, exact_nearest=1, w_method=2)
    ids = [b.add(s) for s in specs]
    best = autotune(b, steps=4, blocks=(64, 256), resort=(0, 4), graph=(1, 4), warmup=1)
    assert len(best.table) == 8 and best.rate > 0
    b.steps(20)
    ref = Batch(gpu_ctx, 8000, axis3=(0, 3), axis2=(3, 1), snap_dt=synth.SNAP_DT, max_clouds=8,
                order=2, dt=1800.0, exact_nearest=1, w_method=2)
    ids2 = [ref.add(s) for s in specs]
    ref.steps(20, TuneChoice(256, 0, 1))
    for c1, c2 in zip(ids, ids2):
        _assert_same(b.collect(c1), ref.collect(c2), c1)


def test_sigma_mixed_lattice_batch(gpu_ctx):
    """The batching workload's mesh (mixed quads and triangles, sigma
    layers, steady fields) at reduced cloud sizes: 3-D clouds equal the
    sigma step alone, 2-D clouds the 2-D step alone."""
    mesh = synth.mixed_lattice_mesh(lon1=20.0, lat1=35.0)
    u, v, w, eta, u2, v2 = synth.mixed_lattice_fields(mesh, lon1=20.0, lat1=35.0)
    gpu_ctx.upload_mesh(mesh)
    gpu_ctx.upload_field(0, u, v, w)
    gpu_ctx.upload_eta(0, eta)
    gpu_ctx.upload_field(1, u2, v2)
    clouds = synth.batch_clouds(mesh, n_clouds=12, max_size=3000, min_days=5, max_days=20)
    specs = [CloudSpec(**c) for c in clouds]
    sigma = dict(start=(0, -1, -1, 0.0, 0.0), end=None, Dmin=0.5, z0b_min=0.1)
    dt = synth.DAY / 8
    b = Batch(gpu_ctx, sum(s.n for s in specs), axis3=(0, 1), axis2=(1, 1), max_clouds=16, order=2, dt=dt,
              exact_nearest=1, w_method=2, sigma=sigma)
    ids = [b.add(s) for s in specs]
    b.steps(max(s.n_steps for s in specs), TuneChoice(128, 4, 4))
    for cid, s in zip(ids, specs):
        got = b.collect(cid)
        if s.is3d:
            want = _alone(gpu_ctx, s, (0, 1), (1, 1), 0.0, 2, dt, None, sigma=sigma)
        else:
            want = _alone2d_zlevel(mesh, s, u2, v2, dt)
        assert (want["status"] == 1).mean() > 0.3, cid
        _assert_same(got, want, (cid, s.is3d))


def _alone2d_zlevel(mesh, spec, u2, v2, dt):
    """A 2-D cloud on the horizontal mesh alone (the sigma step takes 3-D
    fields only): a second context holding the mesh as a z-level mesh."""
    import copy
    m2 = copy.copy(mesh)
    m2.sigma = np.array([], dtype=np.float64)
    m2.vertical_coord = "zlevel"
    ctx = DeviceContext(0)
    try:
        ctx.upload_mesh(m2)
        ctx.upload_field(1, u2, v2)
        return _alone(ctx, spec, (0, 1), (1, 1), 0.0, 2, dt, None)
    finally:
        ctx.close()
