"""Asynchronous snapshot refill (the snapshot ring) on the device (GPU).

kelp_field_upload_async refills a slot on the copy stream while steps run.
Bar: a run whose slots are refilled asynchronously -- right after a step
that still reads the old snapshot was enqueued -- is bit-identical to the
same run with synchronous uploads between the steps.
"""
import numpy as np
import pytest

from kelp_amd import synth
from kelp_amd._native import KelpError

pytestmark = pytest.mark.gpu

DT = 1800.0


@pytest.fixture(scope="module")
def world():
    mesh = synth.sphere_mesh(8000)
    snaps = [synth.jet_rossby_snapshot(mesh, s * synth.SNAP_DT) for s in range(4)]
    lon, lat, z = synth.wet_area_seeds(mesh, 4000)
    return mesh, snaps, lon, lat, z


def _run(ctx, world, refill):
    """Steps over four snapshots through two slots: after each step the
    slot the next bracket no longer needs receives the next snapshot."""
    mesh, snaps, lon, lat, z = world
    f32 = [tuple(a.astype(np.float32) for a in sn[:3]) for sn in snaps]
    ctx.upload_mesh(mesh)
    ctx.upload_field(0, *f32[0])
    ctx.upload_field(1, *f32[1])
    pinned = [tuple(ctx.pinned(a) for a in f) for f in f32]
    p = ctx.particles(lon, lat, z, pid=np.arange(lon.size))
    slot_of = {0: 0, 1: 1}
    steps_per_snap = int(synth.SNAP_DT // DT)
    for k in range(2 * steps_per_snap):
        t = k * DT
        e = int(t // synth.SNAP_DT)
        ia, ib, a0, it, dtn = synth.snapshot_bracket(t, synth.SNAP_DT, 4)
        sa, st = slot_of[e], slot_of[e + 1]
        sb = st if ib >= 0 else -1
        al = [c * DT / dtn for c in (0.0, 0.5)] + [0.0, 0.0]
        ctx.step(p, 2, sa, DT, 1.0, slot_b=sb, slot_t=st, alpha0=a0, alphas=al, exact_nearest=1,
                 w_method=2, step_i=k)
        if (k + 1) % steps_per_snap == 0:
            # the next bracket is (e + 1, e + 2): refill epoch e's slot
            s = slot_of.pop(e)
            if refill == "async":
                ctx.upload_field_async(s, *pinned[e + 2])
            else:
                ctx.upload_field(s, *f32[e + 2])
            slot_of[e + 2] = s
    return p.to_host()


def test_async_refill_equals_sync_upload(gpu_ctx, world):
    a = _run(gpu_ctx, world, "async")
    b = _run(gpu_ctx, world, "sync")
    assert (a["status"] == 1).mean() > 0.9
    for k in ("lon", "lat", "z", "status", "nn", "fr", "u", "v", "w"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_async_refill_needs_a_filled_slot(gpu_ctx, world):
    mesh, snaps, *_ = world
    gpu_ctx.upload_mesh(mesh)
    U, V, W = (x.astype(np.float32) for x in snaps[0][:3])
    with pytest.raises(KelpError, match="must already hold"):
        gpu_ctx.upload_field_async(5, U, V, W)
    gpu_ctx.upload_field(5, U, V, W)
    with pytest.raises(KelpError, match="must already hold"):
        gpu_ctx.upload_field_async(5, U.astype(np.float64), V.astype(np.float64), W.astype(np.float64))
