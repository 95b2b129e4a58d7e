"""The request loop (kelp_amd.serving): continuous batching of tracking
requests against resident fields (GPU).

A stream of requests of different sizes, durations, directions, kinds and
random walks is served from a pool too small to hold them all at once, so
requests wait, are admitted as others finish and their slots are compacted.
Bar: each request's final state equals the same request run alone through
the plain step, bit for bit, and every request completes.
"""
import numpy as np
import pytest

from kelp_amd import synth
from kelp_amd.batching import TuneChoice
from kelp_amd.serving import Pool, Request, Server, latency_summary

from .test_gpu_batch import _alone, _assert_same, _setup_zlevel

pytestmark = pytest.mark.gpu


def test_served_requests_equal_their_runs_alone(gpu_ctx):
    mesh = synth.sphere_mesh(20000)
    snaps = [synth.jet_rossby_snapshot(mesh, s * synth.SNAP_DT) for s in range(3)]
    _setup_zlevel(gpu_ctx, mesh, snaps)
    rng = np.random.default_rng(7)
    lon, lat, z = synth.wet_area_seeds(mesh, 20000, seed=99)
    reqs, off = [], 0
    for r in range(14):
        n = int(rng.integers(50, 2500))
        reqs.append(Request(rid=r, lon=lon[off:off + n], lat=lat[off:off + n], z=z[off:off + n],
                            n_steps=int(rng.integers(3, 15)), direction=float(rng.choice([1.0, -1.0])),
                            is3d=bool(rng.random() < 0.5), kh=float(rng.choice([0.0, 10.0])),
                            seed=int(rng.integers(1, 2 ** 40)), t0=float(rng.uniform(0, 40000.0)),
                            arrival=0.0))
        off += n
    pool = Pool(gpu_ctx, axis3=(0, 3), axis2=(3, 1), snap_dt=synth.SNAP_DT, capacity=5000, max_clouds=6,
                choice=TuneChoice(128, 4, 2), order=2, dt=1800.0, exact_nearest=1, w_method=2)
    srv = Server({"default": pool}, chunk=3)
    res = srv.run(reqs, realtime=False)
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    assert srv.alive_steps > 0
    s = latency_summary(res)
    assert s["p99_s"] >= s["p50_s"] > 0
    by_id = {r.rid: r for r in res}
    for q in reqs:
        from kelp_amd.batching import CloudSpec
        spec = CloudSpec(lon=q.lon, lat=q.lat, z=q.z, n_steps=q.n_steps, direction=q.direction, is3d=q.is3d,
                         kh=q.kh, seed=q.seed, t0=q.t0)
        want = _alone(gpu_ctx, spec, (0, 3), (3, 1), synth.SNAP_DT, 2, 1800.0, None)
        _assert_same(by_id[q.rid].state, want, q.rid)
