"""Edge-walk element tracking (elem_search: walk) on the device (GPU).

Every site's containing element comes from the edge walk started at the
previous site's, the carried element at the first (core/mesh.py:2489
_walk_containing_element, :2582 _loop_walk_containing; integrators.py:5020,
:5415, :5792), and the end-of-step domain verdict from the walk too.  Bar:
bit-identical to the oracle built with the device's math, the carried
elements included, over several steps with kicks and a periodic fold; and
the same trajectories as the local search within 1e-10 (the two can only
pick different elements for a point on a shared edge).
"""
import numpy as np
import pytest

from kelp_amd import synth
from kelp_amd.mesh import build_elem_neighbors
from oracle.oracle import OracleMesh, step_args

pytestmark = pytest.mark.gpu

DT = 1800.0


@pytest.fixture(scope="module")
def world3d():
    mesh = synth.sphere_mesh(12000)
    build_elem_neighbors(mesh)
    snaps = [synth.jet_rossby_snapshot(mesh, s * synth.SNAP_DT) for s in range(3)]
    lon, lat, z = synth.wet_area_seeds(mesh, 5000)
    return mesh, snaps, lon, lat, z


@pytest.fixture(scope="module")
def world2d():
    w = synth.double_gyre_workload(n_particles=5000)
    build_elem_neighbors(w.mesh)
    return w


def _seed_elems(om, lon, lat):
    nn = om.nearest(lon, lat, exact=True)
    return om.containing(lon, lat, nn)


@pytest.mark.parametrize("order", [1, 2, 4])
def test_walk_3d_bitwise(gpu_ctx, world3d, order):
    mesh, snaps, lon, lat, z = world3d
    gpu_ctx.upload_mesh(mesh)
    for s, sn in enumerate(snaps):
        gpu_ctx.upload_field(s, *(a.astype(np.float32) for a in sn[:3]))
    om = OracleMesh(mesh, "kmath")
    elem = _seed_elems(om, lon, lat)
    pid = np.arange(lon.size)
    p = gpu_ctx.particles(lon, lat, z, pid=pid, elem=elem)
    st = dict(lon=lon, lat=lat, z=z, status=np.ones(lon.size, np.int32),
              nn=np.full(lon.size, -1, np.int32), pid=pid, elem=elem)
    kick = dict(kh_x=30.0, kh_y=30.0, seed=5)
    f32 = [tuple(a.astype(np.float32) for a in sn[:3]) for sn in snaps]
    for k in range(8):
        t = k * DT
        ia, ib, a0, it, dtn = synth.snapshot_bracket(t, synth.SNAP_DT, 3)
        c = {1: (0.0,), 2: (0.0, 0.5), 4: (0.0, 0.5, 0.5, 1.0)}[order]
        al = [ck * DT / dtn for ck in c] + [0.0] * (4 - order)
        f = dict(u_a=f32[ia][0], v_a=f32[ia][1], w_a=f32[ia][2], u_t=f32[it][0], v_t=f32[it][1],
                 w_t=f32[it][2], alpha0=a0)
        if ib >= 0:
            f.update(u_b=f32[ib][0], v_b=f32[ib][1], w_b=f32[ib][2])
        args = step_args(order, DT, 1.0, is3d=True, w_method=2, exact=1, alphas=al, kick=kick,
                         step_i=k, walk=True)
        st = om.step(st, f, args)
        gpu_ctx.step(p, order, ia, DT, 1.0, slot_b=ib, slot_t=it, alpha0=a0, alphas=al,
                     exact_nearest=1, w_method=2, kick=kick, step_i=k, walk=True)
    h = p.to_host()
    assert (h["status"] == 1).mean() > 0.9
    for key in ("status", "nn", "fr", "elem"):
        np.testing.assert_array_equal(h[key], st[key], err_msg=key)
    for key in ("lon", "lat", "z", "u", "v", "w"):
        np.testing.assert_array_equal(h[key].view(np.int64), st[key].view(np.int64), err_msg=key)


@pytest.mark.parametrize("order", [2, 4])
def test_walk_2d_periodic_bitwise_and_equals_local(gpu_ctx, world2d, order):
    w = world2d
    mesh = w.mesh
    gpu_ctx.upload_mesh(mesh)
    gpu_ctx.upload_field(0, w.u, w.v)
    om = OracleMesh(mesh, "kmath")
    elem = _seed_elems(om, w.lon, w.lat)
    window = dict(periodic_lon=(5.0, 35.0))
    runs = {}
    for walk in (True, False):
        p = gpu_ctx.particles(w.lon, w.lat, pid=np.arange(w.lon.size), elem=elem if walk else None)
        st = dict(lon=w.lon, lat=w.lat, status=np.ones(w.lon.size, np.int32),
                  nn=np.full(w.lon.size, -1, np.int32), elem=elem if walk else None)
        for k in range(12):
            gpu_ctx.step(p, order, 0, 600.0, 1.0, exact_nearest=1, step_i=k, walk=walk, **window)
            st = om.step(st, dict(u_a=w.u, v_a=w.v), step_args(order, 600.0, 1.0, exact=1, step_i=k,
                                                                walk=walk, **window))
        h = p.to_host()
        for key in ("status", "nn", "fr"):
            np.testing.assert_array_equal(h[key], st[key], err_msg=key)
        for key in ("lon", "lat", "u", "v"):
            np.testing.assert_array_equal(h[key].view(np.int64), st[key].view(np.int64), err_msg=key)
        if walk:
            np.testing.assert_array_equal(h["elem"], st["elem"])
        runs[walk] = h
    np.testing.assert_array_equal(runs[True]["status"], runs[False]["status"])
    np.testing.assert_allclose(runs[True]["lon"], runs[False]["lon"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(runs[True]["lat"], runs[False]["lat"], rtol=0, atol=1e-10)
