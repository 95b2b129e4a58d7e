"""The step kernel's grid-hash tile in LDS (GPU).

Every workgroup of the step kernel stages the hash cells around its alive
particles' start cells in LDS and scans them there; a window outside the
staged box is scanned in HBM.  The answer must not depend on which: this
test drives both paths in one launch -- a dense patch whose sorted
workgroups are stageable, a sparse spread whose workgroups are not, a
random walk strong enough to kick particles out of their workgroup's box,
and hash cells small enough that the exact search often widens (its wide
row scans) -- and holds the device bit for bit to the oracle (built with
the device's math), whose search has no tile.
"""
import numpy as np
import pytest

from kelp_amd import synth
from kelp_amd.mesh import build_grid_hash
from oracle.oracle import OracleMesh, step_args

pytestmark = pytest.mark.gpu

# the staging rule of the kernel (kelp_device.cuh: stage_gh_tile)
TILE_ROWS, TILE_COLS, TILE_NODES, MARGIN = 8, 64, 1536, 2


def _cells(mesh, lon, lat):
    ci = np.clip(((lon - mesh.gh_lon_min) / mesh.gh_cell_size).astype(np.int64), 0, mesh.gh_nx - 1)
    cj = np.clip(((lat - mesh.gh_lat_min) / mesh.gh_cell_size).astype(np.int64), 0, mesh.gh_ny - 1)
    return ci, cj


def _stageable(mesh, lon, lat, status):
    """Per 256-particle workgroup: does the kernel's rule stage its box?"""
    ci, cj = _cells(mesh, lon, lat)
    out = []
    for s in range(0, lon.size, 256):
        a = status[s:s + 256] == 1
        if not a.any():
            out.append(False)
            continue
        i0 = max(ci[s:s + 256][a].min() - MARGIN, 0)
        i1 = min(ci[s:s + 256][a].max() + MARGIN, mesh.gh_nx - 1)
        j0 = max(cj[s:s + 256][a].min() - MARGIN, 0)
        j1 = min(cj[s:s + 256][a].max() + MARGIN, mesh.gh_ny - 1)
        ok = i1 - i0 + 1 <= TILE_COLS and j1 - j0 + 1 <= TILE_ROWS
        if ok:
            off = mesh.gh_offsets
            nodes = sum(off[j * mesh.gh_nx + i1 + 1] - off[j * mesh.gh_nx + i0] for j in range(j0, j1 + 1))
            ok = nodes <= TILE_NODES
        out.append(ok)
    return np.array(out)


def test_tiled_anchor_scans_match_oracle(gpu_ctx):
    mesh = synth.lattice_mesh()
    build_grid_hash(mesh, 0.1)      # small cells: kicks leave the 2-cell margin
    u, v = synth.double_gyre(mesh)
    rng = np.random.default_rng(77)
    n_dense, n_sparse = 24_000, 4_000
    lon = np.concatenate([rng.uniform(12.0, 15.0, n_dense), rng.uniform(1.0, 39.0, n_sparse)])
    lat = np.concatenate([rng.uniform(40.0, 42.0, n_dense), rng.uniform(31.0, 59.0, n_sparse)])
    pid = rng.permutation(lon.size).astype(np.int64)
    gpu_ctx.upload_mesh(mesh)
    gpu_ctx.upload_field(0, u, v)
    p = gpu_ctx.particles(lon, lat, pid=pid)
    om = OracleMesh(mesh, "kmath")
    st = dict(lon=lon, lat=lat, status=np.ones(lon.size, np.int32), nn=np.full(lon.size, -1, np.int32),
              pid=pid)
    kick = dict(kh_x=2.0e4, kh_y=2.0e4, seed=404)
    dt = 21600.0
    staged_frac = []
    for k in range(6):
        if k % 2 == 0:
            gpu_ctx.sort(p)
            h = p.to_host()
            staged_frac.append(_stageable(mesh, h["lon"], h["lat"], h["status"]).mean())
        args = step_args(2, dt, 1.0, kick=kick, step_i=k)
        st = om.step(st, dict(u_a=u, v_a=v), args)
        gpu_ctx.step(p, 2, 0, dt, 1.0, kick=kick, step_i=k)
    gpu_ctx.synchronize()
    h = p.to_host()
    order = np.argsort(h["pid"])
    want = np.argsort(st["pid"])
    # most workgroups of the sorted state stage their tile, some do not
    assert 0.5 < min(staged_frac) and max(staged_frac) < 1.0, staged_frac
    for key in ("lon", "lat", "u", "v"):
        assert np.array_equal(h[key][order].view(np.int64), st[key][want].view(np.int64)), key
    for key in ("status", "nn", "fr"):
        assert np.array_equal(h[key][order], st[key][want]), key
    assert (st["status"] == 1).mean() > 0.9


def test_tile_switch_is_result_neutral(gpu_ctx, monkeypatch):
    """A context created with KELP_GH_TILE=0 scans every window in HBM; its
    steps equal the tiled ones bit for bit (RK4, 3-D, kick and beaching)."""
    from kelp_amd.device import DeviceContext

    mesh = synth.sphere_mesh(20000)
    snap = synth.jet_rossby_snapshot(mesh, 0.0)
    rng = np.random.default_rng(9)
    lon, lat, z = synth.wet_area_seeds(mesh, 200000)
    box = (lon > 20) & (lon < 50) & (lat > -10) & (lat < 10)
    lon, lat, z = lon[box], lat[box], z[box]
    pid = rng.permutation(lon.size).astype(np.int64)
    kick = dict(kh_x=10.0, kh_y=10.0, seed=5)
    beach = dict(criteria=["depth", "on_failure", "stochastic"], depth_min=60.0, t_beach=86400.0,
                 resuspension="auto", min_wet_steps=2)
    monkeypatch.setenv("KELP_GH_TILE", "0")
    plain = DeviceContext(0)
    monkeypatch.delenv("KELP_GH_TILE")
    runs = []
    try:
        for ctx in (gpu_ctx, plain):
            ctx.upload_mesh(mesh)
            ctx.upload_field(0, snap[0], snap[1], snap[2])
            p = ctx.particles(lon, lat, z, pid=pid)
            ctx.sort(p)
            for k in range(8):
                ctx.step(p, 4, 0, 1800.0, 1.0, w_method=2, kick=kick, beach=beach, step_i=k)
            ctx.synchronize()
            h = p.to_host()
            o = np.argsort(h["pid"])
            runs.append({k: v[o] for k, v in h.items()})
    finally:
        plain.close()
    a, b = runs
    for key in ("lon", "lat", "z", "u", "v", "w"):
        assert np.array_equal(a[key].view(np.int64), b[key].view(np.int64)), key
    for key in ("status", "nn", "fr", "ws"):
        assert np.array_equal(a[key], b[key]), key
