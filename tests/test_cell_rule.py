"""The grid-hash cell rule (kelp_amd.cellrule, restating core/mesh.py:464
choose_grid_hash_cell and :386 _gh_sample_stats) with the oracle's
statistics (CPU): the statistics against brute force, and the choices."""
import numpy as np

from kelp_amd import synth
from kelp_amd.cellrule import apply_grid_hash_cell_rule, choose_grid_hash_cell
from kelp_amd.mesh import _cell_index, _grid_geometry, _lon_branch, set_exact_nearest
from oracle.oracle import OracleMesh

from .golden import reference_cases as rc


class OracleStats:
    def __init__(self, mesh):
        self.om = OracleMesh(mesh, "glibc")

    def nearest(self, qlon, qlat):
        return self.om.nearest(qlon, qlat, exact=True)

    def stats(self, *args):
        return self.om.gh_sample_stats(*args)


def _candidate(mesh, cell):
    lon_b, wrap = _lon_branch(mesh.nodes_lon)
    lon_min, lat_min, nx, ny = _grid_geometry(lon_b, mesh.nodes_lat, cell, wrap)
    idx = _cell_index(lon_b, mesh.nodes_lat, cell, lon_min, lat_min, nx, ny, wrap)
    counts = np.bincount(idx, minlength=nx * ny).reshape(ny, nx)
    cum = np.zeros((ny + 1, nx + 1), np.int32)
    cum[1:, 1:] = counts.cumsum(0).cumsum(1)
    return lon_min, lat_min, nx, ny, wrap, counts, cum


def test_sample_stats_against_brute_force():
    """3x3 occupancy = the nodes counted cell by cell; the approximate flag
    = the brute-force nearest node lies outside the 3x3 window; a provable
    sample costs its occupancy + 9 cells + 1 certificate (20)."""
    mesh = synth.lattice_mesh(step=0.5)
    om = OracleMesh(mesh, "glibc")
    q = np.stack([mesh.elem_center_lon, mesh.elem_center_lat], 1)[::3]
    rng = np.random.default_rng(3)
    q = q + rng.uniform(-0.2, 0.2, q.shape)
    nn = om.nearest(q[:, 0], q[:, 1], exact=True)
    d2 = ((np.radians(mesh.nodes_lat)[None, :] - np.radians(q[:, 1])[:, None]) ** 2
          + (np.cos(np.radians(q[:, 1]))[:, None] * np.radians(mesh.nodes_lon[None, :] - q[:, 0][:, None])) ** 2)
    assert (nn == d2.argmin(1)).mean() > 0.99
    seen = set()
    for k in range(6):
        cell = mesh.gh_cell_size / np.sqrt(2.0) ** k
        lon_min, lat_min, nx, ny, wrap, counts, cum = _candidate(mesh, cell)
        occ, cost, unprov, approx = om.gh_sample_stats(q[:, 0], q[:, 1], nn, cum, cell, lon_min, lat_min, nx, ny,
                                                       wrap)
        ci = np.clip(((q[:, 0] - lon_min) / cell).astype(int), 0, nx - 1)
        cj = np.clip(((q[:, 1] - lat_min) / cell).astype(int), 0, ny - 1)
        pad = np.pad(counts, 1)
        brute = sum(pad[cj + 1 + dj, ci + 1 + di] for dj in (-1, 0, 1) for di in (-1, 0, 1))
        np.testing.assert_array_equal(occ, brute)
        cin = np.clip(((mesh.nodes_lon[nn] - lon_min) / cell).astype(int), 0, nx - 1)
        cjn = np.clip(((mesh.nodes_lat[nn] - lat_min) / cell).astype(int), 0, ny - 1)
        far = (np.abs(cin - ci) > 1) | (np.abs(cjn - cj) > 1)
        np.testing.assert_array_equal(approx.astype(bool), far)
        ok = unprov == 0
        seen.update(set(np.unique(unprov).tolist()))
        np.testing.assert_array_equal(cost[ok], occ[ok] + 29.0)
        assert np.all(cost[~ok] >= occ[~ok] + 29.0)
        assert np.all(~ok[far | (occ == 0)])
    assert seen == {0, 1}


def test_rule_keeps_the_legacy_cell_of_the_reference_run():
    """The reference's own run_tracking output on the FESOM2 unit-square mesh
    records grid_hash_cell 1.0 under the occupancy rule (exact mode): the
    restated rule decides the same."""
    case = rc.multiyear_case()
    mesh = case["mesh"]
    set_exact_nearest(mesh, True)
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_golden_3d",
                                "run_tracking_fesom2_multiyear_node.npz"), allow_pickle=False)
    cell, rep = choose_grid_hash_cell(mesh, OracleStats(mesh))
    assert cell == 1.0 and rep["mode"] == "exact"
    assert float(gold["attr__grid_hash_cell"]) == cell
    assert str(gold["attr__grid_hash_cell_rule"]) == "occupancy"
    rep2 = apply_grid_hash_cell_rule(mesh, OracleStats(mesh), "occupancy")
    assert mesh.gh_cell_size == 1.0 and mesh.metadata["grid_hash_cell_rule"] == "occupancy"
    assert rep2["chosen"] == 1.0


def test_approximate_mode_picks_the_smallest_dense_cell():
    """Guard off: the smallest candidate with no empty window, median
    occupancy >= 8 and no added approximate anchors; every candidate below
    it fails one of those."""
    mesh = synth.lattice_mesh(step=0.25)
    set_exact_nearest(mesh, False)
    cell, rep = choose_grid_hash_cell(mesh, OracleStats(mesh), sample=5000)
    rows = rep["candidates"]
    chosen = [r for r in rows if r["cell"] == cell][0]
    assert chosen["empty"] == 0.0 and chosen["median"] >= 8
    base = rows[0]
    for r in rows:
        if r["cell"] < cell:
            assert (r["empty"] > 0.0 or r["median"] < 8 or r["approx"] > base["approx"] + 0.005)


def test_explicit_pin_is_kept():
    mesh = synth.lattice_mesh(step=0.5)
    mesh.metadata["grid_hash_cell"] = 0.7
    rep = apply_grid_hash_cell_rule(mesh, OracleStats(mesh), "occupancy")
    assert rep["rule"] == "explicit"
