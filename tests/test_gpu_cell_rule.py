"""The grid-hash cell rule's per-sample statistics on the device
(kelp_gh_sample_stats) against the oracle (GPU): identical occupancy,
scan-cost proxy, widening and approximate-anchor flags for every candidate
of the ladder, and the same choice, on a global Fibonacci mesh with land
(wrap grid) and on a regional lattice, both guard modes."""
import numpy as np
import pytest

from kelp_amd import synth
from kelp_amd.cellrule import DeviceStats, choose_grid_hash_cell
from kelp_amd.mesh import set_exact_nearest
from oracle.oracle import OracleMesh

pytestmark = pytest.mark.gpu


class OracleStats:
    def __init__(self, mesh):
        self.om = OracleMesh(mesh, "kmath")

    def nearest(self, qlon, qlat):
        return self.om.nearest(qlon, qlat, exact=True)

    def stats(self, *args):
        return self.om.gh_sample_stats(*args)


@pytest.mark.parametrize("which,exact", [("sphere", True), ("sphere", False), ("lattice", True),
                                         ("lattice", False)])
def test_device_stats_and_choice_equal_oracle(gpu_ctx, which, exact):
    mesh = synth.sphere_mesh(40000) if which == "sphere" else synth.lattice_mesh(step=0.25)
    set_exact_nearest(mesh, exact)
    gpu_ctx.upload_mesh(mesh)
    dev, ora = DeviceStats(gpu_ctx), OracleStats(mesh)
    calls = []

    class Both:
        def nearest(self, qlon, qlat):
            a, b = dev.nearest(qlon, qlat), ora.nearest(qlon, qlat)
            np.testing.assert_array_equal(a, b)
            return a

        def stats(self, *args):
            a, b = dev.stats(*args), ora.stats(*args)
            for x, y in zip(a, b):
                assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
            calls.append(1)
            return a

    cell, rep = choose_grid_hash_cell(mesh, Both(), sample=20000)
    cell_o, rep_o = choose_grid_hash_cell(mesh, ora, sample=20000)
    assert cell == cell_o and len(calls) == len(rep["candidates"]) >= 2
    assert rep["candidates"] == rep_o["candidates"]
