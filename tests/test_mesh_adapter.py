"""``from_reference_mesh``: this package's Mesh over a reference-shaped one
(src/kelp/core/mesh.py:41-210) without copying its arrays (CPU)."""
import numpy as np

from kelp_amd import synth
from kelp_amd.mesh import build_grid_hash, from_reference_mesh, make_mesh

from .test_gpu_registry import _as_reference


def test_adapter_shares_arrays_and_memoises():
    mesh = synth.sphere_mesh(3000)
    ref = _as_reference(mesh)
    m = from_reference_mesh(ref)
    assert m is from_reference_mesh(ref)
    for a, b in ((m.nodes_lon, ref.nodes_lon), (m.elem_nodes, ref.elem_nodes),
                 (m.node_to_elems, ref.node_to_elems), (m.gh_offsets, ref._gh_offsets),
                 (m.gh_node_ids, ref._gh_node_ids)):
        assert a is b
    assert m.metadata is ref.metadata
    assert (m.gh_cell_size, m.gh_nx, m.gh_ny, m.gh_far_rows) == (ref._gh_cell_size, ref._gh_nx, ref._gh_ny,
                                                               ref._gh_far_rows)
    # a rebuilt hash on the reference side makes a new adapter
    ref._gh_offsets = ref._gh_offsets.copy()
    m2 = from_reference_mesh(ref)
    assert m2 is not m and m2.gh_offsets is ref._gh_offsets
    # this package's Mesh passes through
    assert from_reference_mesh(mesh) is mesh


def test_adapter_builds_missing_hash_and_compacts_repeated_vertex():
    lon = np.array([0.0, 1.0, 1.0, 0.0, 2.0])
    lat = np.array([0.0, 0.0, 1.0, 1.0, 0.5])
    en = np.array([[0, 1, 2, 3], [1, 4, 2, 1]], np.int32)     # a triangle repeating its first vertex
    own = make_mesh("t", lon, lat, np.array([[0, 1, 2, 3], [1, 4, 2, -1]], np.int32))
    ref = _as_reference(own)
    ref.elem_nodes = en
    ref._gh_offsets = None
    m = from_reference_mesh(ref)
    np.testing.assert_array_equal(m.elem_nodes, [[0, 1, 2, 3], [1, 4, 2, -1]])
    build_grid_hash(own)
    np.testing.assert_array_equal(m.gh_offsets, own.gh_offsets)
    np.testing.assert_array_equal(m.gh_node_ids, own.gh_node_ids)
    assert m.gh_cell_size == own.gh_cell_size
