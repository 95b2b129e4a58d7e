"""Host-side mesh container and its search structures.

Mirrors the reference's ``Mesh`` dataclass (src/kelp/core/mesh.py:41) and the
host-side setup that runs once per mesh: the grid-hash CSR index
(``build_grid_hash``, core/mesh.py:844), its row budget (``_gh_far_rows_for``,
core/mesh.py:1042), the node -> element adjacency (``build_node_to_elems``,
core/mesh.py:2716) and the coastal-element flags (``build_boundary_info``,
core/mesh.py:2765).  These are integer/bookkeeping passes done with numpy in
the same order as the reference, so the CSR layout (and with it the scan order
that breaks nearest-node ties) is identical.  The per-node unit-vector table
is computed on the device at upload (kelp_amd/csrc/kelp_hip.hip).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, Optional

import numpy as np

DEG2RAD = np.pi / 180.0
RAD2DEG = 180.0 / np.pi
R_EARTH_STANDARD = 6_371_000.0

GRID_HASH_MAX_CELLS = 50_000_000     # core/mesh.py:710
GH_FAR_ROWS_MIN = 4                  # core/mesh.py:1005
GH_FAR_ROWS_MAX = 4096               # core/mesh.py:1006


@dataclass
class Mesh:
    """Unstructured mesh (triangles, quads or a mix), node-layout geometry.

    ``elem_nodes`` is ``(E, max_v)`` int32, ``-1`` padded; ``sigma`` holds the
    vertical interfaces (z-levels in metres, positive down; empty for 2-D).
    The grid-hash fields are filled by :func:`build_grid_hash`.
    """

    name: str
    nodes_lon: np.ndarray
    nodes_lat: np.ndarray
    depth: np.ndarray
    elem_nodes: np.ndarray
    elem_center_lon: np.ndarray
    elem_center_lat: np.ndarray
    node_to_elems: np.ndarray
    node_to_elems_count: np.ndarray
    sigma: np.ndarray = field(default_factory=lambda: np.array([], dtype=np.float64))
    vertical_coord: str = "zlevel"
    r_earth: float = R_EARTH_STANDARD
    metadata: Dict[str, Any] = field(default_factory=dict)

    gh_offsets: Optional[np.ndarray] = None
    gh_node_ids: Optional[np.ndarray] = None
    gh_cell_size: float = 0.0
    gh_lon_min: float = 0.0
    gh_lat_min: float = 0.0
    gh_nx: int = 0
    gh_ny: int = 0
    gh_wrap_lon: bool = False
    gh_far_rows: int = 16
    gh_max_edge_deg: Optional[float] = None
    elem_is_coastal: Optional[np.ndarray] = None
    elem_neighbors: Optional[np.ndarray] = None   # (E, max_v), build_elem_neighbors

    @property
    def n_nodes(self) -> int:
        return int(self.nodes_lon.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.elem_nodes.shape[0])

    @property
    def max_v(self) -> int:
        return int(self.elem_nodes.shape[1])

    @property
    def is_3d(self) -> bool:
        return self.sigma.size > 0

    @property
    def z_levels(self) -> np.ndarray:
        return self.sigma


def set_exact_nearest(mesh: Mesh, flag: bool) -> None:
    """The exact nearest-node anchor guard (core/mesh.py:321): on, every
    anchor is the provably nearest node; off, anchors come from the 3x3
    hash window (misses healed).  Stored in ``mesh.metadata``."""
    mesh.metadata["exact_nearest"] = bool(flag)


def exact_nearest_on(mesh: Mesh) -> bool:
    """numerics/integrators.py:3283 _exact_nearest_on."""
    md = getattr(mesh, "metadata", None)
    return bool(md and md.get("exact_nearest", False))


def build_node_to_elems(n_nodes: int, elem_nodes: np.ndarray):
    """Node -> element adjacency ``(N, max_conn)`` (-1 padded) and counts.

    Element ids appear in ascending order per node, each element once even
    when it repeats a vertex (core/mesh.py:2716-2762).
    """
    en = np.asarray(elem_nodes)
    E, mv = en.shape
    valid = en >= 0
    # drop repeated vertices within an element (keep the first occurrence)
    for k in range(1, mv):
        for j in range(k):
            valid[:, k] &= ~((en[:, k] == en[:, j]) & (en[:, j] >= 0))
    elem_ids = np.repeat(np.arange(E, dtype=np.int64)[:, None], mv, axis=1)[valid]
    nodes = en[valid].astype(np.int64)
    counts = np.bincount(nodes, minlength=n_nodes).astype(np.int32)
    max_conn = int(counts.max()) if n_nodes > 0 else 0
    order = np.lexsort((elem_ids, nodes))          # by node, then element id
    nodes_s = nodes[order]
    elems_s = elem_ids[order]
    starts = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(counts, out=starts[1:])
    slot = np.arange(nodes_s.size, dtype=np.int64) - starts[nodes_s]
    n2e = np.full((n_nodes, max(max_conn, 1)), -1, dtype=np.int32)
    n2e[nodes_s, slot] = elems_s.astype(np.int32)
    if max_conn == 0:
        n2e = n2e[:, :0]
    return n2e, counts


def _lon_branch(nodes_lon: np.ndarray):
    """core/mesh.py:793 _gh_lon_branch."""
    lon_0360 = (nodes_lon + 360.0) % 360.0
    lon_signed = (nodes_lon + 180.0) % 360.0 - 180.0
    span_0360 = lon_0360.max() - lon_0360.min()
    span_signed = lon_signed.max() - lon_signed.min()
    wrap = bool(min(span_0360, span_signed) > 350.0)
    if wrap or span_0360 <= span_signed:
        return lon_0360, wrap
    return lon_signed, wrap


def _grid_geometry(lon, lat, cell_size, wrap):
    """core/mesh.py:807 _gh_grid_geometry."""
    lat_min = lat.min() - cell_size
    lat_max = lat.max() + cell_size
    ny = int(np.ceil((lat_max - lat_min) / cell_size))
    if wrap:
        lon_min = 0.0
        nx = int(np.ceil(360.0 / cell_size))
    else:
        lon_min = lon.min() - cell_size
        lon_max = lon.max() + cell_size
        nx = int(np.ceil((lon_max - lon_min) / cell_size))
    return float(lon_min), float(lat_min), nx, ny


def _cell_index(lon, lat, cell_size, lon_min, lat_min, nx, ny, wrap):
    """core/mesh.py:824 _gh_cell_index."""
    if wrap:
        ci = np.minimum((lon / cell_size).astype(np.int32), nx - 1)
    else:
        ci = np.clip(((lon - lon_min) / cell_size).astype(np.int32), 0, nx - 1)
    cj = np.clip(((lat - lat_min) / cell_size).astype(np.int32), 0, ny - 1)
    return cj.astype(np.int64) * nx + ci


def max_edge_deg(mesh: Mesh) -> float:
    """Longest element edge in degrees of great circle (core/mesh.py:1009)."""
    en = mesh.elem_nodes
    if en is None or en.size == 0:
        return 0.0
    lon = np.asarray(mesh.nodes_lon, dtype=np.float64)
    lat = np.asarray(mesh.nodes_lat, dtype=np.float64)
    nv = en.shape[1]
    best = 0.0
    step = 1_000_000
    for e0 in range(0, en.shape[0], step):
        blk = en[e0:e0 + step]
        last_valid = np.clip((blk >= 0).sum(axis=1) - 1, 0, nv - 1)
        pairs = [(blk[:, k], blk[:, (k + 1) % nv]) for k in range(nv)]
        pairs.append((blk[np.arange(blk.shape[0]), last_valid], blk[:, 0]))
        for a, b in pairs:
            ok = (a >= 0) & (b >= 0)
            if not ok.any():
                continue
            a = a[ok]
            b = b[ok]
            dl = ((((lon[b] - lon[a]) + 180.0) % 360.0) - 180.0) * DEG2RAD
            la1 = lat[a] * DEG2RAD
            la2 = lat[b] * DEG2RAD
            cosd = np.sin(la1) * np.sin(la2) + np.cos(la1) * np.cos(la2) * np.cos(dl)
            d = np.arccos(np.clip(cosd, -1.0, 1.0)) * RAD2DEG
            best = max(best, float(d.max()))
    return best


def far_rows_for(mesh: Mesh, cell_size: float) -> int:
    """Exact-search row budget (core/mesh.py:1042 _gh_far_rows_for)."""
    if mesh.gh_max_edge_deg is None:
        mesh.gh_max_edge_deg = max_edge_deg(mesh)
    rows = int(np.ceil(mesh.gh_max_edge_deg / cell_size - 1.0e-9)) + 2
    rows = min(rows, GH_FAR_ROWS_MAX)
    return int(max(rows, GH_FAR_ROWS_MIN))


def build_grid_hash(mesh: Mesh, cell_size: float = 0.0) -> None:
    """Uniform lon/lat grid hash over the nodes (core/mesh.py:844).

    ``cell_size`` 0 = the pinned ``metadata['grid_hash_cell']`` or the auto
    size ``max(0.05, 2*sqrt(area/N))``.  Nodes are bucketed by
    ``cj*nx + ci`` in the contiguous longitude branch; within a cell they
    keep ascending node order (the reference's fill loop), which is the
    scan order that decides nearest-node ties.
    """
    lon, wrap = _lon_branch(np.asarray(mesh.nodes_lon, dtype=np.float64))
    lat = np.asarray(mesh.nodes_lat, dtype=np.float64)
    N = lon.shape[0]
    if not cell_size > 0:
        pinned = (mesh.metadata or {}).get("grid_hash_cell")
        if pinned is not None:
            cell_size = float(pinned)
            if not np.isfinite(cell_size) or not 0.0 < cell_size < 45.0:
                raise ValueError(f"grid_hash_cell must be in (0, 45) degrees, got {pinned!r}")
        else:
            lon_span = lon.max() - lon.min()
            lat_span = lat.max() - lat.min()
            area = max(lon_span * lat_span, 1.0)
            cell_size = max(0.05, np.sqrt(area / N) * 2)
    lon_min, lat_min, nx, ny = _grid_geometry(lon, lat, cell_size, wrap)
    n_cells = nx * ny
    if n_cells > GRID_HASH_MAX_CELLS:
        raise ValueError(f"grid hash: cell size {cell_size:.6g} deg needs {n_cells} cells "
                         f"(> {GRID_HASH_MAX_CELLS}); use a larger cell")
    cell_idx = _cell_index(lon, lat, cell_size, lon_min, lat_min, nx, ny, wrap)
    counts = np.bincount(cell_idx, minlength=n_cells).astype(np.int32)
    offsets = np.zeros(n_cells + 1, dtype=np.int32)
    np.cumsum(counts, out=offsets[1:])
    # stable sort by cell = the reference's ascending fill loop
    node_ids = np.argsort(cell_idx, kind="stable").astype(np.int32)
    mesh.gh_offsets = offsets
    mesh.gh_node_ids = node_ids
    mesh.gh_cell_size = float(cell_size)
    mesh.gh_lon_min = lon_min
    mesh.gh_lat_min = float(lat_min)
    mesh.gh_nx = int(nx)
    mesh.gh_ny = int(ny)
    mesh.gh_wrap_lon = wrap
    mesh.gh_far_rows = far_rows_for(mesh, cell_size)


def build_boundary_info(mesh: Mesh) -> np.ndarray:
    """Coastal elements: those owning an edge no other element shares
    (core/mesh.py:2765).  Cached on ``mesh.elem_is_coastal``."""
    if mesh.elem_is_coastal is not None:
        return mesh.elem_is_coastal
    en = np.asarray(mesh.elem_nodes)
    E, mv = en.shape
    rows = []
    for e in range(E):
        verts = []
        for k in range(mv):
            nd = int(en[e, k])
            if nd < 0 or (verts and nd == verts[0]):
                continue
            verts.append(nd)
        nv = len(verts)
        if nv < 3:
            continue
        seen = set()
        for j in range(nv):
            a, b = verts[j], verts[(j + 1) % nv]
            if a == b:
                continue
            ek = (a, b) if a < b else (b, a)
            if ek in seen:
                continue
            seen.add(ek)
            rows.append((ek[0], ek[1], e))
    coastal = np.zeros(E, dtype=np.bool_)
    if rows:
        arr = np.asarray(rows, dtype=np.int64)
        key = arr[:, 0] * np.int64(mesh.n_nodes + 1) + arr[:, 1]
        _, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
        coastal[arr[cnt[inv] == 1, 2]] = True
    mesh.elem_is_coastal = coastal
    return coastal


def elem_centers(nodes_lon, nodes_lat, elem_nodes):
    """Element centres as the normalised mean of the member unit vectors."""
    lon_r = np.asarray(nodes_lon) * DEG2RAD
    lat_r = np.asarray(nodes_lat) * DEG2RAD
    c = np.cos(lat_r)
    xyz = np.stack((c * np.cos(lon_r), c * np.sin(lon_r), np.sin(lat_r)), axis=1)
    valid = elem_nodes >= 0
    V = xyz[np.where(valid, elem_nodes, 0)] * valid[:, :, None]
    C = V.sum(axis=1)
    C /= np.linalg.norm(C, axis=1, keepdims=True)
    lon = np.degrees(np.arctan2(C[:, 1], C[:, 0]))
    lat = np.degrees(np.arcsin(np.clip(C[:, 2], -1.0, 1.0)))
    return lon, lat


def make_mesh(name, nodes_lon, nodes_lat, elem_nodes, depth=None, z_levels=None,
              r_earth=R_EARTH_STANDARD, cell_size=0.0, metadata=None) -> Mesh:
    """Assemble a :class:`Mesh` with adjacency and grid hash built."""
    nodes_lon = np.ascontiguousarray(nodes_lon, dtype=np.float64)
    nodes_lat = np.ascontiguousarray(nodes_lat, dtype=np.float64)
    elem_nodes = np.ascontiguousarray(elem_nodes, dtype=np.int32)
    N = nodes_lon.shape[0]
    n2e, cnt = build_node_to_elems(N, elem_nodes)
    clon, clat = elem_centers(nodes_lon, nodes_lat, elem_nodes)
    mesh = Mesh(
        name=name, nodes_lon=nodes_lon, nodes_lat=nodes_lat,
        depth=(np.zeros(N) if depth is None else np.ascontiguousarray(depth, dtype=np.float64)),
        elem_nodes=elem_nodes, elem_center_lon=clon, elem_center_lat=clat,
        node_to_elems=n2e, node_to_elems_count=cnt,
        sigma=(np.array([], dtype=np.float64) if z_levels is None
               else np.abs(np.asarray(z_levels, dtype=np.float64))),
        r_earth=float(r_earth), metadata=dict(metadata or {}))
    build_grid_hash(mesh, cell_size)
    return mesh


def build_elem_neighbors(mesh: Mesh) -> np.ndarray:
    """core/mesh.py:2824 build_elem_neighbors: slot k of row e is the element
    sharing the edge (verts[k], verts[(k+1) % nv]) of e's compacted vertex
    ring, -1 on a boundary edge (and in unused slots); an edge shared by
    more than two elements raises.  Cached as ``mesh.elem_neighbors``."""
    cached = getattr(mesh, "elem_neighbors", None)
    if cached is not None:
        return cached
    en = np.asarray(mesh.elem_nodes, dtype=np.int64)
    E, mv = en.shape if en.size else (0, 0)
    out = np.full((E, mv), -1, dtype=np.int32)
    rows = []
    for e in range(E):
        verts = []
        for nd in en[e]:
            nd = int(nd)
            if nd < 0 or (verts and nd == verts[0]):
                continue
            verts.append(nd)
        nv = len(verts)
        if nv < 3:
            continue
        seen = set()
        for k in range(nv):
            a, b = verts[k], verts[(k + 1) % nv]
            if a == b:
                continue
            ek = (a, b) if a < b else (b, a)
            if ek in seen:
                continue
            seen.add(ek)
            rows.append((ek[0], ek[1], e, k))
    if rows:
        arr = np.asarray(rows, dtype=np.int64)
        key = arr[:, 0] * np.int64(mesh.n_nodes + 1) + arr[:, 1]
        order = np.argsort(key, kind="stable")
        key_s, arr_s = key[order], arr[order]
        starts = np.flatnonzero(np.r_[True, key_s[1:] != key_s[:-1]])
        counts = np.diff(np.r_[starts, key_s.size])
        if (counts > 2).any():
            bad = arr_s[starts[counts > 2][0]]
            raise ValueError(f"non-manifold mesh: edge ({bad[0]}, {bad[1]}) is shared by more "
                             "than two elements")
        two = starts[counts == 2]
        e1, k1 = arr_s[two, 2], arr_s[two, 3]
        e2, k2 = arr_s[two + 1, 2], arr_s[two + 1, 3]
        out[e1, k1] = e2
        out[e2, k2] = e1
    mesh.elem_neighbors = out
    return out


def _normalized_connectivity(elem_nodes) -> np.ndarray:
    """``elem_nodes`` as int32 with no repeated vertex in a row (a FESOM-C
    triangle stored as a quad repeats its first vertex, core/mesh.py:52-56):
    the array itself when it already qualifies, else a compacted copy."""
    en = np.asarray(elem_nodes)
    if en.dtype != np.int32:
        en = en.astype(np.int32)
    if en.ndim != 2 or en.shape[1] < 2:
        return en
    rep = np.zeros(en.shape[0], dtype=bool)
    for k in range(1, en.shape[1]):
        for j in range(k):
            rep |= (en[:, k] == en[:, j]) & (en[:, k] >= 0)
    if not rep.any():
        return en
    out = en.copy()
    for e in np.flatnonzero(rep):
        seen = []
        for nd in en[e]:
            nd = int(nd)
            if nd >= 0 and nd not in seen:
                seen.append(nd)
        out[e] = -1
        out[e, :len(seen)] = seen
    return out


def from_reference_mesh(ref) -> Mesh:
    """This package's :class:`Mesh` over a reference ``kelp.core.mesh.Mesh``
    (src/kelp/core/mesh.py:41-210), sharing its arrays (no copies) and its
    ``metadata`` dict (so ``set_exact_nearest`` on either is seen by both).

    The grid hash is the reference's own (``_gh_offsets``, ``_gh_node_ids``,
    ``_gh_cell_size`` ... ``_gh_far_rows``, built by its ``build_grid_hash``,
    core/mesh.py:844) when present, else built here the same way; the
    coastal flags (``_elem_is_coastal``) and element neighbours
    (``_elem_neighbors``) are taken when the reference has built them.  The
    adapter is memoised on the reference object and rebuilt only when one of
    those arrays or the cell changes, so a device context bound to it keeps
    its uploaded mesh and snapshots across calls."""
    if isinstance(ref, Mesh):
        return ref
    gh_off = getattr(ref, "_gh_offsets", None)
    key = (id(ref.nodes_lon), id(ref.elem_nodes), id(gh_off), float(getattr(ref, "_gh_cell_size", 0.0)),
           id(getattr(ref, "_elem_neighbors", None)), id(getattr(ref, "_elem_is_coastal", None)),
           id(ref.sigma))
    memo = getattr(ref, "__dict__", {}).get("_kelp_amd_mesh")
    if memo is not None and memo[0] == key:
        return memo[1]
    md = getattr(ref, "metadata", None)
    if md is None:
        md = {}
    m = Mesh(name=str(getattr(ref, "name", "mesh")),
             nodes_lon=np.asarray(ref.nodes_lon, dtype=np.float64),
             nodes_lat=np.asarray(ref.nodes_lat, dtype=np.float64),
             depth=np.asarray(ref.depth, dtype=np.float64),
             elem_nodes=_normalized_connectivity(ref.elem_nodes),
             elem_center_lon=np.asarray(ref.elem_center_lon, dtype=np.float64),
             elem_center_lat=np.asarray(ref.elem_center_lat, dtype=np.float64),
             node_to_elems=np.asarray(ref.node_to_elems, dtype=np.int32),
             node_to_elems_count=np.asarray(ref.node_to_elems_count, dtype=np.int32),
             sigma=np.asarray(getattr(ref, "sigma", np.array([], dtype=np.float64)), dtype=np.float64),
             vertical_coord=str(getattr(ref, "vertical_coord", "zlevel")),
             r_earth=float(getattr(ref, "r_earth", R_EARTH_STANDARD)),
             metadata=md)
    if gh_off is not None:
        m.gh_offsets = np.asarray(gh_off, dtype=np.int32)
        m.gh_node_ids = np.asarray(ref._gh_node_ids, dtype=np.int32)
        m.gh_cell_size = float(ref._gh_cell_size)
        m.gh_lon_min = float(ref._gh_lon_min)
        m.gh_lat_min = float(ref._gh_lat_min)
        m.gh_nx = int(ref._gh_nx)
        m.gh_ny = int(ref._gh_ny)
        m.gh_wrap_lon = bool(ref._gh_wrap_lon)
        m.gh_far_rows = int(ref._gh_far_rows)
        m.gh_max_edge_deg = getattr(ref, "_gh_max_edge_deg", None)
    else:
        build_grid_hash(m)
    coastal = getattr(ref, "_elem_is_coastal", None)
    if coastal is not None:
        m.elem_is_coastal = np.asarray(coastal, dtype=bool)
    nbr = getattr(ref, "_elem_neighbors", None)
    if nbr is not None:
        m.elem_neighbors = np.asarray(nbr, dtype=np.int32)
    try:
        ref.__dict__["_kelp_amd_mesh"] = (key, m)
    except (AttributeError, TypeError):
        pass
    return m
