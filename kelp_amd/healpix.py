"""HEALPix NEST geometry and the dual mesh of a HEALPix grid (no healpy).

The pixel centres and the 8-neighbourhood follow the HEALPix scheme of
Gorski et al. (2005, ApJ 622:759): a NEST index splits into a base face
(12 faces) and the bit-interleaved (x, y) position inside it; centres come
from the face's ring/phi offsets, neighbours across face edges from the
face-adjacency and axis-swap tables of the scheme.  The dual mesh
(pixel centres = nodes, one cell per pixel corner) restates
src/kelp/readers/healpix.py:93 build_healpix_dual_mesh: four corner cells
per pixel (owner, edge neighbour, diagonal, edge neighbour), kept when every
incident pixel is wet (an absent diagonal = valence-3 corner = triangle),
deduplicated by the smallest member, sorted by (owner, corner).
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh, build_node_to_elems, make_mesh

# face layout of the scheme: ring index (in units of nside) and phi offset
_JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4], dtype=np.int64)
_JPLL = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7], dtype=np.int64)

# neighbour offsets in (x, y), in the order SW, W, NW, N, NE, E, SE, S
_NB_X = np.array([-1, -1, 0, 1, 1, 1, 0, -1], dtype=np.int64)
_NB_Y = np.array([0, 1, 1, 1, 0, -1, -1, -1], dtype=np.int64)

# face across an edge/corner, indexed [direction 0..8][face]; direction
# = 4 + dx + 3*dy for the out-of-face offsets (dx, dy) in {-1, 0, 1}
_NB_FACE = np.array([
    [8, 9, 10, 11, -1, -1, -1, -1, 10, 11, 8, 9],     # S
    [5, 6, 7, 4, 8, 9, 10, 11, 9, 10, 11, 8],         # SE
    [-1, -1, -1, -1, 5, 6, 7, 4, -1, -1, -1, -1],     # E
    [4, 5, 6, 7, 11, 8, 9, 10, 11, 8, 9, 10],         # SW
    [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11],           # centre
    [1, 2, 3, 0, 0, 1, 2, 3, 5, 6, 7, 4],             # NE
    [-1, -1, -1, -1, 7, 4, 5, 6, -1, -1, -1, -1],     # W
    [3, 0, 1, 2, 3, 0, 1, 2, 4, 5, 6, 7],             # NW
    [2, 3, 0, 1, -1, -1, -1, -1, 0, 1, 2, 3],         # N
], dtype=np.int64)
# coordinate transform entering the neighbour face, [direction][face // 4]:
# bit 1 flips x, bit 2 flips y, bit 4 swaps x and y
_NB_SWAP = np.array([
    [0, 0, 3], [0, 0, 6], [0, 0, 0], [0, 0, 5], [0, 0, 0],
    [5, 0, 0], [0, 0, 0], [6, 0, 0], [3, 0, 0],
], dtype=np.int64)


def _order(nside: int) -> int:
    o = int(nside).bit_length() - 1
    if nside < 1 or (1 << o) != nside:
        raise ValueError(f"nside must be a positive power of 2, got {nside}")
    return o


def _compress_bits(v):
    """Even bits of ``v`` packed together (inverse of spreading)."""
    v = v & 0x5555555555555555
    v = (v | (v >> 1)) & 0x3333333333333333
    v = (v | (v >> 2)) & 0x0F0F0F0F0F0F0F0F
    v = (v | (v >> 4)) & 0x00FF00FF00FF00FF
    v = (v | (v >> 8)) & 0x0000FFFF0000FFFF
    v = (v | (v >> 16)) & 0x00000000FFFFFFFF
    return v


def _spread_bits(v):
    v = v & 0x00000000FFFFFFFF
    v = (v | (v << 16)) & 0x0000FFFF0000FFFF
    v = (v | (v << 8)) & 0x00FF00FF00FF00FF
    v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0F
    v = (v | (v << 2)) & 0x3333333333333333
    v = (v | (v << 1)) & 0x5555555555555555
    return v


def nest2xyf(nside, pix):
    order = _order(nside)
    pix = np.asarray(pix, dtype=np.int64)
    face = pix >> (2 * order)
    ipf = pix & ((1 << (2 * order)) - 1)
    return _compress_bits(ipf), _compress_bits(ipf >> 1), face


def xyf2nest(nside, x, y, face):
    order = _order(nside)
    return (np.asarray(face, np.int64) << (2 * order)) + _spread_bits(np.asarray(x, np.int64)) \
        + (_spread_bits(np.asarray(y, np.int64)) << 1)


def pix2ang_nest(nside, pix, lonlat=True):
    """Pixel centres of NEST pixels: (theta, phi) in radians, or (lon, lat)
    in degrees with ``lonlat`` (lon in [0, 360))."""
    nside = int(nside)
    npix = 12 * nside * nside
    ix, iy, face = nest2xyf(nside, pix)
    fact2 = 4.0 / npix
    fact1 = (nside << 1) * fact2
    jr = (_JRLL[face] << _order(nside)) - ix - iy - 1
    nr = np.where(jr < nside, jr, np.where(jr > 3 * nside, 4 * nside - jr, nside))
    tmp2 = (nr * nr).astype(np.float64) * fact2
    z = np.where(jr < nside, 1.0 - tmp2,
                 np.where(jr > 3 * nside, tmp2 - 1.0, (2 * nside - jr) * fact1))
    have_sth = ((jr < nside) & (z > 0.99)) | ((jr > 3 * nside) & (z < -0.99))
    sth = np.sqrt(np.where(have_sth, tmp2 * (2.0 - tmp2), 0.0))
    t = _JPLL[face] * nr + ix - iy
    t = np.where(t < 0, t + 8 * nr, t)
    halfpi = 0.5 * np.pi
    phi = np.where(nr == nside, 0.75 * halfpi * t * fact1, (0.5 * halfpi * t) / np.maximum(nr, 1))
    theta = np.where(have_sth, np.arctan2(sth, z), np.arccos(np.clip(z, -1.0, 1.0)))
    if lonlat:
        return np.degrees(phi), 90.0 - np.degrees(theta)
    return theta, phi


def all_neighbours_nest(nside, pix):
    """The 8 neighbours of NEST pixels, rows SW, W, NW, N, NE, E, SE, S
    (shape (8, n)); -1 where a corner pixel has only 7 neighbours."""
    nside = int(nside)
    ix, iy, face = nest2xyf(nside, pix)
    out = np.empty((8, ix.size), dtype=np.int64)
    for i in range(8):
        x = ix + _NB_X[i]
        y = iy + _NB_Y[i]
        nb = np.full(ix.shape, 4, dtype=np.int64)
        lo_x, hi_x = x < 0, x >= nside
        x = np.where(lo_x, x + nside, np.where(hi_x, x - nside, x))
        nb = nb - lo_x + hi_x
        lo_y, hi_y = y < 0, y >= nside
        y = np.where(lo_y, y + nside, np.where(hi_y, y - nside, y))
        nb = nb - 3 * lo_y + 3 * hi_y
        f = _NB_FACE[nb, face]
        bits = _NB_SWAP[nb, face >> 2]
        x = np.where(bits & 1, nside - x - 1, x)
        y = np.where(bits & 2, nside - y - 1, y)
        sw = (bits & 4) != 0
        x, y = np.where(sw, y, x), np.where(sw, x, y)
        out[i] = np.where(f >= 0, xyf2nest(nside, x, y, np.maximum(f, 0)), -1)
    return out


def dual_mesh(nside, wet_mask=None, depth=None, z_levels=None, name="healpix",
              cell_size=0.0) -> Mesh:
    """Dual mesh of a NEST HEALPix grid (readers/healpix.py:93): nodes are the
    wet pixel centres, elements the pixel-corner cells (quads; the 8
    valence-3 corners give triangles stored as [a, b, c, -1])."""
    nside = int(nside)
    npix = 12 * nside * nside
    wet = np.ones(npix, dtype=bool) if wet_mask is None else np.asarray(wet_mask, dtype=bool)
    dep = None if depth is None else np.abs(np.asarray(depth, dtype=np.float64))
    wet_pix = np.flatnonzero(wet).astype(np.int64)
    wet_ext = np.concatenate([wet, np.zeros(1, dtype=bool)])
    nb = all_neighbours_nest(nside, wet_pix)
    n_sw, n_w, n_nw, n_n, n_ne, n_e, n_se, n_s = nb
    blk = wet_pix
    cand = np.concatenate([
        np.stack([blk, n_ne, n_n, n_nw], axis=1),
        np.stack([blk, n_se, n_e, n_ne], axis=1),
        np.stack([blk, n_sw, n_s, n_se], axis=1),
        np.stack([blk, n_nw, n_w, n_sw], axis=1),
    ], axis=0)
    k = np.repeat(np.arange(4, dtype=np.int64), blk.size)
    edge1, diag, edge2 = cand[:, 1], cand[:, 2], cand[:, 3]
    keep = wet_ext[edge1] & wet_ext[edge2] & (wet_ext[diag] | (diag < 0))
    members = np.where(cand >= 0, cand, np.iinfo(np.int64).max)
    keep &= cand[:, 0] == members.min(axis=1)
    rows = cand[keep]
    keys = (cand[keep, 0] << 2) | k[keep]
    rows = rows[np.argsort(keys, kind="stable")]
    tri = rows[:, 2] < 0
    rows[tri, 2] = rows[tri, 3]
    rows[tri, 3] = -1
    node_pix = np.unique(rows[rows >= 0])
    remap = np.full(npix, -1, dtype=np.int64)
    remap[node_pix] = np.arange(node_pix.size)
    elem_nodes = np.where(rows >= 0, remap[np.clip(rows, 0, npix - 1)], -1).astype(np.int32)
    lon, lat = pix2ang_nest(nside, node_pix, lonlat=True)
    lon = ((lon + 180.0) % 360.0) - 180.0
    node_depth = dep[node_pix] if dep is not None else np.zeros(node_pix.size)
    mesh = make_mesh(name, lon, lat, elem_nodes, depth=node_depth, z_levels=z_levels,
                     cell_size=cell_size,
                     metadata={"n_triangles": int(tri.sum()),
                               "n_quads": int(rows.shape[0] - tri.sum()),
                               "healpix_nside": nside})
    mesh.metadata["healpix_node_pix"] = node_pix
    return mesh
