"""The grid-hash cell rule: choose the hash cell of a mesh by measuring a
ladder of candidate cells on sampled element centres.

Restates ``choose_grid_hash_cell`` / ``apply_grid_hash_cell_rule``
(src/kelp/core/mesh.py:464, :619), which ``run_tracking`` applies by default
(``mesh.grid_hash_cell_rule: occupancy``, src/kelp/tracking.py:3378-3386).
The ladder, the prefix sums of the node counts per candidate and the
choice run here on the host; the per-sample statistics of a candidate (3x3
occupancy, the exact search's scan-cost proxy, the widening and
approximate-anchor flags, core/mesh.py:386 ``_gh_sample_stats``) and the
exact nearest node of every sample run on the device
(``kelp_gh_sample_stats``, ``kelp_nearest``).
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from ._native import check
from .mesh import (GRID_HASH_MAX_CELLS, Mesh, _cell_index, _grid_geometry, _lon_branch, build_grid_hash,
                   exact_nearest_on, far_rows_for)

GRID_HASH_CELL_RULES = ("occupancy", "legacy")


class DeviceStats:
    """The exact nearest node and the per-sample statistics on the device
    context ``ctx`` (which holds the mesh with its current hash)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def nearest(self, qlon, qlat):
        return self.ctx.nearest(qlon, qlat, exact=True)

    def stats(self, qlon, qlat, nn, cum, cell, lon_min, lat_min, nx, ny, wrap):
        t = self.ctx.torch
        dev = f"cuda:{self.ctx.device}"
        return _sample_stats(self.ctx, t.as_tensor(qlon, device=dev), t.as_tensor(qlat, device=dev),
                             t.as_tensor(np.ascontiguousarray(nn, np.int32), device=dev), cum, cell, lon_min,
                             lat_min, nx, ny, wrap)


def _sample_stats(ctx, qlon_d, qlat_d, nn_d, cum: np.ndarray, cell, lon_min, lat_min, nx, ny, wrap):
    t = ctx.torch
    dev = qlon_d.device
    n = qlon_d.numel()
    cum_d = t.as_tensor(np.ascontiguousarray(cum, np.int32), device=dev)
    occ = t.empty(n, dtype=t.int64, device=dev)
    cost = t.empty(n, dtype=t.float64, device=dev)
    unprov = t.empty(n, dtype=t.int8, device=dev)
    approx = t.empty(n, dtype=t.int8, device=dev)
    t.cuda.current_stream().synchronize()
    check(ctx.lib.kelp_gh_sample_stats(ctx.h, n, qlon_d.data_ptr(), qlat_d.data_ptr(), nn_d.data_ptr(),
                                       cum_d.data_ptr(), float(cell), float(lon_min), float(lat_min), int(nx),
                                       int(ny), int(bool(wrap)), occ.data_ptr(), cost.data_ptr(),
                                       unprov.data_ptr(), approx.data_ptr()), "kelp_gh_sample_stats")
    ctx.synchronize()
    return occ.cpu().numpy(), cost.cpu().numpy(), unprov.cpu().numpy(), approx.cpu().numpy()


def choose_grid_hash_cell(mesh: Mesh, ctx=None, *, exact: Optional[bool] = None, sample: int = 200_000,
                          seed: int = 0, min_median: int = 8, approx_tol: float = 0.005,
                          max_far_rows: int = 64, steps: int = 12, min_gain: float = 2.75,
                          max_unprovable: float = 0.20, max_hash_mb: float = 128.0) -> Tuple[float, dict]:
    """core/mesh.py:464: candidates legacy / sqrt(2)^k, k = 0..steps, while
    the row budget stays within ``max_far_rows`` and the grid within the
    cell cap; exact mode (the mesh's exact_nearest flag): the eligible
    candidate (widening share <= ``max_unprovable``, hash <= ``max_hash_mb``)
    of least mean scan cost, kept only if it predicts a ``min_gain`` x gain
    over the legacy cell; approximate mode: the smallest cell with no empty
    window, median occupancy >= ``min_median`` and no more approximate
    anchors than the legacy cell + ``approx_tol``.  ``ctx``: a device
    context holding ``mesh`` (with its current, legacy hash), or any object
    with the ``nearest`` / ``stats`` methods of :class:`DeviceStats`.
    Returns ``(cell, report)``; the mesh is not changed."""
    src = ctx if hasattr(ctx, "stats") else DeviceStats(ctx)
    if mesh.gh_offsets is None:
        build_grid_hash(mesh)
    if exact is None:
        exact = exact_nearest_on(mesh)
    legacy = float(mesh.gh_cell_size)
    mode = "exact" if exact else "approx"

    def keep(reason, n=0):
        return legacy, {"rule": "occupancy", "mode": mode, "legacy": legacy, "chosen": legacy,
                        "candidates": [], "sample": int(n), "reason": reason}

    E = mesh.n_elements
    if E == 0:
        return keep("no elements to sample: the reader's automatic cell")
    if E > sample:
        pick = np.sort(np.random.default_rng(seed).choice(E, sample, replace=False))
    else:
        pick = np.arange(E)
    qlon = np.ascontiguousarray(mesh.elem_center_lon[pick], dtype=np.float64)
    qlat = np.ascontiguousarray(mesh.elem_center_lat[pick], dtype=np.float64)
    finite = np.isfinite(qlon) & np.isfinite(qlat)
    if not finite.all():
        qlon, qlat = qlon[finite], qlat[finite]
    if qlon.size == 0:
        return keep("no finite element centre to sample: the reader's automatic cell")
    # the exact nearest node of every sample, whatever the guard flag says
    nn = src.nearest(qlon, qlat)
    found = nn >= 0
    if not found.all():
        qlon, qlat, nn = qlon[found], qlat[found], nn[found]
    if qlon.size == 0:
        return keep("no sample has a nearest node: the reader's automatic cell", 0)
    lon_b, wrap = _lon_branch(np.asarray(mesh.nodes_lon, dtype=np.float64))
    lat = np.asarray(mesh.nodes_lat, dtype=np.float64)
    rows = []
    for k in range(steps + 1):
        cell = legacy / (np.sqrt(2.0) ** k)
        far_rows = far_rows_for(mesh, cell)
        if far_rows > max_far_rows and k > 0:
            break
        lon_min, lat_min, nx, ny = _grid_geometry(lon_b, lat, cell, wrap)
        if nx * ny > GRID_HASH_MAX_CELLS and k > 0:
            break
        idx = _cell_index(lon_b, lat, cell, lon_min, lat_min, nx, ny, wrap)
        counts = np.bincount(idx, minlength=nx * ny).reshape(ny, nx)
        cum = np.zeros((ny + 1, nx + 1), dtype=np.int32)
        np.cumsum(counts, axis=0, dtype=np.int32, out=cum[1:, 1:])
        np.cumsum(cum[1:, 1:], axis=1, dtype=np.int32, out=cum[1:, 1:])
        occ, cost, unprov, approx = src.stats(qlon, qlat, nn, cum, cell, lon_min, lat_min, nx, ny, wrap)
        unprov_share = float(np.mean(unprov))
        hash_mb = float(4.0 * (nx * ny + 1) / 1e6)
        rows.append({"cell": float(cell), "far_rows": int(far_rows), "n_cells": int(nx * ny),
                     "hash_mb": hash_mb, "median": float(np.median(occ)),
                     "p95": float(np.percentile(occ, 95)), "empty": float(np.mean(occ == 0)),
                     "unprovable": unprov_share, "approx": float(np.mean(approx)),
                     "mean_cost": float(np.mean(cost)),
                     "eligible": bool(k == 0 or (unprov_share <= max_unprovable and hash_mb <= max_hash_mb))})
    base = rows[0]
    if exact:
        eligible = [r for r in rows if r["eligible"]]
        best = min(eligible, key=lambda r: (r["mean_cost"], -r["cell"]))
        gain = base["mean_cost"] / best["mean_cost"] if best["mean_cost"] > 0 else 1.0
        if best is not base and gain < min_gain:
            reason = (f"keep the legacy cell: the cheapest eligible candidate {best['cell']:.4g} deg "
                      f"predicts only {gain:.2f}x (< {min_gain:g}x) on the scan cost")
            best = base
        else:
            reason = (f"smallest mean scan cost {best['mean_cost']:.1f} nodes/query (legacy "
                      f"{base['mean_cost']:.1f}, {gain:.2f}x)")
    else:
        ok = [r for r in rows if r["empty"] == 0.0 and r["median"] >= min_median
              and r["approx"] <= base["approx"] + approx_tol and (r is base or r["hash_mb"] <= max_hash_mb)]
        best = min(ok, key=lambda r: r["cell"]) if ok else base
        reason = ("smallest cell with no empty window, median >= %d and no added approximate anchor" % min_median
                  if ok else "no smaller cell keeps every window non-empty without approximate anchors")
    return float(best["cell"]), {"rule": "occupancy", "mode": mode, "legacy": legacy,
                                 "chosen": float(best["cell"]), "reason": reason, "candidates": rows,
                                 "sample": int(qlon.size)}


def apply_grid_hash_cell_rule(mesh: Mesh, ctx, rule: str = "occupancy", **kw) -> dict:
    """core/mesh.py:619: ``legacy`` keeps the reader's automatic cell;
    ``occupancy`` runs :func:`choose_grid_hash_cell` and pins a different
    choice in ``mesh.metadata['grid_hash_cell']`` (the hash is rebuilt; the
    caller uploads the mesh again).  An explicit pin is never overridden.
    The report is kept in ``mesh.metadata['grid_hash_report']``."""
    rule = str(rule).lower()
    if rule not in GRID_HASH_CELL_RULES:
        raise ValueError(f"grid_hash_cell_rule must be one of {GRID_HASH_CELL_RULES}, got {rule!r}")
    md = mesh.metadata
    if "grid_hash_cell" in md and md.get("grid_hash_cell_rule") in (None, "explicit"):
        md["grid_hash_cell_rule"] = "explicit"
        report = {"rule": "explicit", "chosen": float(mesh.gh_cell_size), "reason": "explicit cell"}
        md["grid_hash_report"] = report
        return report
    if rule == "legacy":
        report = {"rule": "legacy", "chosen": float(mesh.gh_cell_size), "legacy": float(mesh.gh_cell_size),
                  "reason": "legacy rule: the reader's automatic cell"}
    else:
        cell, report = choose_grid_hash_cell(mesh, ctx, **kw)
        if cell != mesh.gh_cell_size:
            md["grid_hash_cell"] = float(cell)
            build_grid_hash(mesh)
        md["grid_hash_cell_mode"] = report["mode"]
    md["grid_hash_cell_rule"] = rule
    md["grid_hash_report"] = report
    return report


__all__ = ["choose_grid_hash_cell", "apply_grid_hash_cell_rule", "GRID_HASH_CELL_RULES", "DeviceStats"]
