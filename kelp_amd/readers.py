"""Reader registry and in-memory datasets, mirroring src/kelp/readers/__init__.py.

The reference's plugin API for model output is a reader class registered by
name (``@register("fesom2")``, readers/__init__.py:46) with
``load_mesh(mesh_cfg)``, ``open_dataset(path)`` and
``read_velocities(ds, timestep, mesh, mode_cfg, ...) -> VelocityField``.
NetCDF/Zarr/xarray are not part of this build, so the registry ships a
dataset type held in memory (``ArrayDataset``: named arrays with dimension
names and a time axis, the shape of an xarray Dataset) and one reader over
it with FESOM2's variable conventions (``fesom2_arrays``): u/v per node
(``unod``/``vnod``, dims ``(time, nz1, nod2)``) or per element
(``u``/``v``, ``(time, nz1, elem)``), W on the interfaces
(``(time, nz, nod2)``, file convention positive up, flipped to positive
down as readers/fesom2.py:1402 does), surface elevation ``ssh``/``eta``/
``zos`` and node scalars for sampling.  Other readers plug in with
``register``.
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from .fields import VelocityField

_READERS: Dict[str, type] = {}


def register(name: str):
    """Decorator registering a reader class by model name."""
    def deco(cls):
        _READERS[name] = cls
        return cls
    return deco


def get_reader(name: str):
    if name not in _READERS:
        raise KeyError(f"No reader registered for '{name}'. Available: {', '.join(sorted(_READERS))}")
    return _READERS[name]


class ArrayDataset:
    """Named arrays with dimension names and a ``time`` coordinate (s)."""

    def __init__(self, variables: Dict[str, Tuple[Sequence[str], np.ndarray]],
                 time: Sequence[float], attrs: Optional[Dict[str, Dict[str, str]]] = None):
        self.vars = {}
        for name, (dims, arr) in variables.items():
            arr = np.asarray(arr)
            if arr.ndim != len(dims):
                raise ValueError(f"{name}: {arr.ndim}-d array for dims {tuple(dims)}")
            self.vars[name] = (tuple(dims), arr)
        self.time = np.asarray(time, dtype=np.float64)
        self.attrs = attrs or {}
        for name, (dims, arr) in self.vars.items():
            if dims and dims[0] == "time" and arr.shape[0] != self.time.size:
                raise ValueError(f"{name}: {arr.shape[0]} records for {self.time.size} times")

    @property
    def sizes(self) -> Dict[str, int]:
        out = {"time": int(self.time.size)}
        for dims, arr in self.vars.values():
            out.update(dict(zip(dims, arr.shape)))
        return out

    def __contains__(self, name) -> bool:
        return name in self.vars

    def dims(self, name):
        return self.vars[name][0]

    def record(self, name, t: int) -> np.ndarray:
        """``name`` at time index ``t`` (the array itself when it has no
        time axis)."""
        dims, arr = self.vars[name]
        return arr[t] if dims and dims[0] == "time" else arr


def extract_time_info(ds: ArrayDataset):
    """readers/__init__.py extract_time_info: (file_times, file_dt)."""
    t = ds.time
    if t.size == 0:
        return None, None
    dt = float(t[1] - t[0]) if t.size > 1 else None
    return t.copy(), dt


def _first(ds, names):
    for n in names:
        if n and n in ds:
            return n
    return None


@register("fesom2_arrays")
class Fesom2ArrayReader:
    """FESOM2 variable conventions over an :class:`ArrayDataset`
    (readers/fesom2.py:1402 read_velocities)."""

    @staticmethod
    def open_dataset(source):
        if isinstance(source, ArrayDataset):
            return source
        raise TypeError("fesom2_arrays reads ArrayDataset objects (no file formats in this build)")

    @staticmethod
    def _uv_names(ds, mode_cfg):
        vu = getattr(mode_cfg, "var_u", None)
        vv = getattr(mode_cfg, "var_v", None)
        if vu and vv:
            return vu, vv
        for pair in (("u", "v"), ("unod", "vnod"), ("ua", "va"), ("u_vel", "v_vel")):
            if pair[0] in ds and pair[1] in ds:
                return pair
        raise KeyError("no u/v variables (u/v, unod/vnod, ua/va, u_vel/v_vel) in the dataset")

    @classmethod
    def probe_uv_layout(cls, ds, mode_cfg, mesh) -> str:
        u_name, _ = cls._uv_names(ds, mode_cfg)
        n = ds.vars[u_name][1].shape[-1]
        if n == mesh.n_nodes and n != mesh.n_elements:
            return "node"
        if n == mesh.n_elements:
            return "element"
        raise ValueError(f"{u_name}: {n} values per record match neither the nodes nor the elements")

    @classmethod
    def read_velocities(cls, ds, timestep, mesh, mode_cfg=None, diffusion=None, sample_vars=None,
                        velocity_float32=False) -> VelocityField:
        u_name, v_name = cls._uv_names(ds, mode_cfg)
        layout = cls.probe_uv_layout(ds, mode_cfg, mesh)
        is3d = bool(getattr(mode_cfg, "Fdrift3D", False))
        vd = np.float32 if velocity_float32 else np.float64
        u = np.asarray(ds.record(u_name, timestep))
        v = np.asarray(ds.record(v_name, timestep))
        if not is3d:
            layer = getattr(mode_cfg, "sigma_layer", None) or 0
            if u.ndim == 2:
                u, v = u[layer], v[layer]
        u = np.ascontiguousarray(u, vd)
        v = np.ascontiguousarray(v, vd)
        eta = np.zeros(mesh.n_nodes)
        e_name = _first(ds, [getattr(mode_cfg, "var_eta", None), "ssh", "eta", "zos"])
        if e_name is not None:
            eta = np.asarray(ds.record(e_name, timestep), dtype=np.float64)
        w = None
        if is3d:
            w_name = _first(ds, [getattr(mode_cfg, "var_w", None)] if getattr(mode_cfg, "var_w", None)
                            else ["w", "wo", "wvel"])
            if w_name is not None:
                if "nz1" in ds.dims(w_name):
                    raise ValueError(f"vertical velocity {w_name!r} lives on the nz1 mid-layers")
                w = np.ascontiguousarray(ds.record(w_name, timestep), vd)
                pos = getattr(mode_cfg, "w_positive", "auto")
                if pos == "auto":
                    pos = ds.attrs.get(w_name, {}).get("positive", "up")
                if pos == "up":
                    w = np.negative(w)
                if w.shape != (mesh.sigma.size, mesh.n_nodes):
                    raise ValueError(f"vertical velocity {w_name!r}: expected ({mesh.sigma.size}, "
                                     f"{mesh.n_nodes}), got {w.shape}")
        scalars = None
        if sample_vars:
            scalars = {}
            for name in sample_vars:
                if name not in ds:
                    raise KeyError(f"sample variable {name!r} not in the dataset")
                scalars[name] = np.asarray(ds.record(name, timestep), dtype=np.float64)
        az = None
        var_az = (getattr(diffusion, "var_az", None)
                  if diffusion is not None and getattr(diffusion, "enabled", False)
                  and getattr(diffusion, "az_enabled", False) else None)
        if is3d and var_az:
            # readers/fesom2.py:1596-1636: interface-based (the W layout),
            # float64, NaN and infinities as 0
            if var_az not in ds:
                raise ValueError(f"diffusion.var_az={var_az!r} not found in the dataset")
            if "nz1" in ds.dims(var_az):
                raise ValueError(f"diffusion.var_az={var_az!r} lives on the mid-layers -- Az must be "
                                 "interface-based (dim nz, same layout as w)")
            az = np.asarray(ds.record(var_az, timestep), dtype=np.float64)
            if az.shape != (mesh.sigma.size, mesh.n_nodes):
                raise ValueError(f"diffusion.var_az={var_az!r}: expected ({mesh.sigma.size}, "
                                 f"{mesh.n_nodes}), got {az.shape}")
            az = np.ascontiguousarray(np.nan_to_num(az, nan=0.0, posinf=0.0, neginf=0.0))
        return VelocityField(u=u, v=v, eta=eta, w=w, uv_layout=layout, scalars=scalars, az=az)


def sanitize_connectivity(nv) -> np.ndarray:
    """readers/fesomc.py:52 _sanitize_connectivity: 1-based padded
    connectivity (a triangle repeats its first vertex) to 0-based with -1
    fill and no repeated vertex."""
    conn = np.asarray(nv, dtype=np.int64)
    conn = np.where(conn <= 0, -1, conn - 1)
    clean = -np.ones(conn.shape, dtype=np.int32)
    for i in range(conn.shape[0]):
        seen = []
        for node in conn[i]:
            node = int(node)
            if node < 0 or node in seen:
                continue
            seen.append(node)
        clean[i, :len(seen)] = seen
    return clean


def _layer_major(arr, dims, n_layers, n_items, name, vert=("sigma", "sigmam1", "nz", "nz1", "depth",
                                                            "level")):
    """readers/fesomc.py:166/:206: a (layer, item) array from either axis
    order, by dimension name first, else by shape."""
    arr = np.asarray(arr)
    if arr.ndim != 2:
        raise ValueError(f"{name} must be 2-D in 3-D mode, got shape {arr.shape}")
    if dims is not None and len(dims) == 2:
        if dims[0] in vert and dims[1] not in vert:
            out = arr
        elif dims[1] in vert and dims[0] not in vert:
            out = arr.T
        else:
            out = None
        if out is not None:
            if out.shape != (n_layers, n_items):
                raise ValueError(f"{name}: expected ({n_layers}, {n_items}), got {out.shape}")
            return np.ascontiguousarray(out)
    if arr.shape == (n_layers, n_items):
        return np.ascontiguousarray(arr)
    if arr.shape == (n_items, n_layers):
        return np.ascontiguousarray(arr.T)
    raise ValueError(f"{name} has shape {arr.shape}; expected ({n_layers}, {n_items}) or the transpose")


@register("fesomc_arrays")
class FesomCArrayReader:
    """FESOM-C variable conventions (readers/fesomc.py) over an
    :class:`ArrayDataset`: nodes ``lon``/``lat``/``depth``, connectivity
    ``nv`` (1-based, triangles repeating their first vertex), element
    centres ``lon_elem``/``lat_elem``, sigma interfaces ``sigma_lev`` (1 at
    the surface .. 0 at the bottom), u/v per element and layer, the
    dia-sigma flux ``w`` on the interfaces (file convention positive up,
    flipped to positive down), surface elevation ``eta``."""

    @staticmethod
    def open_dataset(source):
        if isinstance(source, ArrayDataset):
            return source
        raise TypeError("fesomc_arrays reads ArrayDataset objects (no file formats in this build)")

    @staticmethod
    def load_mesh(ds, r_earth=None, cell_size=0.0):
        """readers/fesomc.py:246 load_mesh: a sigma mesh with the grid hash
        built (Earth radius 6,400 km unless given, the FESOM-C value)."""
        from .mesh import Mesh, build_grid_hash, build_node_to_elems
        lon = np.asarray(ds.record("lon", 0), dtype=np.float64)
        lat = np.asarray(ds.record("lat", 0), dtype=np.float64)
        depth = (np.asarray(ds.record("depth", 0), dtype=np.float64) if "depth" in ds
                 else np.zeros_like(lon))
        conn = sanitize_connectivity(ds.record("nv", 0))
        if "lon_elem" in ds and "lat_elem" in ds:
            clon = np.asarray(ds.record("lon_elem", 0), dtype=np.float64)
            clat = np.asarray(ds.record("lat_elem", 0), dtype=np.float64)
        else:
            clon = np.empty(conn.shape[0])
            clat = np.empty(conn.shape[0])
            for i, row in enumerate(conn):
                v = row[row >= 0]
                lo = lon[v]
                if lo.max() - lo.min() <= 180.0:
                    clon[i] = lo.mean()
                else:
                    d = ((lo - lo[0] + 180.0) % 360.0) - 180.0
                    clon[i] = ((lo[0] + d.mean() + 180.0) % 360.0) - 180.0
                clat[i] = lat[v].mean()
        sigma = (np.asarray(ds.record("sigma_lev", 0), dtype=np.float64) if "sigma_lev" in ds
                 else np.array([0.0]))
        n2e, cnt = build_node_to_elems(lon.size, conn)
        mesh = Mesh(name="fesomc", nodes_lon=lon, nodes_lat=lat, depth=depth, elem_nodes=conn,
                    elem_center_lon=clon, elem_center_lat=clat, node_to_elems=n2e,
                    node_to_elems_count=cnt, sigma=sigma, vertical_coord="sigma",
                    r_earth=float(r_earth or 6_400_000.0),
                    metadata={"sigma_transform": "linear_fesomc"})
        build_grid_hash(mesh, cell_size)
        return mesh

    @staticmethod
    def _uv_names(ds, mode_cfg):
        vu = getattr(mode_cfg, "var_u", None)
        vv = getattr(mode_cfg, "var_v", None)
        return (vu or "u"), (vv or "v")

    @classmethod
    def probe_uv_layout(cls, ds, mode_cfg, mesh) -> str:
        u_name, _ = cls._uv_names(ds, mode_cfg)
        dims = ds.dims(u_name)
        if "nele" in dims or "elem" in dims:
            return "element"
        if "node" in dims or "nod2" in dims:
            return "node"
        return Fesom2ArrayReader.probe_uv_layout(ds, mode_cfg, mesh)

    @classmethod
    def read_velocities(cls, ds, timestep, mesh, mode_cfg=None, diffusion=None, sample_vars=None,
                        velocity_float32=False) -> VelocityField:
        """readers/fesomc.py:514 read_velocities (3-D)."""
        if not getattr(mode_cfg, "Fdrift3D", False):
            raise NotImplementedError("fesomc_arrays: 2-D reductions of sigma fields")
        vd = np.float32 if velocity_float32 else np.float64
        e_name = getattr(mode_cfg, "var_eta", None) or "eta"
        eta = (np.asarray(ds.record(e_name, timestep), dtype=np.float64) if e_name in ds
               else np.zeros(mesh.n_nodes))
        u_name, v_name = cls._uv_names(ds, mode_cfg)
        w_name = getattr(mode_cfg, "var_w", None) or "w"
        if w_name == "w_cart":
            raise ValueError("FESOM-C 3-D: 'w_cart' lacks the dia-sigma part; track with 'w'")
        layout = cls.probe_uv_layout(ds, mode_cfg, mesh)
        n_items = mesh.n_nodes if layout == "node" else mesh.n_elements
        K = mesh.sigma.size
        drop_time = lambda name: ds.dims(name)[1:] if ds.dims(name)[0] == "time" else ds.dims(name)
        u = _layer_major(ds.record(u_name, timestep), drop_time(u_name), K - 1, n_items, u_name)
        v = _layer_major(ds.record(v_name, timestep), drop_time(v_name), K - 1, n_items, v_name)
        w = _layer_major(ds.record(w_name, timestep), drop_time(w_name), K, mesh.n_nodes, w_name)
        pos = getattr(mode_cfg, "w_positive", "auto")
        if pos == "auto":
            pos = ds.attrs.get(w_name, {}).get("positive", "up")
        w = (-1.0 if pos == "up" else 1.0) * np.asarray(w, dtype=np.float64)
        scalars = None
        if sample_vars:
            scalars = {name: np.asarray(ds.record(name, timestep), dtype=np.float64)
                       for name in sample_vars}
        return VelocityField(u=u.astype(vd, copy=False), v=v.astype(vd, copy=False), eta=eta,
                             w=w.astype(vd, copy=False), uv_layout=layout, scalars=scalars)


_HP_U = ("avg_uoe", "uo", "uos", "u")
_HP_V = ("avg_von", "vo", "vos", "v")
_HP_W = ("avg_wo", "wo", "w", "wvel")
_HP_ETA = ("avg_zos", "zos", "ssh", "eta", "sossheig")
_HP_DEPTH = ("deptho", "bathymetry", "bathy_meter")


@register("healpix_arrays")
class HealpixArrayReader:
    """readers/healpix.py:1291 HealpixReader over an ``ArrayDataset`` of a
    HEALPix store: u/v ``(time, level, value)`` with land and below-bottom
    levels NaN (finite-prefix columns), W ON the interfaces (positive up,
    like the products), surface elevation per cell; the NEST ordering and
    the geographic frame (no ``mesh.rotation``).  Nodes are the wet pixel
    centres of the dual mesh (``healpix.dual_mesh``), data gathered at the
    node pixels."""

    @staticmethod
    def _name(ds, aliases, override=None):
        if override:
            if override not in ds:
                raise KeyError(f"healpix store: variable {override!r} not found")
            return override
        return _first(ds, aliases)

    @classmethod
    def load_mesh(cls, ds, levels=None, nside=None, order="nest", mode_cfg=None, name="healpix"):
        """readers/healpix.py:1295 load_mesh: the wet mask = finite u and v
        at the first record on the tracked level (the surface unless
        ``mode.sigma_layer``), the bathymetry from a depth variable (which
        also refines the mask) or else ``levels[min(finite count, K-1)]``
        per cell, then the dual mesh on those pixels."""
        from . import healpix
        crs = ds.attrs.get("crs", {})
        nside = int(nside if nside is not None else crs.get("healpix_nside"))
        order = crs.get("healpix_order", order)
        if order != "nest":
            raise NotImplementedError("healpix_arrays: RING-ordered stores")
        u_name = cls._name(ds, _HP_U, getattr(mode_cfg, "var_u", None))
        v_name = cls._name(ds, _HP_V, getattr(mode_cfg, "var_v", None))
        if u_name is None or v_name is None:
            raise KeyError("healpix store: no u/v variable found")
        u0, v0 = np.asarray(ds.record(u_name, 0), np.float64), np.asarray(ds.record(v_name, 0), np.float64)
        lev = getattr(mode_cfg, "sigma_layer", None) or 0
        u0s = u0[lev] if u0.ndim == 2 else u0
        v0s = v0[lev] if v0.ndim == 2 else v0
        npix = 12 * nside * nside
        if u0s.size != npix:
            raise ValueError(f"healpix store: {u_name!r} has {u0s.size} cells, expected {npix}")
        wet = np.isfinite(u0s) & np.isfinite(v0s)
        levels = None if levels is None else np.abs(np.asarray(levels, np.float64))
        if getattr(mode_cfg, "Fdrift3D", False) and levels is None:
            raise ValueError("healpix 3-D mode needs the vertical interface depths (mesh.levels)")
        dep = None
        d_name = _first(ds, _HP_DEPTH)
        if d_name is not None:
            dd = np.asarray(ds.vars[d_name][1], np.float64).reshape(-1)
            dep = np.abs(np.nan_to_num(dd, nan=0.0))
            wet &= dep > 0.0
        elif levels is not None and u0.ndim == 2:
            nfin = np.minimum(np.isfinite(u0).sum(axis=0), levels.size - 1)
            dep = levels[nfin]
        mesh = healpix.dual_mesh(nside, wet_mask=wet, depth=dep, z_levels=levels, name=name)
        mesh.metadata["healpix_npix"] = npix
        return mesh

    @staticmethod
    def probe_uv_layout(ds, mode_cfg=None, mesh=None) -> str:
        return "node"

    @classmethod
    def read_velocities(cls, ds, timestep, mesh, mode_cfg=None, diffusion=None, sample_vars=None,
                        velocity_float32=False) -> VelocityField:
        """readers/healpix.py:1872 read_velocities (dual geometry): K-1
        u/v layers and K W interfaces gathered at the node pixels (a layer
        variable's K-th plane must be empty there), W negated when the
        store is positive up, every non-finite velocity zeroed (ground
        below each column's finite prefix), eta NaN as 0."""
        if sample_vars:
            raise NotImplementedError("healpix_arrays: record-point sampling")
        idx = np.asarray(mesh.metadata["healpix_node_pix"], np.int64)
        is3d = bool(getattr(mode_cfg, "Fdrift3D", False))
        u_name = cls._name(ds, _HP_U, getattr(mode_cfg, "var_u", None))
        v_name = cls._name(ds, _HP_V, getattr(mode_cfg, "var_v", None))
        vd = np.float32 if velocity_float32 else np.float64
        u = np.asarray(ds.record(u_name, timestep), np.float64)
        v = np.asarray(ds.record(v_name, timestep), np.float64)
        w = None
        if is3d:
            K = int(mesh.sigma.size)
            w_name = cls._name(ds, _HP_W, getattr(mode_cfg, "var_w", None))
            if w_name is None:
                raise KeyError("healpix store: no vertical velocity")
            wr = np.asarray(ds.record(w_name, timestep), np.float64)
            if wr.shape[0] < K:
                raise ValueError(f"healpix reader: {w_name!r} has {wr.shape[0]} levels for {K} interfaces")
            w = wr[:K][:, idx]
            pos = getattr(mode_cfg, "w_positive", "auto")
            if pos not in ("up", "down"):
                pos = ds.attrs.get(w_name, {}).get("positive", "up")
            if pos == "up":
                w = np.negative(w)
            cols = []
            for name, a in ((u_name, u), (v_name, v)):
                if a.shape[0] > K - 1 and np.isfinite(a[K - 1][idx]).any():
                    raise ValueError(f"healpix reader: {name!r} carries data in the plane below the "
                                     "deepest interface at tracked nodes")
                c = a[:K - 1][:, idx]
                fin = np.isfinite(c)
                if (~fin[:-1] & fin[1:]).any():
                    raise ValueError(f"healpix reader: {name!r} has finite data below a non-finite level")
                cols.append(c)
            u, v = cols
            w = w.astype(vd)
        else:
            lev = getattr(mode_cfg, "sigma_layer", None) or 0
            u = (u[lev] if u.ndim == 2 else u)[idx]
            v = (v[lev] if v.ndim == 2 else v)[idx]
        surf = (u[0] if is3d else u)
        surf_v = (v[0] if is3d else v)
        bad = ~(np.isfinite(surf) & np.isfinite(surf_v))
        if bad.mean() > 0.05:
            raise ValueError(f"healpix reader: {bad.mean():.1%} of wet nodes carry non-finite u/v at "
                             f"timestep {timestep}")
        u = u.astype(vd)
        v = v.astype(vd)
        u[~np.isfinite(u)] = 0.0
        v[~np.isfinite(v)] = 0.0
        if w is not None:
            w[~np.isfinite(w)] = 0.0
        eta = np.zeros(mesh.n_nodes)
        e_name = cls._name(ds, _HP_ETA, getattr(mode_cfg, "var_eta", None))
        if e_name is not None:
            eta = np.nan_to_num(np.asarray(ds.record(e_name, timestep), np.float64).reshape(-1)[idx],
                                nan=0.0, posinf=0.0, neginf=0.0)
        return VelocityField(u=np.ascontiguousarray(u), v=np.ascontiguousarray(v), eta=eta,
                             w=None if w is None else np.ascontiguousarray(w), uv_layout="node")
