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
        return VelocityField(u=u, v=v, eta=eta, w=w, uv_layout=layout, scalars=scalars)
