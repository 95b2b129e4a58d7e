"""ctypes access to the CPU oracle (test infrastructure only).

Used by tests/, the smoke check and the benchmark's CPU-baseline leg as the
checker; never by the product path.  ``flavour`` selects the build:
``"glibc"`` (host libm, as the reference's Numba JIT) or ``"kmath"`` (the
device kernels' math, for bit-exact comparison with the GPU).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")


def build() -> None:
    """Compile both oracle flavours with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OrMesh(ctypes.Structure):
    _fields_ = [
        ("n_nodes", c_int32), ("n_elems", c_int32), ("max_v", c_int32), ("max_conn", c_int32),
        ("nodes_lon", c_void_p), ("nodes_lat", c_void_p), ("depth", c_void_p),
        ("elem_nodes", c_void_p), ("n2e", c_void_p), ("n2e_count", c_void_p),
        ("node_xyz", c_void_p), ("gh_offsets", c_void_p), ("gh_node_ids", c_void_p),
        ("gh_cell", c_double), ("gh_lon_min", c_double), ("gh_lat_min", c_double),
        ("gh_nx", c_int32), ("gh_ny", c_int32), ("gh_wrap", c_int32), ("gh_far_rows", c_int32),
        ("r_earth", c_double),
        ("n_levels", c_int32), ("z_levels", c_void_p),
    ]


class OrFields(ctypes.Structure):
    _fields_ = [
        ("u_a", c_void_p), ("v_a", c_void_p), ("u_b", c_void_p), ("v_b", c_void_p),
        ("u_t", c_void_p), ("v_t", c_void_p),
        ("w_a", c_void_p), ("w_b", c_void_p), ("w_t", c_void_p),
        ("dtype", c_int32), ("alpha0", c_double),
    ]


_LIBS = {}


def lib(flavour: str = "glibc") -> ctypes.CDLL:
    if flavour in _LIBS:
        return _LIBS[flavour]
    path = os.path.join(BUILD, f"liboracle_{flavour}.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    vp = c_void_p
    L.or_step2d.argtypes = [POINTER(OrMesh), POINTER(OrFields), c_int, vp, c_double, c_double,
                            c_int, c_int64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.or_step2d.restype = None
    L.or_nearest_batch.argtypes = [POINTER(OrMesh), c_int64, vp, vp, c_int, vp]
    L.or_contain_batch.argtypes = [POINTER(OrMesh), c_int64, vp, vp, vp, vp]
    L.or_node_xyz.argtypes = [c_int64, vp, vp, vp]
    L.or_math_probe.argtypes = [c_int, c_int64, vp, vp, vp]
    L.or_nearest_exact.argtypes = [POINTER(OrMesh), c_double, c_double]
    L.or_nearest_exact.restype = c_int
    L.or_find_containing_local.argtypes = [POINTER(OrMesh), c_double, c_double, c_int]
    L.or_find_containing_local.restype = c_int
    L.or_element_contains.argtypes = [POINTER(OrMesh), c_int, c_double, c_double]
    L.or_element_contains.restype = c_int
    _LIBS[flavour] = L
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class OracleMesh:
    """Holds the arrays of one mesh and its ``or_mesh_t`` view."""

    def __init__(self, mesh, flavour: str = "glibc"):
        self.flavour = flavour
        self.L = lib(flavour)
        self.mesh = mesh
        self._keep = k = dict(
            lon=_c(mesh.nodes_lon, np.float64), lat=_c(mesh.nodes_lat, np.float64),
            depth=_c(mesh.depth, np.float64), en=_c(mesh.elem_nodes, np.int32),
            n2e=_c(mesh.node_to_elems, np.int32), cnt=_c(mesh.node_to_elems_count, np.int32),
            off=_c(mesh.gh_offsets, np.int32), ids=_c(mesh.gh_node_ids, np.int32),
            z=_c(mesh.sigma if mesh.sigma.size else np.zeros(1), np.float64))
        xyz = np.empty((mesh.n_nodes, 3), dtype=np.float64)
        self.L.or_node_xyz(mesh.n_nodes, k["lon"].ctypes.data, k["lat"].ctypes.data, xyz.ctypes.data)
        k["xyz"] = xyz
        self.node_xyz = xyz
        self.s = OrMesh(
            mesh.n_nodes, mesh.n_elements, mesh.max_v, k["n2e"].shape[1],
            k["lon"].ctypes.data, k["lat"].ctypes.data, k["depth"].ctypes.data,
            k["en"].ctypes.data, k["n2e"].ctypes.data, k["cnt"].ctypes.data,
            xyz.ctypes.data, k["off"].ctypes.data, k["ids"].ctypes.data,
            mesh.gh_cell_size, mesh.gh_lon_min, mesh.gh_lat_min,
            mesh.gh_nx, mesh.gh_ny, int(mesh.gh_wrap_lon), mesh.gh_far_rows,
            mesh.r_earth, int(mesh.sigma.size), k["z"].ctypes.data)

    def nearest(self, lon, lat, exact=True) -> np.ndarray:
        lon = _c(lon, np.float64)
        lat = _c(lat, np.float64)
        out = np.empty(lon.shape, dtype=np.int32)
        self.L.or_nearest_batch(ctypes.byref(self.s), lon.size, lon.ctypes.data, lat.ctypes.data,
                                int(exact), out.ctypes.data)
        return out

    def containing(self, lon, lat, nn) -> np.ndarray:
        lon = _c(lon, np.float64)
        lat = _c(lat, np.float64)
        nn = _c(nn, np.int32)
        out = np.empty(lon.shape, dtype=np.int32)
        self.L.or_contain_batch(ctypes.byref(self.s), lon.size, lon.ctypes.data, lat.ctypes.data,
                                nn.ctypes.data, out.ctypes.data)
        return out

    def element_contains(self, e, lon, lat) -> bool:
        return bool(self.L.or_element_contains(ctypes.byref(self.s), int(e), float(lon), float(lat)))

    def step2d(self, state: dict, order, u_a, v_a, dt, time_dir, u_b=None, v_b=None, u_t=None,
               v_t=None, alpha0=0.0, alphas=(0.0, 0.0, 0.0, 0.0), exact=True) -> dict:
        """One fused 2-D step (returns new arrays; ``state`` holds lon, lat,
        status, nn)."""
        dt_code = 1 if np.asarray(u_a).dtype == np.float32 else 0
        npd = np.float32 if dt_code else np.float64
        arrs = [_c(x, npd) if x is not None else None for x in (u_a, v_a, u_b, v_b, u_t, v_t)]
        ua, va, ub, vb, ut, vt = arrs
        ub = ua if ub is None else ub
        vb = va if vb is None else vb
        ut = ua if ut is None else ut
        vt = va if vt is None else vt
        f = OrFields(ua.ctypes.data, va.ctypes.data, ub.ctypes.data, vb.ctypes.data,
                     ut.ctypes.data, vt.ctypes.data, None, None, None, dt_code, float(alpha0))
        al = np.zeros(4, dtype=np.float64)
        al[:len(alphas)] = alphas
        lon = _c(state["lon"], np.float64)
        lat = _c(state["lat"], np.float64)
        st = _c(state["status"], np.int32)
        nn = _c(state["nn"], np.int32)
        P = lon.size
        out = dict(lon=np.empty(P), lat=np.empty(P), status=np.empty(P, np.int32),
                   nn=np.empty(P, np.int32), u=np.empty(P), v=np.empty(P), fr=np.empty(P, np.int8))
        self.L.or_step2d(ctypes.byref(self.s), ctypes.byref(f), int(order), al.ctypes.data,
                         float(dt), float(time_dir), int(exact), P,
                         lon.ctypes.data, lat.ctypes.data, st.ctypes.data, nn.ctypes.data,
                         out["lon"].ctypes.data, out["lat"].ctypes.data, out["status"].ctypes.data,
                         out["nn"].ctypes.data, out["u"].ctypes.data, out["v"].ctypes.data,
                         out["fr"].ctypes.data)
        return out


def math_probe(fn: int, x, y=None, flavour="glibc") -> np.ndarray:
    x = _c(x, np.float64)
    y = x if y is None else _c(y, np.float64)
    out = np.empty_like(x)
    lib(flavour).or_math_probe(int(fn), x.size, x.ctypes.data, y.ctypes.data, out.ctypes.data)
    return out
