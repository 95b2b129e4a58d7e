"""ctypes access to the CPU oracle (test infrastructure only).

Used by tests/, the smoke check and the benchmark's CPU-baseline leg as the
checker; never by the product path.  ``flavour`` selects the build:
``"glibc"`` (host libm, as the reference's Numba JIT), ``"kmath"`` (the
device kernels' math, for bit-exact comparison with the GPU) or
``"glibc_omp"`` (the glibc build with OpenMP over particles, whose thread
count :func:`set_threads` sets).
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
        ("n_levels", c_int32), ("z_levels", c_void_p), ("elem_coastal", c_void_p),
        ("elem_center_lon", c_void_p), ("elem_center_lat", c_void_p),
        ("vertical_coord", c_int32), ("elem_neighbors", c_void_p),
    ]


class OrStepArgs(ctypes.Structure):
    _fields_ = [
        ("order", c_int32), ("is3d", c_int32), ("w_method", c_int32), ("exact", c_int32),
        ("dt", c_double), ("time_dir", c_double), ("alphas", c_double * 4),
        ("kick", c_int32), ("beach", c_int32), ("kh_x", c_double), ("kh_y", c_double),
        ("seed", ctypes.c_uint64), ("real_id", ctypes.c_uint64), ("step_i", ctypes.c_uint64),
        ("use_depth", c_int32), ("use_coastal", c_int32), ("use_stochastic", c_int32),
        ("on_failure", c_int32), ("resuspension", c_int32), ("min_wet_steps", c_int32),
        ("depth_min", c_double), ("zone_depth", c_double), ("t_beach", c_double),
        ("rewet_hysteresis", c_double), ("t_resuspend", c_double),
        ("beach_seed", ctypes.c_uint64), ("beach_dt", c_double),
        ("az", c_int32), ("az_const", c_double), ("az_field", c_void_p), ("use_eta", c_int32),
        ("periodic_lon", c_int32), ("periodic_lat", c_int32),
        ("px0", c_double), ("px1", c_double), ("py0", c_double), ("py1", c_double),
        ("sig_eta", c_void_p * 5), ("Dmin", c_double), ("z0b_min", c_double),
        ("walk", c_int32),
        ("rk45_tol_m", c_double), ("rk45_h_min", c_double), ("rk45_inv_dtn", c_double),
        ("rk45_max_sub", ctypes.c_int64),
    ]


RESUSPENSION_CODES = {"never": 0, "auto": 1, "stochastic": 2, "auto_or_stochastic": 3}


def step_args(order, dt, time_dir, is3d=False, w_method=2, exact=True, alphas=(0, 0, 0, 0),
              kick=None, beach=None, step_i=0, real_id=0, periodic_lon=None,
              periodic_lat=None, sigma=None, walk=False, rk45=None) -> OrStepArgs:
    """``sigma`` (sigma meshes): dict(etas=[stage 0..3 surfaces, end
    surface] as (N,) arrays, Dmin, z0b_min); the arrays are kept alive by
    the returned structure."""
    a = OrStepArgs()
    a.walk = int(bool(walk))
    if int(order) == 5:
        # integrators.py:6524-6533: tolerance 1 m, h_min dt 1e-4, the cap
        r = dict(rk45 or {})
        a.rk45_tol_m = float(r.get("tol_m", 1.0))
        h_min = r.get("h_min")
        a.rk45_h_min = float(h_min) if h_min is not None else float(dt) * 1e-4
        a.rk45_max_sub = max(10_000, 2 * int(float(dt) / a.rk45_h_min) + 16)
        a.rk45_inv_dtn = float(r.get("inv_dtn", 0.0))
    if sigma is not None:
        etas = [np.ascontiguousarray(e, dtype=np.float64) for e in sigma["etas"]]
        a._sig_keep = etas
        for k in range(5):
            a.sig_eta[k] = etas[k].ctypes.data
        a.Dmin = float(sigma.get("Dmin", 0.5))
        a.z0b_min = float(sigma.get("z0b_min", 0.1))
    a.order, a.is3d, a.w_method, a.exact = int(order), int(is3d), int(w_method), int(exact)
    a.dt, a.time_dir = float(dt), float(time_dir)
    for k in range(4):
        a.alphas[k] = float(alphas[k]) if k < len(alphas) else 0.0
    a.step_i, a.real_id = int(step_i), int(real_id)
    if periodic_lon is not None:
        a.periodic_lon, a.px0, a.px1 = 1, float(periodic_lon[0]), float(periodic_lon[1])
    if periodic_lat is not None:
        a.periodic_lat, a.py0, a.py1 = 1, float(periodic_lat[0]), float(periodic_lat[1])
    if kick:
        a.kick = 1
        a.kh_x, a.kh_y, a.seed = float(kick["kh_x"]), float(kick["kh_y"]), int(kick["seed"])
        if kick.get("az_const"):
            a.az, a.az_const = 1, float(kick["az_const"])
        if kick.get("az_field") is not None:
            # the step payload's Az (K, N) at the interfaces, mesh order
            azf = np.ascontiguousarray(kick["az_field"], dtype=np.float64)
            a._az_keep = azf
            a.az, a.az_field = 1, azf.ctypes.data
    if beach:
        a.beach = 1
        crit = set(beach["criteria"])
        a.use_depth = int("depth" in crit)
        a.use_coastal = int("coastal_element" in crit)
        a.use_stochastic = int("stochastic" in crit)
        a.on_failure = int("on_failure" in crit)
        a.resuspension = RESUSPENSION_CODES[beach.get("resuspension", "never")]
        a.min_wet_steps = int(beach.get("min_wet_steps", 1))
        a.depth_min = float(beach.get("depth_min", 0.3))
        zone = beach.get("zone_depth")
        a.zone_depth = float(2.0 * a.depth_min if zone is None else zone)
        a.t_beach = float(beach.get("t_beach") or 0.0)
        a.rewet_hysteresis = float(beach.get("rewet_hysteresis", 0.1))
        a.t_resuspend = float(beach.get("t_resuspend") or 0.0)
        a.beach_dt = float(beach.get("dt", dt))
        a.beach_seed = int(beach.get("seed", a.seed))
        a.use_eta = int(bool(beach.get("use_eta", False)))
    return a


class OrFields(ctypes.Structure):
    _fields_ = [
        ("u_a", c_void_p), ("v_a", c_void_p), ("u_b", c_void_p), ("v_b", c_void_p),
        ("u_t", c_void_p), ("v_t", c_void_p),
        ("w_a", c_void_p), ("w_b", c_void_p), ("w_t", c_void_p),
        ("dtype", c_int32), ("alpha0", c_double), ("uv_layout", c_int32),
        ("eta_a", c_void_p), ("eta_b", c_void_p),
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
    L.or_gh_sample_stats.argtypes = [POINTER(OrMesh), c_int64, vp, vp, vp, vp, vp, vp, vp, vp]
    L.or_gh_sample_stats.restype = None
    L.or_set_threads.argtypes = [c_int]
    L.or_set_threads.restype = c_int
    L.or_step2d.argtypes = [POINTER(OrMesh), POINTER(OrFields), c_int, vp, c_double, c_double,
                            c_int, c_int64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.or_step2d.restype = None
    L.or_step.argtypes = [POINTER(OrMesh), POINTER(OrFields), POINTER(OrStepArgs), c_int64,
                          vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.or_step.restype = None
    L.or_hash5.argtypes = [ctypes.c_uint64] * 5
    L.or_hash5.restype = ctypes.c_uint64
    L.or_u01.argtypes = [ctypes.c_uint64] * 5
    L.or_u01.restype = c_double
    L.or_normal_pair.argtypes = [ctypes.c_uint64] * 5 + [POINTER(c_double), POINTER(c_double)]
    L.or_normal_pair.restype = None
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
    L.or_interp_bary_tinterp.argtypes = [POINTER(OrMesh), POINTER(OrFields), c_double, c_double,
                                         c_int, c_double, c_int, c_double,
                                         POINTER(c_double), POINTER(c_double), POINTER(c_int)]
    L.or_interp_bary_tinterp.restype = c_int
    L.or_sample.argtypes = [POINTER(OrMesh), vp, vp, c_double, c_int, c_int, c_int64, vp, vp, vp,
                            vp, vp, vp]
    L.or_sample.restype = None
    L.or_column_height_batch.argtypes = [POINTER(OrMesh), vp, c_double, c_int, c_int64, vp, vp, vp,
                                         vp, vp]
    L.or_column_height_batch.restype = None
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
            z=_c(mesh.sigma if mesh.sigma.size else np.zeros(1), np.float64),
            coastal=(None if getattr(mesh, "elem_is_coastal", None) is None
                     else _c(mesh.elem_is_coastal, np.uint8)),
            clon=_c(mesh.elem_center_lon, np.float64), clat=_c(mesh.elem_center_lat, np.float64),
            nbr=(None if getattr(mesh, "elem_neighbors", None) is None
                 else _c(mesh.elem_neighbors, np.int32)))
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
            mesh.r_earth, int(mesh.sigma.size), k["z"].ctypes.data,
            None if k["coastal"] is None else k["coastal"].ctypes.data,
            k["clon"].ctypes.data, k["clat"].ctypes.data,
            1 if getattr(mesh, "vertical_coord", "zlevel") == "sigma" else 0,
            None if k["nbr"] is None else k["nbr"].ctypes.data)

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

    def interp2d(self, lon, lat, nn, u, v):
        """Barycentric/Wachspress interpolation of node values (u, v) at
        each point around its anchor ``nn`` (interpolate.py:1341, steady,
        time_dir +1).  Returns (U, V, ok, elem)."""
        u = _c(u, np.float64)
        v = _c(v, np.float64)
        f = OrFields(u.ctypes.data, v.ctypes.data, u.ctypes.data, v.ctypes.data,
                     u.ctypes.data, v.ctypes.data, None, None, None, 0, 0.0)
        n = len(lon)
        U, V = np.empty(n), np.empty(n)
        ok, el = np.empty(n, bool), np.empty(n, np.int32)
        for i in range(n):
            a, b, e = c_double(), c_double(), c_int()
            ok[i] = bool(self.L.or_interp_bary_tinterp(
                ctypes.byref(self.s), ctypes.byref(f), float(lon[i]), float(lat[i]), int(nn[i]),
                1.0, -1, 0.0, ctypes.byref(a), ctypes.byref(b), ctypes.byref(e)))
            U[i], V[i], el[i] = a.value, b.value, e.value
        return U, V, ok, el

    def sample(self, values_a, lon, lat, z, status, nn, values_b=None, alpha=0.0,
               method="leastsq"):
        """numerics/sampling.py:727 sample_field of a node scalar ((N,),
        (K-1, N) or (K, N)), blended a -> b by ``alpha``, at each particle."""
        codes = {"nearest": 0, "leastsq": 1, "barycentric": 2, "idw": 3}
        fa = _c(values_a, np.float64)
        fb = fa if values_b is None else _c(values_b, np.float64)
        n_planes = 1 if fa.ndim == 1 else fa.shape[0]
        lon, lat = _c(lon, np.float64), _c(lat, np.float64)
        z = _c(np.zeros(lon.size) if z is None else z, np.float64)
        st, nn_ = _c(status, np.int32), _c(nn, np.int32)
        out = np.empty(lon.size)
        self.L.or_sample(ctypes.byref(self.s), fa.ctypes.data, fb.ctypes.data, float(alpha),
                         int(n_planes), codes[method], lon.size, lon.ctypes.data, lat.ctypes.data,
                         z.ctypes.data, st.ctypes.data, nn_.ctypes.data, out.ctypes.data)
        return out

    def column_height(self, lon, lat, nn, eta, Dmin=0.5, method="idw", alive=None):
        """vertical_sigma.py:1063 sigma_column_height: max(depth + eta,
        Dmin) interpolated at each particle by ``method`` around anchor
        ``nn`` (mesh ids); 1.0 where ``alive`` is False."""
        codes = {"nearest": 0, "leastsq": 1, "barycentric": 2, "idw": 3}
        lon, lat = _c(lon, np.float64), _c(lat, np.float64)
        nn_ = _c(nn, np.int32)
        al = _c(np.ones(lon.size) if alive is None else alive, np.int32)
        e = _c(eta, np.float64)
        out = np.empty(lon.size)
        self.L.or_column_height_batch(ctypes.byref(self.s), e.ctypes.data, float(Dmin), codes[method],
                                      lon.size, lon.ctypes.data, lat.ctypes.data, nn_.ctypes.data,
                                      al.ctypes.data, out.ctypes.data)
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


    def step(self, state: dict, fields: dict, args: OrStepArgs) -> dict:
        """One full step (deterministic + kick + beaching) on copies of the
        ``state`` arrays (lon, lat, z, status, nn, fr, pid, ws).  ``fields``:
        u_a, v_a[, w_a] and optionally the b/t snapshots, alpha0 and
        uv_layout ("node", or "element" for u/v at element centres)."""
        ua = np.asarray(fields["u_a"])
        dt_code = 1 if ua.dtype == np.float32 else 0
        npd = np.float32 if dt_code else np.float64
        keep = {}
        for nm in ("u", "v", "w"):
            a = fields.get(nm + "_a")
            if a is None:
                continue
            keep[nm + "_a"] = _c(a, npd)
            for sfx in ("b", "t"):
                x = fields.get(f"{nm}_{sfx}")
                keep[f"{nm}_{sfx}"] = keep[nm + "_a"] if x is None else _c(x, npd)
        ptr = lambda k: keep[k].ctypes.data if k in keep else None
        for nm in ("eta_a", "eta_b"):
            if fields.get(nm) is not None:
                keep[nm] = _c(fields[nm], np.float64)
        f = OrFields(ptr("u_a"), ptr("v_a"), ptr("u_b"), ptr("v_b"), ptr("u_t"), ptr("v_t"),
                     ptr("w_a"), ptr("w_b"), ptr("w_t"), dt_code, float(fields.get("alpha0", 0.0)),
                     1 if fields.get("uv_layout", "node") == "element" else 0,
                     ptr("eta_a"), ptr("eta_b"))
        P = len(state["lon"])
        out = dict(lon=np.array(state["lon"], np.float64), lat=np.array(state["lat"], np.float64),
                   z=np.array(state.get("z", np.zeros(P)), np.float64),
                   status=np.array(state["status"], np.int32), nn=np.array(state["nn"], np.int32),
                   fr=np.array(state.get("fr", np.zeros(P)), np.int8),
                   u=np.zeros(P), v=np.zeros(P), w=np.zeros(P),
                   pid=np.array(state.get("pid", np.arange(P)), np.int64),
                   ws=np.array(state.get("ws", np.zeros(P)), np.int32))
        if state.get("elem") is not None:
            out["elem"] = np.array(state["elem"], np.int32)
        self.L.or_step(ctypes.byref(self.s), ctypes.byref(f), ctypes.byref(args), P,
                       *[out[k].ctypes.data for k in ("lon", "lat", "z", "status", "nn", "fr",
                                                      "u", "v", "w", "pid", "ws")],
                       out["elem"].ctypes.data if "elem" in out else None)
        return out

    def nearest_exact(self, lon, lat):
        return self.nearest(lon, lat, exact=True)

    def gh_sample_stats(self, qlon, qlat, nn, cum, cell, lon_min, lat_min, nx, ny, wrap):
        """core/mesh.py:386 _gh_sample_stats against a candidate grid:
        (occ, cost, unprov, approx) per sample."""
        g = OrMesh.from_buffer_copy(self.s)
        g.gh_cell, g.gh_lon_min, g.gh_lat_min = float(cell), float(lon_min), float(lat_min)
        g.gh_nx, g.gh_ny, g.gh_wrap = int(nx), int(ny), int(bool(wrap))
        qlon = _c(qlon, np.float64)
        qlat = _c(qlat, np.float64)
        nn = _c(nn, np.int32)
        cum = _c(cum, np.int32)
        n = qlon.size
        occ = np.empty(n, np.int64)
        cost = np.empty(n, np.float64)
        unprov = np.empty(n, np.int8)
        approx = np.empty(n, np.int8)
        self.L.or_gh_sample_stats(ctypes.byref(g), n, qlon.ctypes.data, qlat.ctypes.data, nn.ctypes.data,
                                  cum.ctypes.data, occ.ctypes.data, cost.ctypes.data, unprov.ctypes.data,
                                  approx.ctypes.data)
        return occ, cost, unprov, approx

    def walk_locate(self, lon, lat, start) -> np.ndarray:
        """core/mesh.py:2631 find_containing_elements_walk (elements only)."""
        lon = _c(lon, np.float64)
        lat = _c(lat, np.float64)
        start = _c(start, np.int32)
        self.L.or_walk_locate.restype = ctypes.c_int
        self.L.or_walk_locate.argtypes = [POINTER(OrMesh), c_double, c_double, ctypes.c_int]
        return np.array([self.L.or_walk_locate(ctypes.byref(self.s), float(x), float(y), int(e))
                         for x, y, e in zip(lon, lat, start)], np.int32)


def set_threads(n: int, flavour: str = "glibc_omp") -> int:
    """Threads of the threaded build's particle loops; returns the count in effect."""
    return int(lib(flavour).or_set_threads(int(n)))


def math_probe(fn: int, x, y=None, flavour="glibc") -> np.ndarray:
    x = _c(x, np.float64)
    y = x if y is None else _c(y, np.float64)
    out = np.empty_like(x)
    lib(flavour).or_math_probe(int(fn), x.size, x.ctypes.data, y.ctypes.data, out.ctypes.data)
    return out


def _blend(a, b, alpha):
    """core/io.py:68 _blend: (1 - alpha) a + alpha b."""
    return (1.0 - alpha) * np.asarray(a, dtype=np.float64) + alpha * np.asarray(b, dtype=np.float64)


def stage_field(a, b, t, alpha0, alpha):
    """A stage's field value (integrators.py:1187 _make_stage_vels over the
    bracket blend core/io.py:386): the bracket (a, b) blended by alpha0,
    then towards t by alpha with the 1e-12 short-circuits."""
    base = a if (b is None or alpha0 == 0.0) else _blend(a, b, alpha0)
    if alpha < 1e-12:
        return np.asarray(base, dtype=np.float64)
    if alpha > 1.0 - 1e-12:
        return np.asarray(t, dtype=np.float64)
    return _blend(base, t, alpha)


def sigma_stage_etas(order, steady, eta_start, eta_end, a=None, b=None, t=None, alpha0=0.0,
                     alphas=(0.0, 0.0, 0.0, 0.0)):
    """The surfaces of one sigma step (integrators.py:5426-5498 RK2,
    :5806-5816 RK4): stage 0 = the start surface; later stages the stage
    field's eta, or on a steady step integrators.py:1369 _eta_blend of the
    start and end surfaces at c; [stage 0..3, end]."""
    e0 = np.asarray(eta_start, dtype=np.float64)
    ee = e0 if eta_end is None else np.asarray(eta_end, dtype=np.float64)
    out = [e0, e0, e0, e0, ee]
    cs = {2: (0.5,), 4: (0.5, 0.5, 1.0)}.get(int(order), ())
    for k, c in enumerate(cs, start=1):
        if steady:
            out[k] = e0 if c <= 0.0 else (ee if c >= 1.0 else _blend(e0, ee, c))
        else:
            out[k] = stage_field(a, b, t, alpha0, alphas[k])
    return out


def pchip4(f0, f1, f2, f3, times, tq, flavour="glibc"):
    """temporal_reconstruct.py:184 _pchip_value_scalar over whole arrays (the
    C oracle's or_pchip4); float64 result of f1's shape."""
    L = lib(flavour)
    arrs = [np.ascontiguousarray(f, np.float64) for f in (f0, f1, f2, f3)]
    out = np.empty(arrs[1].shape, np.float64)
    L.or_pchip4.argtypes = [c_int64] + [ctypes.c_void_p] * 4 + [c_double] * 5 + [ctypes.c_void_p]
    L.or_pchip4.restype = None
    L.or_pchip4(out.size, *[a.ctypes.data for a in arrs], *[float(t) for t in times], float(tq),
                out.ctypes.data)
    return out


def bf16_round(a) -> np.ndarray:
    """float32 -> bfloat16 -> float32, round to nearest even (IEEE 754
    roundTiesToEven on the top 16 bits of the float32 pattern; a NaN stays a
    quiet NaN): the values a bfloat16 snapshot holds, as float32 arrays."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def f16s_scales(rows) -> np.ndarray:
    """Power-of-two scale of each row of float32 values (KELP_DT_F16S): the
    largest finite |value| m = f 2^q (f in [0.5, 1)) gives 2^(q - 15), an
    empty or all-zero row 1; exponents clamped to the float32 normal range."""
    out = np.ones(len(rows), dtype=np.float32)
    for k, r in enumerate(rows):
        a = np.abs(np.asarray(r, dtype=np.float32))
        a = a[np.isfinite(a)]
        m = np.float32(a.max()) if a.size else np.float32(0.0)
        if m > 0:
            _, q = np.frexp(m)
            out[k] = np.ldexp(np.float32(1.0), int(min(max(int(q) - 15, -126), 127)))
    return out


def f16s_round(u, v, w=None):
    """The values a float16-with-level-scale snapshot holds, as float32
    arrays: u and v of layer L share the scale of the larger of their row
    maxima, every W interface row has its own; value * (1 / scale) rounded to
    the nearest even half, read back as half * scale.  2-D: u, v of shape
    (N,)."""
    u = np.asarray(u, dtype=np.float32)
    v = np.asarray(v, dtype=np.float32)
    u2, v2 = np.atleast_2d(u), np.atleast_2d(v)
    su, sv = f16s_scales(list(u2)), f16s_scales(list(v2))
    suv = np.maximum(su, sv)

    def q(a, s):
        inv = (np.float32(1.0) / s).astype(np.float32)
        h = (a * inv[:, None]).astype(np.float16)
        return (h.astype(np.float32) * s[:, None]).astype(np.float32)

    uq, vq = q(u2, suv).reshape(u.shape), q(v2, suv).reshape(v.shape)
    if w is None:
        return uq, vq
    w = np.asarray(w, dtype=np.float32)
    return uq, vq, q(w, f16s_scales(list(w)))
