"""Device context: HBM-resident mesh, snapshot slots and particle buffers.

A thin Python owner of one ``kelp_ctx`` (include/kelp_hip.h).  Particle state
lives in device tensors allocated through PyTorch-ROCm (plumbing only: every
operation on it is a HIP kernel of libkelp_hip.so); the mesh and the field
snapshots are uploaded once and stay in HBM.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from . import _native
from ._native import (KELP_DT_BF16, KELP_DT_F16S, KELP_DT_F32, KELP_DT_F64, MeshDesc, Particles, RecordRows,
                      StepArgs, check)
from .mesh import Mesh


def _ptr(a) -> Optional[int]:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()  # torch tensor


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("kelp_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


@dataclass
class DeviceParticles:
    """Structure-of-arrays particle state in device memory."""
    lon: object
    lat: object
    z: object
    status: object
    nn: object
    fr: object
    u: object
    v: object
    w: object
    pid: object
    ws: object
    alive_count: object = None
    ctx: object = None      # owning DeviceContext (node id translation)
    elem: object = None     # carried containing element (elem_search walk)

    @property
    def n(self) -> int:
        return int(self.lon.shape[0])

    def as_struct(self) -> Particles:
        return Particles(self.n, _ptr(self.lon), _ptr(self.lat), _ptr(self.z), _ptr(self.status),
                         _ptr(self.nn), _ptr(self.fr), _ptr(self.u), _ptr(self.v), _ptr(self.w),
                         _ptr(self.pid), _ptr(self.ws), _ptr(self.alive_count), _ptr(self.elem))

    def to_host(self) -> Dict[str, np.ndarray]:
        """Host copies of every array; ``nn`` in the mesh's node numbering.
        Waits for the owning context's stream first (the copies run on
        torch's stream, which does not order against it)."""
        if self.ctx is not None:
            self.ctx.synchronize()
        out = {k: getattr(self, k).cpu().numpy() for k in
               ("lon", "lat", "z", "status", "fr", "u", "v", "w", "pid", "ws")}
        nn = self.nn.clone()
        if self.ctx is not None:
            # the clone ran on torch's stream, the id map runs on the
            # context's: order them explicitly
            self.ctx.torch.cuda.current_stream().synchronize()
            check(self.ctx.lib.kelp_nodes_to_mesh_ids(self.ctx.h, nn.numel(), nn.data_ptr()),
                  "kelp_nodes_to_mesh_ids")
            self.ctx.synchronize()
        out["nn"] = nn.cpu().numpy()
        if self.elem is not None:
            el = self.elem.clone()
            if self.ctx is not None:
                self.ctx.torch.cuda.current_stream().synchronize()
                check(self.ctx.lib.kelp_elems_to_mesh_ids(self.ctx.h, el.numel(), el.data_ptr()),
                      "kelp_elems_to_mesh_ids")
                self.ctx.synchronize()
            out["elem"] = el.cpu().numpy()
        return out


RESUSPENSION_CODES = {"never": 0, "auto": 1, "stochastic": 2, "auto_or_stochastic": 3}
#: step argument slot of a sigma surface given by DeviceContext.sigma_surface
SURFACE_GIVEN = -2


def make_step_args(order, slot_a, dt, time_dir, slot_b=-1, slot_t=-1, alpha0=0.0,
                   alphas=(0.0, 0.0, 0.0, 0.0), exact_nearest=True, w_method=2, kick=None,
                   beach=None, step_i=0, real_id=0, periodic_lon=None,
                   periodic_lat=None, reuse_anchor=False, sigma=None, walk=False, rk45=None) -> StepArgs:
    a = StepArgs()
    a.order = int(order)
    if rk45 is not None:
        # order 5: dict(tol_m, h_min, inv_dtn); 0 or missing = the defaults
        a.rk45_tol_m = float(rk45.get("tol_m") or 0.0)
        a.rk45_h_min = float(rk45.get("h_min") or 0.0)
        a.rk45_inv_dtn = float(rk45.get("inv_dtn") or 0.0)
    a.slot_a, a.slot_b, a.slot_t = int(slot_a), int(slot_b), int(slot_t)
    a.alpha0 = float(alpha0)
    for k in range(4):
        a.alphas[k] = float(alphas[k]) if k < len(alphas) else 0.0
    a.dt = float(dt)
    a.time_dir = float(time_dir)
    a.exact_nearest = int(exact_nearest) if not isinstance(exact_nearest, bool) else int(exact_nearest)
    a.w_method = int(w_method)
    a.step_i = int(step_i)
    a.real_id = int(real_id)
    a.reuse_anchor = int(bool(reuse_anchor))
    a.walk = int(bool(walk))
    a.sig_start_a = a.sig_start_b = a.sig_start_t = -1
    a.sig_end_a = a.sig_end_b = a.sig_end_t = -1
    if sigma is not None:
        # surfaces as (slot_a, slot_b, slot_t, alpha0, alpha) nested blends
        sa_, sb_, st_, a0_, al_ = sigma["start"]
        a.sig_start_a, a.sig_start_b, a.sig_start_t = int(sa_), int(sb_), int(st_)
        a.sig_start_alpha0, a.sig_start_alpha = float(a0_), float(al_)
        if sigma.get("end") is not None:
            ea_, eb_, et_, e0_, el_ = sigma["end"]
            a.sig_end_a, a.sig_end_b, a.sig_end_t = int(ea_), int(eb_), int(et_)
            a.sig_end_alpha0, a.sig_end_alpha = float(e0_), float(el_)
        a.Dmin = float(sigma.get("Dmin", 0.5))
        a.z0b_min = float(sigma.get("z0b_min", 0.1))
    if periodic_lon is not None:
        a.periodic_lon, a.px0, a.px1 = 1, float(periodic_lon[0]), float(periodic_lon[1])
    if periodic_lat is not None:
        a.periodic_lat, a.py0, a.py1 = 1, float(periodic_lat[0]), float(periodic_lat[1])
    if kick:
        a.kick = 1
        a.kh_x = float(kick["kh_x"])
        a.kh_y = float(kick["kh_y"])
        a.seed = int(kick["seed"])
        if kick.get("az_const"):
            a.az, a.az_const = 1, float(kick["az_const"])
        if kick.get("az_field") is not None:
            # (slot_a, slot_b or -1, alpha0): the step payload's Az bracket
            fa, fb, fal = kick["az_field"]
            a.az, a.az_field = 1, 1
            a.az_slot_a, a.az_slot_b, a.az_alpha0 = int(fa), int(fb), float(fal)
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


class DeviceContext:
    """One GPU context: mesh + field slots + stream."""

    def __init__(self, device: int = 0):
        self.lib = _native.load()
        self.torch = _torch()
        h = ctypes.c_void_p()
        check(self.lib.kelp_ctx_create(int(device), ctypes.byref(h)), "kelp_ctx_create")
        self.h = h
        self.device = int(device)
        self.mesh: Optional[Mesh] = None
        self._slot_keys: Dict[int, object] = {}

    def close(self) -> None:
        if self.h:
            self.lib.kelp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- mesh ---------------------------------------------------------------
    def upload_mesh(self, mesh: Mesh) -> None:
        if mesh.gh_offsets is None:
            raise ValueError("mesh has no grid hash (call kelp_amd.mesh.build_grid_hash)")
        keep = dict(
            lon=np.ascontiguousarray(mesh.nodes_lon, np.float64),
            lat=np.ascontiguousarray(mesh.nodes_lat, np.float64),
            depth=np.ascontiguousarray(mesh.depth, np.float64),
            en=np.ascontiguousarray(mesh.elem_nodes, np.int32),
            n2e=np.ascontiguousarray(mesh.node_to_elems, np.int32),
            cnt=np.ascontiguousarray(mesh.node_to_elems_count, np.int32),
            off=np.ascontiguousarray(mesh.gh_offsets, np.int32),
            ids=np.ascontiguousarray(mesh.gh_node_ids, np.int32),
            z=np.ascontiguousarray(mesh.sigma, np.float64),
            coastal=(None if mesh.elem_is_coastal is None
                     else np.ascontiguousarray(mesh.elem_is_coastal, np.uint8)),
            clon=np.ascontiguousarray(mesh.elem_center_lon, np.float64),
            clat=np.ascontiguousarray(mesh.elem_center_lat, np.float64),
            nbr=(None if getattr(mesh, "elem_neighbors", None) is None
                 else np.ascontiguousarray(mesh.elem_neighbors, np.int32)),
        )
        d = MeshDesc(
            mesh.n_nodes, mesh.n_elements, mesh.max_v, keep["n2e"].shape[1],
            keep["lon"].ctypes.data, keep["lat"].ctypes.data, keep["depth"].ctypes.data,
            keep["en"].ctypes.data, keep["n2e"].ctypes.data, keep["cnt"].ctypes.data,
            keep["off"].ctypes.data, keep["ids"].ctypes.data,
            mesh.gh_cell_size, mesh.gh_lon_min, mesh.gh_lat_min,
            mesh.gh_nx, mesh.gh_ny, int(mesh.gh_wrap_lon), mesh.gh_far_rows,
            mesh.r_earth, int(keep["z"].size), keep["z"].ctypes.data if keep["z"].size else None,
            None if keep["coastal"] is None else keep["coastal"].ctypes.data,
            keep["clon"].ctypes.data, keep["clat"].ctypes.data,
            1 if mesh.vertical_coord == "sigma" else 0,
            None if keep["nbr"] is None else keep["nbr"].ctypes.data)
        check(self.lib.kelp_mesh_upload(self.h, ctypes.byref(d)), "kelp_mesh_upload")
        self.mesh = mesh
        self._slot_keys.clear()

    def node_xyz(self) -> np.ndarray:
        out = np.empty((self.mesh.n_nodes, 3), dtype=np.float64)
        check(self.lib.kelp_mesh_node_xyz(self.h, out.ctypes.data), "kelp_mesh_node_xyz")
        return out

    # -- fields -------------------------------------------------------------
    def upload_eta(self, slot: int, eta) -> None:
        """Attach the snapshot's surface elevation (N,) to ``slot``."""
        eta = np.ascontiguousarray(eta, np.float64)
        if eta.shape != (self.mesh.n_nodes,):
            raise ValueError(f"eta of shape {eta.shape} does not match the mesh's {self.mesh.n_nodes} nodes")
        check(self.lib.kelp_field_upload_eta(self.h, int(slot), eta.ctypes.data), "kelp_field_upload_eta")

    def sigma_surface(self, which: int, eta) -> None:
        """The surface a sigma step starts (``which`` 0) or ends (1) on,
        given as an (N,) array (kelp_sigma_surface_upload); ``None``
        forgets it.  A step selects it with the slot ``SURFACE_GIVEN``."""
        if eta is None:
            check(self.lib.kelp_sigma_surface_upload(self.h, int(which), None), "kelp_sigma_surface_upload")
            return
        eta = np.ascontiguousarray(eta, np.float64)
        if eta.shape != (self.mesh.n_nodes,):
            raise ValueError(f"sigma surface of shape {eta.shape} does not match the mesh's "
                             f"{self.mesh.n_nodes} nodes")
        check(self.lib.kelp_sigma_surface_upload(self.h, int(which), eta.ctypes.data),
              "kelp_sigma_surface_upload")

    def upload_az(self, slot: int, az) -> None:
        """Attach the snapshot's vertical diffusivity (K, N) at the
        interfaces (mesh node order) to ``slot`` (kelp_field_upload_az)."""
        az = np.ascontiguousarray(az, np.float64)
        if az.ndim != 2 or az.shape[1] != self.mesh.n_nodes:
            raise ValueError(f"Az of shape {az.shape} is not (K, {self.mesh.n_nodes})")
        check(self.lib.kelp_field_upload_az(self.h, int(slot), az.shape[0], az.ctypes.data),
              "kelp_field_upload_az")

    def upload_scalar(self, slot: int, var: int, values) -> None:
        """Node scalar ``var`` of the snapshot in ``slot`` for sampling:
        (N,), (K-1, N) or (K, N) doubles in mesh node order."""
        values = np.ascontiguousarray(values, np.float64)
        if values.shape[-1] != self.mesh.n_nodes:
            raise ValueError(f"scalar of shape {values.shape} does not match the mesh's nodes")
        n_planes = 1 if values.ndim == 1 else values.shape[0]
        check(self.lib.kelp_scalar_upload(self.h, int(slot), int(var), n_planes, values.ctypes.data),
              "kelp_scalar_upload")

    def sample(self, p: "DeviceParticles", var: int, slot_a: int, slot_b: int = -1,
               alpha: float = 0.0, method: str = "leastsq", out=None):
        """Sampled scalar ``var`` at every particle, particle order of ``p``:
        a host array, or written into the device tensor ``out`` (no wait)."""
        codes = {"nearest": 0, "leastsq": 1, "barycentric": 2, "idw": 3}
        t = self.torch
        dev_out = out if out is not None else t.empty((p.n,), dtype=t.float64, device=p.lon.device)
        s = p.as_struct()
        check(self.lib.kelp_sample(self.h, ctypes.byref(s), int(var), int(slot_a), int(slot_b),
                                   float(alpha), codes[method], dev_out.data_ptr()), "kelp_sample")
        if out is not None:
            return out
        self.synchronize()
        return dev_out.cpu().numpy()

    def sample_vars(self, p: "DeviceParticles", variables, slot_a: int, slot_b: int = -1,
                    alpha: float = 0.0, method: str = "leastsq", out=None):
        """Several scalars at once: a (len(variables), n) device tensor
        (``out`` if given), row k = variable variables[k]; no wait."""
        codes = {"nearest": 0, "leastsq": 1, "barycentric": 2, "idw": 3}
        t = self.torch
        vs = np.ascontiguousarray(variables, np.int32)
        if out is None:
            out = t.empty((vs.size, p.n), dtype=t.float64, device=p.lon.device)
        s = p.as_struct()
        check(self.lib.kelp_sample_vars(self.h, ctypes.byref(s), int(vs.size), vs.ctypes.data,
                                        int(slot_a), int(slot_b), float(alpha), codes[method],
                                        out.data_ptr()), "kelp_sample_vars")
        return out

    def upload_field(self, slot: int, u, v, w=None, key=None, uv_layout="node", storage=None) -> None:
        """Upload one snapshot into ``slot``; skipped when ``key`` matches the
        slot's current content.  ``uv_layout``: "node" (u, v per node) or
        "element" (u, v per element centre, interpolated by least squares).
        ``storage="bf16"``: the values are rounded to bfloat16 on the device;
        ``storage="f16s"``: to float16 with a power-of-two scale per level
        (both node layout, z-level meshes); else stored in their own dtype."""
        if key is not None and self._slot_keys.get(slot) is key:
            return
        u = np.asarray(u)
        dt = KELP_DT_F32 if u.dtype == np.float32 else KELP_DT_F64
        if storage == "bf16":
            dt = KELP_DT_BF16
        elif storage == "f16s":
            dt = KELP_DT_F16S
        elif storage not in (None, "native"):
            raise ValueError(f"unknown storage {storage!r}")
        npd = np.float64 if dt == KELP_DT_F64 else np.float32
        u = np.ascontiguousarray(u, npd)
        v = np.ascontiguousarray(v, npd)
        n_planes = 1 if u.ndim == 1 else u.shape[0]
        if uv_layout not in ("node", "element"):
            raise ValueError(f"unknown uv_layout {uv_layout!r}")
        m = self.mesh.n_nodes if uv_layout == "node" else self.mesh.n_elements
        if u.shape[-1] != m or v.shape != u.shape:
            raise ValueError(f"u/v of shape {u.shape}/{v.shape} do not match the mesh "
                             f"({m} {uv_layout} values per plane)")
        wp, nw = None, 0
        if w is not None:
            w = np.ascontiguousarray(w, npd)
            if w.shape[-1] != self.mesh.n_nodes:
                raise ValueError(f"w of shape {w.shape} does not match the mesh's {self.mesh.n_nodes} nodes")
            wp, nw = w.ctypes.data, (1 if w.ndim == 1 else w.shape[0])
        fn = self.lib.kelp_field_upload if uv_layout == "node" else self.lib.kelp_field_upload_elem
        check(fn(self.h, int(slot), dt, n_planes, u.ctypes.data, v.ctypes.data, wp, nw),
              "kelp_field_upload")
        self._slot_keys[slot] = key

    def pinned(self, a):
        """A page-locked host copy of ``a`` (torch tensor), the form in which
        :meth:`upload_field_async` copies run asynchronously."""
        t = self.torch
        return t.from_numpy(np.ascontiguousarray(a)).pin_memory()

    def upload_field_async(self, slot: int, u, v, w=None, eta=None, uv_layout="node", key=None,
                           storage=None) -> None:
        """Refill ``slot`` (already holding a snapshot of this shape) on the
        copy stream while steps run (kelp_field_upload_async).  Arrays:
        numpy or pinned torch CPU tensors (see :meth:`pinned`); they are kept
        alive until the next refill of the slot."""
        t = self.torch

        def host(a, dtype=None):
            if a is None:
                return None
            if isinstance(a, t.Tensor):
                return a
            return np.ascontiguousarray(a, dtype)

        u = host(u)
        f32 = (u.dtype == t.float32) if isinstance(u, t.Tensor) else (u.dtype == np.float32)
        f32 = f32 or storage in ("bf16", "f16s")
        npd = np.float32 if f32 else np.float64
        u, v, w = host(u, npd), host(v, npd), host(w, npd)
        e = host(eta, np.float64)
        n_planes = 1 if u.ndim == 1 else int(u.shape[0])
        nw = 0 if w is None else (1 if w.ndim == 1 else int(w.shape[0]))
        ptr = lambda a: None if a is None else (a.data_ptr() if isinstance(a, t.Tensor) else a.ctypes.data)
        keep = getattr(self, "_async_keep", None)
        if keep is None:
            keep = self._async_keep = {}
        if slot in keep:
            check(self.lib.kelp_field_upload_wait(self.h, int(slot)), "kelp_field_upload_wait")
        layout = 0 if uv_layout == "node" else 1
        code = (KELP_DT_BF16 if storage == "bf16" else KELP_DT_F16S if storage == "f16s"
                else KELP_DT_F32 if f32 else KELP_DT_F64)
        check(self.lib.kelp_field_upload_async(self.h, int(slot), code,
                                               layout, n_planes, ptr(u), ptr(v), ptr(w), nw, ptr(e)),
              "kelp_field_upload_async")
        keep[slot] = (u, v, w, e)
        self._slot_keys[slot] = key

    def reconstruct(self, dst: int, src, times, tq: float) -> None:
        """Slot ``dst`` <- the PCHIP reconstruction of slots ``src`` (4) at
        ``times`` evaluated at ``tq`` (kelp_field_reconstruct); no wait."""
        s = np.ascontiguousarray(src, np.int32)
        t = np.ascontiguousarray(times, np.float64)
        if s.size != 4 or t.size != 4:
            raise ValueError("reconstruct needs four slots and four times")
        check(self.lib.kelp_field_reconstruct(self.h, int(dst), s.ctypes.data, t.ctypes.data, float(tq)),
              "kelp_field_reconstruct")
        self._slot_keys[dst] = None

    # -- particles ----------------------------------------------------------
    def particles(self, lon, lat, z=None, status=None, nn=None, fr=None, pid=None,
                  ws=None, elem=None) -> DeviceParticles:
        """Device particle state; ``nn`` and ``elem`` (carried elements for
        elem_search walk, -1 = unknown) in the mesh's numbering."""
        t = self.torch
        dev = f"cuda:{self.device}"
        P = len(lon)

        def f64(a, fill=0.0):
            if a is None:
                return t.full((P,), fill, dtype=t.float64, device=dev)
            return t.as_tensor(np.ascontiguousarray(a, np.float64), device=dev)

        def i32(a, fill):
            if a is None:
                return t.full((P,), fill, dtype=t.int32, device=dev)
            return t.as_tensor(np.ascontiguousarray(a, np.int32), device=dev)

        fr_t = (t.zeros((P,), dtype=t.int8, device=dev) if fr is None
                else t.as_tensor(np.ascontiguousarray(fr, np.int8), device=dev))
        pid_t = (t.arange(P, dtype=t.int64, device=dev) if pid is None
                 else t.as_tensor(np.ascontiguousarray(pid, np.int64), device=dev))
        nn_t = i32(nn, -1)
        el_t = None if elem is None else i32(elem, -1)
        arrays = dict(lon=f64(lon), lat=f64(lat), z=f64(z), status=i32(status, 1), u=f64(None), v=f64(None),
                      w=f64(None), ws=i32(ws, 0))
        # every tensor above was filled on torch's stream; the id maps below
        # and every later call run on the context's stream
        t.cuda.current_stream().synchronize()
        if nn is not None:
            check(self.lib.kelp_nodes_from_mesh_ids(self.h, nn_t.numel(), nn_t.data_ptr()),
                  "kelp_nodes_from_mesh_ids")
            self.synchronize()
        if elem is not None:
            check(self.lib.kelp_elems_from_mesh_ids(self.h, el_t.numel(), el_t.data_ptr()),
                  "kelp_elems_from_mesh_ids")
            self.synchronize()
        return DeviceParticles(nn=nn_t, fr=fr_t, pid=pid_t, ctx=self, elem=el_t, **arrays)

    # -- operations ---------------------------------------------------------
    def step(self, p: DeviceParticles, order: int, slot_a: int, dt: float, time_dir: float,
             slot_b: int = -1, slot_t: int = -1, alpha0: float = 0.0, alphas=(0.0, 0.0, 0.0, 0.0),
             exact_nearest: bool = True, w_method: int = 2, kick=None, beach=None,
             step_i: int = 0, real_id: int = 0, periodic_lon=None, periodic_lat=None,
             reuse_anchor: bool = False, sigma=None, walk: bool = False, rk45=None) -> None:
        """One step in place.  ``kick``: dict(kh_x, kh_y, seed[, az_const])
        for the split uniform-Kh random walk (and the vertical Az kick);
        ``beach``: dict of the beaching options (criteria, depth_min, ...);
        ``sigma`` (sigma meshes): dict(start=(slot_a, slot_b, slot_t, alpha0,
        alpha), end=same or None, Dmin, z0b_min), the surfaces at the step's
        start and end as blends of the slots' eta.  ``order=5`` is rk45
        with ``rk45=dict(tol_m, h_min, inv_dtn)``."""
        a = make_step_args(order, slot_a, dt, time_dir, slot_b, slot_t, alpha0, alphas,
                           exact_nearest, w_method, kick, beach, step_i, real_id,
                           periodic_lon, periodic_lat, reuse_anchor, sigma, walk, rk45)
        s = p.as_struct()
        check(self.lib.kelp_step(self.h, ctypes.byref(s), ctypes.byref(a)), "kelp_step")

    def record_stage(self, p: DeviceParticles, prev_status, pid_offset: int, n_rows: int, time_dir: float,
                     rows: Dict[str, object], sampled=None) -> None:
        """One masked record of every particle into the device rows ``rows``
        (lon, lat, z, u, v, w or None, status, fr, samples or None: device
        tensors of n_rows entries, samples n_vars x n_rows), at row pid -
        pid_offset (kelp_record_stage); ``sampled``: the n_vars x n device
        rows of :meth:`sample_vars`.  No wait."""
        s = p.as_struct()
        ptr = lambda t: 0 if t is None else t.data_ptr()
        out = RecordRows(*(ptr(rows.get(k)) for k in ("lon", "lat", "z", "u", "v", "w", "status", "fr",
                                                        "samples")))
        n_vars = 0 if sampled is None else int(sampled.shape[0])
        check(self.lib.kelp_record_stage(self.h, ctypes.byref(s), prev_status.data_ptr(), int(pid_offset),
                                         int(n_rows), float(time_dir), n_vars, ptr(sampled), ctypes.byref(out)),
              "kelp_record_stage")

    def activate(self, p: DeviceParticles, release_step, pid_offset: int, step: int,
                 prev_status=None) -> None:
        """Release NOT_RELEASED particles whose release step has come
        (``release_step``: device int64 tensor indexed by pid - pid_offset,
        or None) and copy the resulting status into ``prev_status`` (device
        int32 tensor or None); no wait."""
        s = p.as_struct()
        rs = 0 if release_step is None else release_step.data_ptr()
        nr = 0 if release_step is None else int(release_step.numel())
        ps = 0 if prev_status is None else prev_status.data_ptr()
        check(self.lib.kelp_activate(self.h, ctypes.byref(s), rs, int(pid_offset), nr, int(step), ps),
              "kelp_activate")

    def sort(self, p: DeviceParticles) -> None:
        """Reorder the particle arrays by grid-hash cell (see kelp_sort_particles)."""
        s = p.as_struct()
        check(self.lib.kelp_sort_particles(self.h, ctypes.byref(s)), "kelp_sort_particles")

    def nearest(self, lon, lat, exact: bool = True) -> np.ndarray:
        t = self.torch
        dev = f"cuda:{self.device}"
        lo = t.as_tensor(np.ascontiguousarray(lon, np.float64), device=dev)
        la = t.as_tensor(np.ascontiguousarray(lat, np.float64), device=dev)
        out = t.empty(lo.shape, dtype=t.int32, device=dev)
        check(self.lib.kelp_nearest(self.h, lo.numel(), _ptr(lo), _ptr(la), int(exact), _ptr(out)),
              "kelp_nearest")
        self.synchronize()
        return out.cpu().numpy()

    def containing(self, lon, lat, nn) -> np.ndarray:
        t = self.torch
        dev = f"cuda:{self.device}"
        lo = t.as_tensor(np.ascontiguousarray(lon, np.float64), device=dev)
        la = t.as_tensor(np.ascontiguousarray(lat, np.float64), device=dev)
        n = t.as_tensor(np.ascontiguousarray(nn, np.int32), device=dev)
        out = t.empty(lo.shape, dtype=t.int32, device=dev)
        check(self.lib.kelp_containing(self.h, lo.numel(), _ptr(lo), _ptr(la), _ptr(n), _ptr(out)),
              "kelp_containing")
        self.synchronize()
        return out.cpu().numpy()

    def synchronize(self) -> None:
        check(self.lib.kelp_synchronize(self.h), "kelp_synchronize")

    def event_record(self, slot: int) -> None:
        check(self.lib.kelp_event_record(self.h, int(slot)), "kelp_event_record")

    def event_elapsed_ms(self, a: int, b: int) -> float:
        ms = ctypes.c_float()
        check(self.lib.kelp_event_elapsed(self.h, int(a), int(b), ctypes.byref(ms)),
              "kelp_event_elapsed")
        return float(ms.value)
