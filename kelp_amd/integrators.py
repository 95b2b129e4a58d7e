"""Integrator registry on the device path, mirroring
src/kelp/numerics/integrators.py:6635-6660.

``euler_step`` / ``rk2_step`` / ``rk4_step`` keep the reference signature
``step(state, vel, mesh, dt, time_dir, layer=0, **kwargs) -> ParticleState``
(integrators.py:5150, :5611, :6013) and its keyword arguments (``vel_next``,
``dt_to_next``, ``w_interp``, ``uv_interp``, ``diffusion``), and return a
NEW state (inputs untouched).  Each call runs one ``kelp_step`` launch: the
anchors, containment, interpolation, time blend, RK stages and the split
uniform-Kh kick in one HIP kernel.  The mesh and every snapshot the calls
name stay resident in HBM between calls (uploaded on first use, keyed by the
Python object); the particle state crosses PCIe both ways on each call —
``kelp_amd.tracking`` keeps it on the device for whole runs instead.

Anchor mode: with the mesh's exact-nearest guard on (``set_exact_nearest``,
always on under ``run_tracking``) every anchor is the exact nearest node;
with it off the reference's routes all anchor through the 3x3 hash window
with misses healed to the exact nearest (the unfused batch path, and the
fused paths wherever they are enabled, integrators.py:3289-3320), which is
anchor mode 2 of the kernel.

u/v at the nodes are interpolated barycentrically in the containing element
(``uv_interp="barycentric"``), u/v at element centres by least squares over
the anchor's element ring (``uv_interp="leastsq"``, the reference default);
W by ``nearest``, ``leastsq``, ``barycentric`` or ``idw``.
Sigma meshes take the reference's ``sigma_eta_start`` / ``sigma_eta_end``
/ ``Dmin`` / ``z0b_min`` keywords; a PCHIP ``recon`` bundle (Euler, RK2) is
reconstructed on the device; a state that carries ``elem`` runs the edge
walk (node-layout u/v, z-level).  The mesh may be this package's
:class:`~kelp_amd.mesh.Mesh` or the reference's own ``Mesh`` (adapted
without copies by :func:`~kelp_amd.mesh.from_reference_mesh`), and the
state, fields and diffusion arguments may be the reference's objects.
Combinations the device path does not implement raise
``NotImplementedError`` (edge_leastsq, EM/M1 diffusion, PCHIP with RK4);
there is no CPU fallback.
"""
from __future__ import annotations

import os
from collections import OrderedDict
from typing import Optional

import numpy as np

from .diffusion import DiffusionStepArgs
from .fields import VelocityField
from .mesh import Mesh, build_elem_neighbors, exact_nearest_on, from_reference_mesh
from .particles import ParticleState

_C_COEFFS = {1: (0.0,), 2: (0.0, 0.5), 4: (0.0, 0.5, 0.5, 1.0)}
W_METHODS = {"nearest": 0, "leastsq": 1, "barycentric": 2, "idw": 3}
SIGMA_W_METHODS = {"nearest": 0, "leastsq": 1, "barycentric": 2, "idw": 3}
#: library slots (kMaxSlots); the last two hold PCHIP stage fields
MAX_SLOTS = 16
RECON_SLOTS = (MAX_SLOTS - 2, MAX_SLOTS - 1)
N_SLOTS = RECON_SLOTS[0]


def _is_lazy(vel) -> bool:
    """A two-snapshot bracket (core/io.py:386 LazyLerpField, this package's
    or the reference's)."""
    return hasattr(vel, "vel_a") and hasattr(vel, "vel_b")


def _validate_time_aware_kwargs(vel_next, dt_to_next):
    """integrators.py:965: both or neither."""
    if (vel_next is None) != (dt_to_next is None):
        raise ValueError("vel_next and dt_to_next must be provided together "
                         f"(got vel_next={vel_next!r}, dt_to_next={dt_to_next!r})")


def _is_3d(vel) -> bool:
    """integrators.py:949: W present and u per layer."""
    base = vel.vel_a if _is_lazy(vel) else vel
    return base.w is not None and np.ndim(base.u) == 2


def stage_alphas(order: int, dt: float, dt_to_next: Optional[float]):
    """Per-stage fraction towards the target snapshot, ``c_k * dt /
    dt_to_next`` (integrators.py:1187 _make_stage_vels); the kernel applies
    the reference's endpoint short-circuits (alpha < 1e-12 -> the step-start
    field, alpha > 1 - 1e-12 -> the target)."""
    if order == 5:
        return [0.0] * 4          # rk45 takes t / dt_to_next per substep instead
    cs = _C_COEFFS[order]
    if dt_to_next is None:
        return [0.0] * 4
    if dt > dt_to_next * (1.0 + 1e-12):
        raise ValueError(f"dt={dt} steps past the snapshot bracket (dt_to_next={dt_to_next})")
    al = [c * dt / dt_to_next for c in cs]
    return al + [0.0] * (4 - len(al))


class Engine:
    """One device context shared by the registry functions: the resident
    mesh and an LRU of snapshot slots keyed by the field object."""

    def __init__(self, device: Optional[int] = None):
        from .device import DeviceContext
        if device is None:
            device = int(os.environ.get("KELP_DEVICE", "0"))
        self.ctx = DeviceContext(device)
        self._mesh = None
        self._slots: "OrderedDict[int, tuple]" = OrderedDict()   # id(field) -> (slot, field, dtype)

    def bind_mesh(self, mesh: Mesh) -> None:
        if self._mesh is not mesh:
            self.ctx.upload_mesh(mesh)
            self._mesh = mesh
            self._slots.clear()

    def rebind_mesh(self, mesh: Mesh) -> None:
        """Upload ``mesh`` again (its search structures grew)."""
        self._mesh = None
        self.bind_mesh(mesh)

    def slot_for(self, vel: VelocityField, dtype) -> int:
        key = id(vel)
        hit = self._slots.get(key)
        if hit is not None and hit[1] is vel and hit[2] == dtype:
            self._slots.move_to_end(key)
            return hit[0]
        if hit is not None:
            slot = hit[0]
            del self._slots[key]
        elif len(self._slots) < N_SLOTS:
            used = {s for s, _, _ in self._slots.values()}
            slot = min(set(range(N_SLOTS)) - used)
        else:
            _, (slot, _, _) = self._slots.popitem(last=False)
        w = None if vel.w is None else np.asarray(vel.w, dtype)
        self.ctx.upload_field(slot, np.asarray(vel.u, dtype), np.asarray(vel.v, dtype), w,
                              uv_layout=vel.uv_layout)
        if getattr(vel, "az", None) is not None and w is not None:
            self.ctx.upload_az(slot, vel.az)
        if self._mesh is not None and self._mesh.vertical_coord == "sigma" and getattr(vel, "eta", None) is not None:
            self.ctx.upload_eta(slot, vel.eta)
        self._slots[key] = (slot, vel, dtype)
        return slot


_ENGINE: Optional[Engine] = None


def get_engine() -> Engine:
    global _ENGINE
    if _ENGINE is None:
        _ENGINE = Engine()
    return _ENGINE


def _field_dtype(*fields):
    """float32 when every snapshot is stored float32 (the kernel blends in
    f64 after promoting each value), else float64."""
    for f in fields:
        if f is not None and np.asarray(f.u).dtype != np.float32:
            return np.float64
    return np.float32


def _payload_az(vel):
    """The step payload's Az: a bracket's only when both snapshots carry
    one (core/io.py:475-492), else the snapshot's own."""
    if _is_lazy(vel):
        if getattr(vel.vel_a, "az", None) is None or getattr(vel.vel_b, "az", None) is None:
            return None
        return vel.vel_a.az
    return getattr(vel, "az", None)


def _check_supported(vel, mesh, uv_interp, w_interp, diffusion, is3d, walk):
    base = vel.vel_a if _is_lazy(vel) else vel
    layout = getattr(base, "uv_layout", "element")
    ok = ((uv_interp == "barycentric" and layout == "node")
          or (uv_interp == "leastsq" and layout == "element"))
    if not ok:
        raise NotImplementedError(
            f"device path: uv_interp={uv_interp!r} with uv_layout={layout!r} is not "
            "implemented (node-layout barycentric, element-layout leastsq)")
    if walk and (layout != "node" or (is3d and mesh.vertical_coord == "sigma")):
        raise NotImplementedError("device path: the edge walk (a state carrying elem) with "
                                  "element-layout u/v or on a sigma mesh")
    if is3d:
        if mesh.vertical_coord not in ("zlevel", "sigma"):
            raise NotImplementedError(f"device path: vertical_coord={mesh.vertical_coord!r}")
        methods = SIGMA_W_METHODS if mesh.vertical_coord == "sigma" else W_METHODS
        if w_interp not in methods:
            raise NotImplementedError(f"device path: w_interp={w_interp!r} is not implemented "
                                      f"({', '.join(methods)})")
        if getattr(base, "uv_vertical", "layer") != "layer":
            raise NotImplementedError(f"device path: uv_vertical={base.uv_vertical!r}")
    if diffusion is not None:
        if diffusion.scheme != "uniform":
            raise NotImplementedError(f"device path: diffusion scheme {diffusion.scheme!r}")
        if diffusion.composition not in ("split", "split_rk"):
            raise NotImplementedError(f"device path: composition {diffusion.composition!r}")
        if diffusion.az_enabled:
            has_field = _payload_az(vel) is not None
            if diffusion.az_requires_field and not has_field:
                # integrators.py:884-889
                raise ValueError("diffusion.var_az is configured but the velocity payload carries "
                                 "no az field")
            if not is3d and (has_field or diffusion.az_const > 0.0):
                # integrators.py:893
                raise ValueError("diffusion.az_enabled requires a 3-D run (Fdrift3D)")
            if has_field and mesh.vertical_coord == "sigma":
                # integrators.py:902-906
                raise ValueError("diffusion.var_az is not supported on sigma meshes (v1)")


def _recon_stage_slots(eng, recon, order, dt, dtype):
    """temporal_reconstruct.py:298 make_stage_vels_recon on the device: the
    four snapshots of ``recon = (fields, times, q, direction)`` resident,
    stage c's field (at q + direction c dt) reconstructed into the reserved
    slots; returns (slot_a, slot_t, alphas) for the step."""
    if order not in (1, 2):
        raise NotImplementedError(f"device path: pchip reconstruction with order {order} "
                                  "(euler and rk2)")
    rs, rt, rq, rd = recon
    if len(rs) != 4 or len(rt) != 4:
        raise ValueError("recon needs four snapshots and four times")
    src = [eng.slot_for(f, dtype) for f in rs]
    t4 = [float(t) for t in rt]
    cs = _C_COEFFS[order]
    for dst, c in zip(RECON_SLOTS, cs):
        eng.ctx.reconstruct(dst, src, t4, float(rq) + float(rd) * c * float(dt))
    if order == 1:
        return RECON_SLOTS[0], -1, [0.0, 0.0, 0.0, 0.0]
    return RECON_SLOTS[0], RECON_SLOTS[1], [0.0, 1.0, 0.0, 0.0]


def _sigma_args(eng, vel, kwargs, order, dt, dt_to_next, slot_a, slot_b, slot_t, alpha0, alphas, recon):
    """The surfaces of a sigma step (integrators.py:1381 _sigma_end_eta,
    :1423 _sigma_given): the given ``sigma_eta_start`` / ``sigma_eta_end``
    arrays, else the step field's own surface at the start and its blend
    towards ``vel_next`` (or the PCHIP stage at c = 1) at the end, else the
    start surface."""
    from .device import SURFACE_GIVEN
    ctx = eng.ctx
    n = eng._mesh.n_nodes
    given_s = kwargs.get("sigma_eta_start")
    given_e = kwargs.get("sigma_eta_end")
    for g in (given_s, given_e):
        if g is not None and np.shape(g) != (n,):
            raise ValueError(f"sigma step: supplied surface has shape {np.shape(g)}, the velocity "
                             f"field's eta ({n},)")
    if given_s is None or given_e is None:
        bases = [vel.vel_a, vel.vel_b] if _is_lazy(vel) else [vel]
        if any(getattr(b, "eta", None) is None for b in bases):
            raise ValueError("sigma step: the velocity field carries no eta and no surface was given")
    if given_s is not None:
        ctx.sigma_surface(0, given_s)
        start = (SURFACE_GIVEN, -1, -1, 0.0, 0.0)
    else:
        start = (slot_a, slot_b, slot_t, alpha0, 0.0)
    end = None
    if given_e is not None:
        ctx.sigma_surface(1, given_e)
        end = (SURFACE_GIVEN, -1, -1, 0.0, 0.0)
    elif recon is not None:
        end = (slot_a, slot_b, slot_t, 0.0, 1.0)
    elif slot_t >= 0 and dt_to_next is not None and dt_to_next > 0:
        end = (slot_a, slot_b, slot_t, alpha0, float(dt) / float(dt_to_next))
    elif given_s is not None:
        end = start
    return dict(start=start, end=end, Dmin=float(kwargs.get("Dmin", 0.5)),
                z0b_min=float(kwargs.get("z0b_min", 0.1)))


def _step(order: int, state: ParticleState, vel, mesh, dt: float, time_dir: float,
          kwargs) -> ParticleState:
    mesh = from_reference_mesh(mesh)
    recon = kwargs.get("recon")
    walk = getattr(state, "elem", None) is not None
    if kwargs.get("elem_search", "walk" if walk else "local") not in ("local", "walk"):
        raise ValueError(f"elem_search must be 'local' or 'walk', got {kwargs.get('elem_search')!r}")
    uv_interp = kwargs.get("uv_interp", "leastsq")
    w_interp = kwargs.get("w_interp", "leastsq")
    vel_next = kwargs.get("vel_next")
    dt_to_next = kwargs.get("dt_to_next")
    diffusion: Optional[DiffusionStepArgs] = kwargs.get("diffusion")
    _validate_time_aware_kwargs(vel_next, dt_to_next)
    is3d = _is_3d(vel)
    sig3d = is3d and mesh.vertical_coord == "sigma"
    rk45 = None
    if order == 5:
        # integrators.py:6485-6513: 2-D element-layout least squares only
        base = vel.vel_a if _is_lazy(vel) else vel
        if uv_interp != "leastsq":
            raise NotImplementedError(f"rk45 supports uv_interp='leastsq' only (v1), got {uv_interp!r}")
        if is3d:
            raise NotImplementedError("rk45 is 2-D only (v1)")
        if base.uv_layout != "element":
            raise ValueError("rk45 requires element-layout velocities (v1)")
        if walk:
            raise ValueError("rk45 does not support elem_search='walk' (v1)")
        if recon is not None:
            raise NotImplementedError("device path: rk45 with pchip reconstruction")
        tol = float(kwargs.get("rk45_tol_m", 1.0))
        h_min = kwargs.get("rk45_h_min")
        rk45 = dict(tol_m=tol, h_min=None if h_min is None else float(h_min),
                    inv_dtn=0.0 if dt_to_next is None else 1.0 / float(dt_to_next))
    _check_supported(vel, mesh, uv_interp, w_interp, diffusion, is3d, walk)
    if recon is not None and sig3d:
        raise NotImplementedError("device path: pchip reconstruction on a sigma mesh")
    if dt <= 0:
        raise ValueError(f"dt must be positive (direction goes in time_dir), got {dt}")

    if not np.any(state.status == 1):
        # integrators.py:272 _zero_dead_velocity
        if state.u is None:
            return state
        return state.replace(u=np.zeros_like(state.u), v=np.zeros_like(state.v),
                             w=None if state.w is None else np.zeros_like(state.w))

    eng = get_engine()
    if walk and mesh.elem_neighbors is None:
        build_elem_neighbors(mesh)          # core/mesh.py:2675: built lazily on first use
        eng.rebind_mesh(mesh)
    eng.bind_mesh(mesh)
    lazy = _is_lazy(vel)
    va = vel.vel_a if lazy else vel
    vb = vel.vel_b if lazy else None
    dtype = _field_dtype(va, vb, vel_next, *(recon[0] if recon is not None else ()))
    alpha0 = vel.alpha if lazy else 0.0
    if recon is not None:
        slot_a, slot_t, alphas = _recon_stage_slots(eng, recon, order, dt, dtype)
        slot_b, alpha0 = -1, 0.0
    else:
        slot_a = eng.slot_for(va, dtype)
        slot_b = eng.slot_for(vb, dtype) if vb is not None else -1
        slot_t = eng.slot_for(vel_next, dtype) if vel_next is not None else -1
        alphas = stage_alphas(order, dt, dt_to_next)
    sigma = None
    if sig3d:
        sigma = _sigma_args(eng, vel, kwargs, order, dt, dt_to_next, slot_a, slot_b, slot_t, alpha0,
                            alphas, recon)

    P = state.n_particles
    pid_off = int(diffusion.pid_offset) if diffusion is not None and hasattr(diffusion, "pid_offset") else 0
    ctx = eng.ctx
    p = ctx.particles(state.lon, state.lat, state.z, status=state.status,
                      nn=state.nearest_node, fr=state.failure_reason,
                      pid=pid_off + np.arange(P, dtype=np.int64),
                      elem=state.elem if walk else None)
    kick = None
    step_i, real_id = 0, 0
    if diffusion is not None:
        kick = dict(kh_x=diffusion.kh_x, kh_y=diffusion.kh_y, seed=diffusion.seed)
        if diffusion.az_enabled and diffusion.az_const > 0.0:
            kick["az_const"] = float(diffusion.az_const)
        if diffusion.az_enabled and is3d and _payload_az(vel) is not None:
            # integrators.py:873 vel.az: the payload's bracket, blended per value
            kick["az_field"] = (eng.slot_for(va, dtype), eng.slot_for(vb, dtype) if vb is not None else -1,
                                vel.alpha if lazy else 0.0)
        step_i, real_id = int(diffusion.step_i), int(diffusion.real_id)
    exact = 1 if exact_nearest_on(mesh) else 2
    wm = (SIGMA_W_METHODS if sig3d else W_METHODS).get(w_interp, 2)
    ctx.step(p, order, slot_a, dt, time_dir, slot_b=slot_b, slot_t=slot_t, alpha0=alpha0,
             alphas=alphas, exact_nearest=exact, w_method=wm,
             kick=kick, step_i=step_i, real_id=real_id, rk45=rk45, sigma=sigma, walk=walk)
    h = p.to_host()
    out = dict(lon=h["lon"], lat=h["lat"], z=h["z"] if is3d else state.z,
               status=h["status"], nearest_node=h["nn"], failure_reason=h["fr"],
               u=h["u"], v=h["v"], w=h["w"] if is3d else None)
    if walk:
        out["elem"] = h["elem"]
    return state.replace(**out)


def euler_step(state, vel, mesh, dt, time_dir, layer=0, **kwargs):
    """integrators.py:5150 euler_step."""
    return _step(1, state, vel, mesh, dt, time_dir, kwargs)


def rk2_step(state, vel, mesh, dt, time_dir, layer=0, **kwargs):
    """integrators.py:5611 rk2_step (midpoint)."""
    return _step(2, state, vel, mesh, dt, time_dir, kwargs)


def rk4_step(state, vel, mesh, dt, time_dir, layer=0, **kwargs):
    """integrators.py:6013 rk4_step."""
    return _step(4, state, vel, mesh, dt, time_dir, kwargs)


def rk45_step(state, vel, mesh, dt, time_dir, layer=0, **kwargs):
    """integrators.py:6448 rk45_step: adaptive Dormand-Prince 5(4) per
    particle over exactly ``dt`` (``rk45_tol_m`` metres per substep, default
    1; ``rk45_h_min`` seconds, default dt 1e-4), 2-D element-layout least
    squares, steady or blended towards ``vel_next``."""
    h_min = kwargs.get("rk45_h_min")
    if h_min is not None and float(h_min) <= 0.0:
        raise ValueError("numerics.rk45_h_min must be > 0 seconds")
    if float(kwargs.get("rk45_tol_m", 1.0)) <= 0.0:
        raise ValueError("numerics.rk45_tol_m must be > 0 m")
    return _step(5, state, vel, mesh, dt, time_dir, kwargs)


INTEGRATORS = {
    "euler": euler_step,
    "rk2": rk2_step,
    "rk4": rk4_step,
    "rk45": rk45_step,
}

ORDERS = {"euler": 1, "rk2": 2, "rk4": 4, "rk45": 5}


def get_integrator(name: str):
    """integrators.py:6650 get_integrator."""
    if name not in INTEGRATORS:
        raise ValueError(f"Unknown integrator '{name}'. Available: {sorted(INTEGRATORS)}")
    return INTEGRATORS[name]


__all__ = ["euler_step", "rk2_step", "rk4_step", "rk45_step", "get_integrator", "INTEGRATORS",
           "stage_alphas", "Engine", "get_engine"]
