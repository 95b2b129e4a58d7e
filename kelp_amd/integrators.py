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
W by ``nearest``, ``leastsq`` or ``barycentric``.  Combinations the device
path does not implement raise ``NotImplementedError`` (edge_leastsq, W idw,
sigma meshes, EM/M1 diffusion, an Az kick from an Az field); there is no CPU
fallback.
"""
from __future__ import annotations

import os
from collections import OrderedDict
from typing import Optional

import numpy as np

from .diffusion import DiffusionStepArgs
from .fields import LazyLerpField, VelocityField
from .mesh import Mesh, exact_nearest_on
from .particles import ParticleState

_C_COEFFS = {1: (0.0,), 2: (0.0, 0.5), 4: (0.0, 0.5, 0.5, 1.0)}
W_METHODS = {"nearest": 0, "leastsq": 1, "barycentric": 2}
N_SLOTS = 16


def _validate_time_aware_kwargs(vel_next, dt_to_next):
    """integrators.py:965: both or neither."""
    if (vel_next is None) != (dt_to_next is None):
        raise ValueError("vel_next and dt_to_next must be provided together "
                         f"(got vel_next={vel_next!r}, dt_to_next={dt_to_next!r})")


def _is_3d(vel) -> bool:
    """integrators.py:949: W present and u per layer."""
    base = vel.vel_a if isinstance(vel, LazyLerpField) else vel
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
    if isinstance(vel, LazyLerpField):
        if vel.vel_a.az is None or vel.vel_b.az is None:
            return None
        return vel.vel_a.az
    return getattr(vel, "az", None)


def _check_supported(vel, mesh, uv_interp, w_interp, diffusion, is3d):
    base = vel.vel_a if isinstance(vel, LazyLerpField) else vel
    ok = ((uv_interp == "barycentric" and base.uv_layout == "node")
          or (uv_interp == "leastsq" and base.uv_layout == "element"))
    if not ok:
        raise NotImplementedError(
            f"device path: uv_interp={uv_interp!r} with uv_layout={base.uv_layout!r} is not "
            "implemented (node-layout barycentric, element-layout leastsq)")
    if is3d:
        if mesh.vertical_coord != "zlevel":
            raise NotImplementedError("device path: sigma meshes are not implemented")
        if w_interp not in W_METHODS:
            raise NotImplementedError(f"device path: w_interp={w_interp!r} is not implemented "
                                      "(nearest, leastsq, barycentric)")
        if base.uv_vertical != "layer":
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


def _step(order: int, state: ParticleState, vel, mesh: Mesh, dt: float, time_dir: float,
          kwargs) -> ParticleState:
    if kwargs.get("recon") is not None:
        raise NotImplementedError("device path: pchip reconstruction is not implemented")
    if kwargs.get("elem_search", "local") != "local" or state.elem is not None:
        raise NotImplementedError("device path: elem_search='walk' is not implemented")
    uv_interp = kwargs.get("uv_interp", "leastsq")
    w_interp = kwargs.get("w_interp", "leastsq")
    vel_next = kwargs.get("vel_next")
    dt_to_next = kwargs.get("dt_to_next")
    diffusion: Optional[DiffusionStepArgs] = kwargs.get("diffusion")
    _validate_time_aware_kwargs(vel_next, dt_to_next)
    is3d = _is_3d(vel)
    rk45 = None
    if order == 5:
        # integrators.py:6485-6513: 2-D element-layout least squares only
        base = vel.vel_a if isinstance(vel, LazyLerpField) else vel
        if uv_interp != "leastsq":
            raise NotImplementedError(f"rk45 supports uv_interp='leastsq' only (v1), got {uv_interp!r}")
        if is3d:
            raise NotImplementedError("rk45 is 2-D only (v1)")
        if base.uv_layout != "element":
            raise ValueError("rk45 requires element-layout velocities (v1)")
        tol = float(kwargs.get("rk45_tol_m", 1.0))
        h_min = kwargs.get("rk45_h_min")
        rk45 = dict(tol_m=tol, h_min=None if h_min is None else float(h_min),
                    inv_dtn=0.0 if dt_to_next is None else 1.0 / float(dt_to_next))
    _check_supported(vel, mesh, uv_interp, w_interp, diffusion, is3d)
    if dt <= 0:
        raise ValueError(f"dt must be positive (direction goes in time_dir), got {dt}")

    if not np.any(state.status == 1):
        # integrators.py:272 _zero_dead_velocity
        if state.u is None:
            return state
        return state.replace(u=np.zeros_like(state.u), v=np.zeros_like(state.v),
                             w=None if state.w is None else np.zeros_like(state.w))

    eng = get_engine()
    eng.bind_mesh(mesh)
    lazy = isinstance(vel, LazyLerpField)
    va = vel.vel_a if lazy else vel
    vb = vel.vel_b if lazy else None
    dtype = _field_dtype(va, vb, vel_next)
    slot_a = eng.slot_for(va, dtype)
    slot_b = eng.slot_for(vb, dtype) if vb is not None else -1
    slot_t = eng.slot_for(vel_next, dtype) if vel_next is not None else -1
    alpha0 = vel.alpha if lazy else 0.0
    alphas = stage_alphas(order, dt, dt_to_next)

    P = state.n_particles
    pid_off = int(diffusion.pid_offset) if diffusion is not None else 0
    ctx = eng.ctx
    p = ctx.particles(state.lon, state.lat, state.z, status=state.status,
                      nn=state.nearest_node, fr=state.failure_reason,
                      pid=pid_off + np.arange(P, dtype=np.int64))
    kick = None
    step_i, real_id = 0, 0
    if diffusion is not None:
        kick = dict(kh_x=diffusion.kh_x, kh_y=diffusion.kh_y, seed=diffusion.seed)
        if diffusion.az_enabled and diffusion.az_const > 0.0:
            kick["az_const"] = float(diffusion.az_const)
        if diffusion.az_enabled and is3d and _payload_az(vel) is not None:
            # integrators.py:873 vel.az: the payload's bracket, blended per value
            kick["az_field"] = (slot_a, slot_b, alpha0)
        step_i, real_id = int(diffusion.step_i), int(diffusion.real_id)
    exact = 1 if exact_nearest_on(mesh) else 2
    ctx.step(p, order, slot_a, dt, time_dir, slot_b=slot_b, slot_t=slot_t, alpha0=alpha0,
             alphas=alphas, exact_nearest=exact, w_method=W_METHODS.get(w_interp, 2),
             kick=kick, step_i=step_i, real_id=real_id, rk45=rk45)
    h = p.to_host()
    return state.replace(lon=h["lon"], lat=h["lat"], z=h["z"] if is3d else state.z,
                         status=h["status"], nearest_node=h["nn"], failure_reason=h["fr"],
                         u=h["u"], v=h["v"], w=h["w"] if is3d else None)


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
