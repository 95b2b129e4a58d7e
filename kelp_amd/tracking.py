"""The tracking loop on the device, mirroring src/kelp/tracking.py.

``run_single_realization`` restates the per-realization loop of the
reference (tracking.py:1805-2833) around ``kelp_step``: the run clock and
the time bracket of every step (``resolve_time_bracket``, tracking.py:2947,
and the payload of ``_build_vel_payload``, :3136), the step length that
lands exactly on a record (the snapshot band, :1667/:2530-2556), delayed
release (core/release.py:112), the record schedule on the global step and
the record masks of output format 2 (core/release.py:132/:170, the backward
velocity flip of tracking.py:2772-2777).  The particle state stays in HBM
for the whole run; the integrator step, the kicks, the periodic fold and
the beaching hook of one step are one kernel launch; a record copies the
state to the host.  ``run_tracking(config)`` is the config-driven entry
(tracking.py:3288).
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .config import AppConfig, ModeConfig, NumericsConfig, resolve_time_params
from .fields import VelocityField
from .integrators import ORDERS, W_METHODS
from .mesh import Mesh, exact_nearest_on, set_exact_nearest
from .particles import ParticleState, ParticleStatus

_C = {1: (0.0,), 2: (0.0, 0.5), 4: (0.0, 0.5, 0.5, 1.0), 5: ()}
N_SLOTS = 16
#: slots of the library (kMaxSlots); with PCHIP reconstruction the last two
#: hold the reconstructed stage fields and the snapshot ring uses the rest
MAX_SLOTS = 16
RECON_SLOTS = (MAX_SLOTS - 2, MAX_SLOTS - 1)
SAMPLE_METHODS = {"nearest": 0, "leastsq": 1, "barycentric": 2, "idw": 3}
SIGMA_W_METHODS = {"nearest": 0, "leastsq": 1, "barycentric": 2, "idw": 3}


def eta_at_time(t, file_times, n, on_exhausted, direction, snap, slots):
    """tracking.py:3097 _eta_at_time (linear): the surface at time ``t`` as
    a blend ``(slot_a, slot_b, -1, alpha, 0)`` of snapshot surfaces, or
    None outside the records."""
    br = resolve_time_bracket(t, file_times, n, on_exhausted, direction, snap)
    if br is None:
        return None
    ia, ib, ta, tb, tw = br
    interval = tb - ta
    alpha = (tw - ta) / interval if interval > 0 else 0.0
    if ia == ib or abs(alpha) < 1e-12:
        return (slots.slot(ia), -1, -1, 0.0, 0.0)
    if abs(alpha - 1.0) < 1e-12:
        return (slots.slot(ib), -1, -1, 0.0, 0.0)
    return (slots.slot(ia), slots.slot(ib), -1, alpha, 0.0)


def snapshot_band(dt) -> float:
    """tracking.py:1667 _snapshot_band."""
    return 1e-5 * max(abs(float(dt)), 1.0)


def _cycle_period(file_times, n):
    """The legacy cycle rule: the span of the records plus the trailing
    interval."""
    last_dt = file_times[-1] - file_times[-2] if n > 1 else 1.0
    return (file_times[-1] - file_times[0]) + last_dt


def wrap_sim_time(sim_time, file_times, n, on_exhausted):
    """tracking.py:2919 _wrap_sim_time (legacy cycle rule)."""
    if on_exhausted != "cycle":
        return sim_time
    t0 = file_times[0]
    return t0 + ((sim_time - t0) % _cycle_period(file_times, n))


def resolve_time_bracket(sim_time, file_times, n, on_exhausted, direction=1, snap=0.0):
    """tracking.py:2947 _resolve_time_bracket (legacy cycle rule):
    ``(idx_a, idx_b, ta, tb, t_wrapped)`` or None outside the records in
    stop mode."""
    t_start, t_end = file_times[0], file_times[-1]
    tw = wrap_sim_time(sim_time, file_times, n, on_exhausted)
    if on_exhausted != "cycle" and (tw < t_start - 1e-6 or tw > t_end + 1e-6):
        return None
    idx_b = int(np.searchsorted(file_times, tw, side="right"))
    idx_a = idx_b - 1
    if on_exhausted == "cycle":
        ia, ib = idx_a % n, idx_b % n
    else:
        ia, ib = max(0, min(idx_a, n - 1)), max(0, min(idx_b, n - 1))
    ta, tb = file_times[ia], file_times[ib]
    last_dt = file_times[-1] - file_times[-2] if n > 1 else 1.0
    if on_exhausted == "cycle" and tb <= ta:
        tb = ta + last_dt
    if direction > 0 and snap > 0.0 and ia != ib and tb - tw <= snap:
        s = ib
        s_next = (s + 1) % n if on_exhausted == "cycle" else min(s + 1, n - 1)
        ta_new, tb_new = file_times[s], file_times[s_next]
        if on_exhausted == "cycle" and tb_new <= ta_new:
            tb_new = ta_new + last_dt
        return s, s_next, ta_new, tb_new, ta_new
    if direction < 0 and tw <= ta + max(1e-6, snap):
        ib_new = ia
        ia_new = (ia - 1) % n if on_exhausted == "cycle" else ia - 1
        if ia_new < 0 or ia_new == ib_new:
            return ib_new, ib_new, ta, ta, tw
        ta_new, tb_new = file_times[ia_new], file_times[ib_new]
        if tb_new <= ta_new:
            tb_new = ta_new + last_dt
        return ia_new, ib_new, ta_new, tb_new, tb_new
    return ia, ib, ta, tb, tw


def index_to_sim_time(idx, file_times, n, on_exhausted, dt=None) -> float:
    """tracking.py:2868 _index_to_sim_time (legacy cycle rule): the time of
    the unwrapped record index ``idx``."""
    if n <= 0:
        return 0.0
    if on_exhausted == "cycle":
        period = (float(dt) if dt is not None else 1.0) if n == 1 else _cycle_period(file_times, n)
        n_full, rem = divmod(int(idx), n)
        return n_full * period + float(file_times[rem])
    return float(file_times[max(0, min(int(idx), n - 1))])


def record_clock_origin(use_tinterp, sim_time, start_time_idx, file_times, n, on_exhausted, dt,
                        direction, start_step_i) -> float:
    """tracking.py:1197 _record_clock_origin: the time of the start of
    global step 0, so the record after global step g is labelled
    ``T_g0 + (g + 1) dt direction`` in a resumed run as in the whole run."""
    if use_tinterp:
        return float(sim_time) - start_step_i * dt * direction
    idx0 = int(start_time_idx) - int(start_step_i) * int(direction)
    if file_times is not None and len(file_times) >= n and n > 0:
        if n == 1:
            return float(file_times[0]) + float(idx0) * float(dt)
        return index_to_sim_time(idx0, file_times, n, on_exhausted)
    return float(idx0) * float(dt)


def step_end_on_record(sim_time, pos, dt_to_next, direction, file_times, n, on_exhausted):
    """tracking.py:2930 _step_end_on_record (legacy cycle rule)."""
    delta = pos - wrap_sim_time(sim_time, file_times, n, on_exhausted)
    if on_exhausted == "cycle":
        period = _cycle_period(file_times, n)
        delta -= period * np.round(delta / period)
    return sim_time + delta + dt_to_next * direction


@dataclass
class Payload:
    """The step's field: bracket snapshots ``ia`` (and ``ib`` blended by
    ``alpha0``), the target snapshot ``it`` (or None) and ``dt_to_next``."""
    ia: int
    ib: Optional[int]
    alpha0: float
    it: Optional[int]
    dt_to_next: Optional[float]


def vel_payload(bracket, direction) -> Payload:
    """tracking.py:3136 _build_vel_payload as snapshot indices."""
    ia, ib, ta, tb, tw = bracket
    if ia == ib:
        return Payload(ia, None, 0.0, None, None)
    interval = tb - ta
    alpha = (tw - ta) / interval if interval > 0 else 0.0
    if abs(alpha) < 1e-12:
        base = Payload(ia, None, 0.0, None, None)
    elif abs(alpha - 1.0) < 1e-12:
        base = Payload(ib, None, 0.0, None, None)
    else:
        base = Payload(ia, ib, alpha, None, None)
    if direction >= 0:
        base.it, base.dt_to_next = ib, tb - tw
        return base
    dt_prev = tw - ta
    if dt_prev <= 1e-6:
        return base
    base.it, base.dt_to_next = ia, dt_prev
    return base


@dataclass
class TrajectoryRecord:
    """In-memory trajectories of one realization, ``(particle, record)``
    arrays with the reference writer's variable names (writers.py:1209)."""
    time: List[float] = field(default_factory=list)
    rows: Dict[str, List[np.ndarray]] = field(default_factory=dict)
    meta: Dict[str, object] = field(default_factory=dict)

    def append(self, rec: Dict[str, np.ndarray], time_val: float) -> None:
        self.time.append(float(time_val))
        for k, v in rec.items():
            self.rows.setdefault(k, []).append(np.asarray(v))

    def as_arrays(self) -> Dict[str, np.ndarray]:
        out = {k: np.stack(v, axis=1) for k, v in self.rows.items()}
        out["time"] = np.asarray(self.time)
        out.update({k: v for k, v in self.meta.items() if isinstance(v, np.ndarray)})
        return out

    def save_npz(self, path: str) -> None:
        np.savez(path, **self.as_arrays())


class _SnapshotSlots:
    """Device slots holding the snapshots of the current brackets, keyed by
    time index, least recently used evicted (the reference's _VelCache,
    tracking.py:584, on HBM)."""

    def __init__(self, ctx, reader, ds, mesh, mode, numerics, sample_vars, need_eta,
                 reserve_recon=False):
        self.ctx, self.reader, self.ds, self.mesh = ctx, reader, ds, mesh
        self.n_slots = min(N_SLOTS, RECON_SLOTS[0]) if reserve_recon else N_SLOTS
        self.mode, self.numerics = mode, numerics
        self.sample_vars = sample_vars
        self.need_eta = need_eta
        self.map: Dict[int, int] = {}
        self.order: List[int] = []
        self.fields: Dict[int, VelocityField] = {}
        self.in_use: set = set()

    def field(self, idx: int) -> VelocityField:
        if idx not in self.fields:
            self.fields[idx] = self.reader.read_velocities(
                self.ds, idx, self.mesh, self.mode, diffusion=self.numerics.diffusion,
                sample_vars=self.sample_vars, velocity_float32=self.numerics.velocity_float32)
        return self.fields[idx]

    def slot(self, idx: int) -> int:
        self.in_use.add(idx)
        if idx in self.map:
            self.order.remove(idx)
            self.order.append(idx)
            return self.map[idx]
        slot = self._victim()
        self._fill(slot, idx, refill=False)
        return slot

    def begin_step(self) -> None:
        """A new step: the snapshots it acquires are protected from the
        prefetch's eviction until the next step."""
        self.in_use = set()

    def _victim(self) -> int:
        if len(self.map) < self.n_slots:
            return min(set(range(self.n_slots)) - set(self.map.values()))
        for old in self.order:
            if old not in self.in_use:
                self.order.remove(old)
                self.fields.pop(old, None)
                return self.map.pop(old)
        raise RuntimeError("every snapshot slot is in use by this step")

    def _fill(self, slot: int, idx: int, refill: bool) -> None:
        vf = self.field(idx)
        eta = None
        if self.need_eta:
            eta = vf.eta if vf.eta is not None else np.zeros(self.mesh.n_nodes)
        if refill:
            # the snapshot ring: the copy runs behind the enqueued steps
            self.ctx.upload_field_async(slot, vf.u, vf.v, vf.w, eta=eta, uv_layout=vf.uv_layout)
        else:
            self.ctx.upload_field(slot, vf.u, vf.v, vf.w, uv_layout=vf.uv_layout)
            if eta is not None:
                self.ctx.upload_eta(slot, eta)
        if vf.az is not None and vf.w is not None:
            self.ctx.upload_az(slot, vf.az)
        for k, name in enumerate(self.sample_vars or ()):
            self.ctx.upload_scalar(slot, k, vf.scalars[name])
        self.map[idx] = slot
        self.order.append(idx)

    def prefetch(self, idx: int) -> None:
        """tracking.py:584 _VelCache read-ahead on the device: snapshot
        ``idx`` is uploaded into a slot no snapshot of this step uses, as an
        asynchronous refill when that slot already holds a snapshot of the
        same shape (the usual case once the run has filled its slots)."""
        if idx in self.map:
            return
        full = len(self.map) >= self.n_slots
        if full and all(old in self.in_use for old in self.order):
            return          # every slot serves this step: no read-ahead
        slot = self._victim()
        self._fill(slot, idx, refill=full)


def _beach_args(numerics: NumericsConfig, dt):
    b = numerics.beaching
    if not b.enabled:
        return None
    seed = b.seed if b.seed is not None else numerics.diffusion.seed
    return dict(criteria=list(b.criteria), depth_min=b.depth_min, use_eta=b.use_eta,
                zone_depth=b.zone_depth, t_beach=b.t_beach, resuspension=b.resuspension,
                rewet_hysteresis=b.rewet_hysteresis, min_wet_steps=b.min_wet_steps,
                t_resuspend=b.t_resuspend, seed=seed, dt=dt)


def _kick_args(numerics: NumericsConfig):
    d = numerics.diffusion
    if not d.enabled:
        return None
    if d.scheme != "uniform":
        raise NotImplementedError(f"device path: diffusion scheme {d.scheme!r}")
    kick = dict(kh_x=d.kh_zonal, kh_y=d.kh_meridional, seed=d.seed)
    if d.az_enabled and d.az_const > 0.0:
        kick["az_const"] = d.az_const
    return kick


def _check_path(numerics: NumericsConfig, mesh: Mesh, is3d: bool, uv_layout: str) -> None:
    if numerics.integrator not in ORDERS:
        raise NotImplementedError(f"device path: integrator {numerics.integrator!r}")
    if numerics.temporal_reconstruction not in ("linear", "pchip"):
        raise ValueError(f"numerics.temporal_reconstruction: unknown {numerics.temporal_reconstruction!r}")
    if numerics.elem_search not in ("local", "walk"):
        raise ValueError(f"numerics.elem_search: unknown {numerics.elem_search!r}")
    if numerics.integrator == "rk45":
        # integrators.py:6485-6513
        if is3d or numerics.uv_interp != "leastsq" or uv_layout != "element":
            raise NotImplementedError("rk45 is 2-D element-layout least squares only (v1)")
        if numerics.elem_search == "walk":
            raise ValueError("rk45 does not support elem_search='walk' (v1)")
    if numerics.elem_search == "walk" and (uv_layout != "node" or mesh.vertical_coord == "sigma"):
        raise NotImplementedError("device path: elem_search='walk' with element-layout u/v or on a "
                                  "sigma mesh")
    want = "node" if numerics.uv_interp == "barycentric" else "element"
    if numerics.uv_interp not in ("barycentric", "leastsq") or uv_layout != want:
        raise NotImplementedError(f"device path: uv_interp={numerics.uv_interp!r} with "
                                  f"{uv_layout}-layout u/v")
    if is3d:
        if mesh.vertical_coord not in ("zlevel", "sigma"):
            raise NotImplementedError(f"device path: vertical_coord {mesh.vertical_coord!r}")
        methods = SIGMA_W_METHODS if mesh.vertical_coord == "sigma" else W_METHODS
        if numerics.w_interp not in methods:
            raise NotImplementedError(f"device path: w_interp={numerics.w_interp!r}")
    if numerics.beaching.enabled and numerics.beaching.dry_mask_var:
        raise NotImplementedError("device path: beaching dry_mask_var")


def resolve_recon_stencil(bracket, n):
    """tracking.py:3075 _resolve_recon_stencil: the four snapshots
    ``(ia - 1, ia, ib, ib + 1)`` around a bracket and its query time, or None
    on a degenerate bracket, across a cycle seam or at the first/last
    interval of the records (the step then blends linearly)."""
    ia, ib, _ta, _tb, q = bracket
    if ib != ia + 1 or ia - 1 < 0 or ib + 1 >= n:
        return None
    return (ia - 1, ia, ib, ib + 1), q


def _sigma_surfaces(use_tinterp, sim_time, sim_time_end, dt, direction, file_times, n,
                    on_exhausted, snap, slots, pay, order, sa, sb, st, alphas, step_len, dtn,
                    file_idx):
    """The surfaces a sigma step starts and ends on (tracking.py:2562-2594:
    ``sigma_eta_start``/``sigma_eta_end``, one function of time), with the
    integrators' defaults where the tracking layer supplies none
    (integrators.py:1381 _sigma_end_eta, :5806-5812)."""
    start = end = None
    if use_tinterp:
        start = eta_at_time(sim_time, file_times, n, on_exhausted, direction, snap, slots)
        t_end = sim_time_end if sim_time_end is not None else sim_time + dt * direction
        end = eta_at_time(t_end, file_times, n, on_exhausted, direction, snap, slots)
    else:
        nxt = file_idx + direction
        if on_exhausted == "cycle":
            nxt %= n
        if 0 <= nxt < n:
            end = (slots.slot(nxt), -1, -1, 0.0, 0.0)
    if start is None:
        # the step field's own surface (stage 0 of the velocity blend)
        start = (sa, sb, st, pay.alpha0, 0.0)
    if end is None and st >= 0:
        # not steady: the velocity blend at the step end
        a_end = alphas[3] if order == 4 else step_len / dtn
        end = (sa, sb, st, pay.alpha0, a_end)
    return dict(start=start, end=end)


def activate_released(status: np.ndarray, release_step, global_step_i: int):
    """core/release.py:112: NOT_RELEASED -> ALIVE once the global step
    reaches the particle's release step."""
    if release_step is None:
        return status
    act = (status == ParticleStatus.NOT_RELEASED) & (int(global_step_i) >= np.asarray(release_step))
    if not act.any():
        return status
    out = status.copy()
    out[act] = ParticleStatus.ALIVE
    return out


def mask_record(h: Dict[str, np.ndarray], prev_status: np.ndarray, direction: int,
                is3d: bool) -> Dict[str, np.ndarray]:
    """core/release.py:132 mask_dormant_for_record + :170
    mask_inactive_velocity + the backward flip (tracking.py:2772)."""
    out = dict(h)
    dormant = h["status"] == ParticleStatus.NOT_RELEASED
    bad = (h["status"] != ParticleStatus.ALIVE) | (prev_status != ParticleStatus.ALIVE)
    for k in ("lon", "lat", "z"):
        a = np.array(h[k], dtype=np.float64, copy=True)
        a[dormant] = np.nan
        out[k] = a
    for k in ("u", "v", "w"):
        a = np.array(h[k], dtype=np.float64, copy=True)
        a[bad | dormant] = np.nan
        if direction < 0:
            a = -a
        out[k] = a
    return out


class RecordStager:
    """The records of a run kept in HBM and copied to the host in batches:
    each record is masked and put in particle-id order on the device
    (``kelp_record_stage``: core/release.py:132 mask_dormant_for_record,
    :170 mask_inactive_velocity, the backward flip of tracking.py:2772),
    into a block of ``capacity`` records per variable; a full block (and
    :meth:`flush`) waits for the device once and moves the block with one
    copy per variable.  The rows reaching ``rec`` are those :func:`mask_record`
    forms from the host state, bit for bit."""

    def __init__(self, ctx, p, pid_offset: int, rec: "TrajectoryRecord", sample_vars,
                 budget_bytes: int = 256 << 20, max_records: int = 64):
        t = ctx.torch
        dev = p.lon.device
        self.ctx, self.p, self.rec = ctx, p, rec
        self.pid_offset = int(pid_offset)
        self.P = int(p.n)
        self.names = list(sample_vars or [])
        nv = len(self.names)
        per = self.P * (8 * 6 + 4 + 1 + 8 * nv)
        self.R = int(max(1, min(max_records, budget_bytes // max(per, 1))))
        f64 = lambda: t.empty((self.R, self.P), dtype=t.float64, device=dev)
        self.buf = {k: f64() for k in ("lon", "lat", "z", "u", "v")}
        if p.w is not None:
            self.buf["w"] = f64()
        self.buf["status"] = t.empty((self.R, self.P), dtype=t.int32, device=dev)
        self.buf["fr"] = t.empty((self.R, self.P), dtype=t.int8, device=dev)
        if nv:
            self.buf["samples"] = t.empty((self.R, nv, self.P), dtype=t.float64, device=dev)
        self.sample_rows = t.empty((nv, self.P), dtype=t.float64, device=dev) if nv else None
        self.times: List[float] = []

    def append(self, p, prev_status, direction: int, sampled, time_val: float) -> None:
        c = len(self.times)
        rows = {k: b[c] for k, b in self.buf.items()}
        self.ctx.record_stage(p, prev_status, self.pid_offset, self.P, float(direction), rows, sampled)
        self.times.append(float(time_val))
        if len(self.times) == self.R:
            self.flush()

    def flush(self) -> None:
        c = len(self.times)
        if c == 0:
            return
        self.ctx.synchronize()
        host = {k: b[:c].cpu().numpy() for k, b in self.buf.items()}
        for j in range(c):
            row = dict(lon=host["lon"][j], lat=host["lat"][j], z=host["z"][j], status=host["status"][j],
                       failure_reason=host["fr"][j], vel_u=host["u"][j], vel_v=host["v"][j])
            if "w" in host:
                row["vel_w"] = host["w"][j]
            for k, name in enumerate(self.names):
                row["sample__" + name] = host["samples"][j, k]
            self.rec.append(row, self.times[j])
        self.times = []


def run_single_realization(mesh: Mesh, ds, reader, init_state: ParticleState, mode: ModeConfig,
                           numerics: NumericsConfig, *, start_time_idx: int = 0,
                           n_steps: int, n_time_total: Optional[int] = None,
                           temporal_interp: Optional[bool] = None, file_times=None,
                           file_dt: Optional[float] = None, sample_vars=None,
                           realization_id: int = 0, particle_offset: int = 0,
                           start_step_i: int = 0, on_exhausted: str = "stop",
                           start_wet_streak=None, ctx=None, resort_every: int = 0,
                           writer: Optional[TrajectoryRecord] = None,
                           sim_time_start: Optional[float] = None,
                           checkpoint_freq: Optional[int] = None,
                           checkpoint_path: Optional[str] = None,
                           checkpoint_sink=None) -> TrajectoryRecord:
    """One realization on the device (tracking.py:1805).  Returns the
    records, ``(particle, record)``: lon, lat, z, status, failure_reason,
    vel_u/vel_v/vel_w (NaN where undefined), time.

    Checkpoints (tracking.py:2785-2862): with ``checkpoint_path`` and
    ``checkpoint_freq`` (default ``numerics.checkpoint_freq``) > 0 the
    device state is written every ``checkpoint_freq`` global steps and after
    the last step that ran (:func:`write_checkpoint`, the reference's
    restart variables); :func:`load_restart` returns the arguments that
    resume the run bit for bit.  ``checkpoint_sink`` (a callable taking the
    :func:`checkpoint_payload`) replaces the file at ``checkpoint_path``,
    e.g. to keep the checkpoint inside the trajectory file."""
    from .device import DeviceContext
    sample_vars = list(sample_vars or [])
    if sample_vars and numerics.sample_interp not in SAMPLE_METHODS:
        raise ValueError(f"Unknown sample_interp {numerics.sample_interp!r}")
    direction = 1 if mode.time_dir >= 0 else -1
    order = ORDERS.get(numerics.integrator)
    if file_times is None:
        from .readers import extract_time_info
        file_times, ax_dt = extract_time_info(ds)
        file_dt = file_dt if file_dt is not None else ax_dt
    file_times = None if file_times is None else np.asarray(file_times, dtype=np.float64)
    n_time_total = n_time_total if n_time_total is not None else len(file_times)
    dt = float(numerics.dt if numerics.dt is not None else file_dt)
    temporal_interp = numerics.temporal_interp if temporal_interp is None else temporal_interp
    use_tinterp = bool(temporal_interp) and file_times is not None
    is3d = bool(mode.Fdrift3D)
    uv_layout = reader.probe_uv_layout(ds, mode, mesh)
    _check_path(numerics, mesh, is3d, uv_layout)
    multistage = order > 1
    rk_snap = use_tinterp and multistage
    snap = snapshot_band(dt) if rk_snap else 5e-10 * max(dt, 1.0)
    exact = 1 if exact_nearest_on(mesh) else 2
    sig3d = is3d and mesh.vertical_coord == "sigma"
    # tracking.py:2159: PCHIP in time, on multistage steps of a time-
    # interpolated run; each stage reads its own reconstructed field
    recon_on = use_tinterp and multistage and numerics.temporal_reconstruction == "pchip"
    if recon_on:
        if order != 2:
            raise NotImplementedError(f"device path: pchip reconstruction with {numerics.integrator} (rk2 only)")
        if sig3d:
            raise NotImplementedError("device path: pchip reconstruction on a sigma mesh")
        if numerics.beaching.enabled and numerics.beaching.use_eta:
            raise NotImplementedError("device path: pchip reconstruction with beaching use_eta")
    w_method = ((SIGMA_W_METHODS if sig3d else W_METHODS)[numerics.w_interp] if is3d else 2)
    kick = _kick_args(numerics)
    d = numerics.diffusion
    az_field = bool(is3d and d.enabled and d.az_enabled and d.var_az)
    if az_field and sig3d:
        # integrators.py:902-906
        raise ValueError("diffusion.var_az is not supported on sigma meshes (v1) -- use the "
                         "constant az_const source instead")
    beach = _beach_args(numerics, dt)
    periodic = dict(periodic_lon=numerics.periodic_lon, periodic_lat=numerics.periodic_lat)

    walk = numerics.elem_search == "walk"
    if walk and mesh.elem_neighbors is None:
        from .mesh import build_elem_neighbors
        build_elem_neighbors(mesh)
        if ctx is not None and ctx.mesh is mesh:
            ctx.upload_mesh(mesh)
    if beach is not None and {"coastal_element", "stochastic"} & set(beach["criteria"]):
        # tracking.py:2085: the coastal-element flags of the mesh boundary
        if mesh.elem_is_coastal is None:
            from .mesh import build_boundary_info
            build_boundary_info(mesh)
            if ctx is not None and ctx.mesh is mesh:
                ctx.upload_mesh(mesh)
    ctx = ctx or DeviceContext(0)
    if ctx.mesh is not mesh:
        ctx.upload_mesh(mesh)
    slots = _SnapshotSlots(ctx, reader, ds, mesh, mode, numerics, sample_vars,
                           need_eta=bool(beach and beach["use_eta"]) or sig3d,
                           reserve_recon=recon_on)
    P = init_state.n_particles
    pid = particle_offset + np.arange(P, dtype=np.int64)
    ws = (np.zeros(P, np.int32) if start_wet_streak is None
          else np.asarray(start_wet_streak, np.int32))
    fr0 = (np.zeros(P, np.int8) if init_state.failure_reason is None
           else init_state.failure_reason)
    # moving particles start with nn unknown (-1): every anchor the run
    # stores is then its own, which lets each step reuse the previous step's
    # end anchor; beached and outside particles keep theirs (the beaching
    # hook reads it, e.g. in a run resumed from a checkpoint)
    nn0 = None
    if init_state.nearest_node is not None:
        moving = ((init_state.status == ParticleStatus.ALIVE)
                  | (init_state.status == ParticleStatus.NOT_RELEASED))
        nn0 = np.where(moving, -1, init_state.nearest_node).astype(np.int32)
    elem0 = None
    if walk:
        # tracking.py:2236-2253: the carried elements are seeded by one
        # global search (a resumed state brings its own)
        elem0 = init_state.elem
        if elem0 is None:
            nn_s = ctx.nearest(init_state.lon, init_state.lat, exact=exact == 1)
            elem0 = ctx.containing(init_state.lon, init_state.lat, nn_s)
    p = ctx.particles(init_state.lon, init_state.lat, init_state.z, status=init_state.status,
                      nn=nn0, fr=fr0, pid=pid, ws=ws, elem=elem0)
    torch = ctx.torch
    release_step = init_state.release_step
    rel_dev = None if release_step is None else torch.as_tensor(
        np.asarray(release_step, np.int64), device=p.lon.device)
    prev_status = torch.empty_like(p.status)
    rec = writer if writer is not None else TrajectoryRecord()
    sim_time = None
    if use_tinterp:
        if not 0 <= start_time_idx < n_time_total and on_exhausted != "cycle":
            raise ValueError(f"start_time_idx {start_time_idx} is outside the records [0, {n_time_total})")
        sim_time = (file_times[start_time_idx % n_time_total] if sim_time_start is None
                    else float(sim_time_start))
    t_g0 = record_clock_origin(use_tinterp, sim_time, start_time_idx, file_times, n_time_total,
                               on_exhausted, dt, direction, start_step_i)

    ck_freq = int(numerics.checkpoint_freq if checkpoint_freq is None else checkpoint_freq)
    ck_on = (bool(checkpoint_path) or checkpoint_sink is not None) and ck_freq > 0
    last_g = None

    stager = RecordStager(ctx, p, particle_offset, rec, sample_vars)

    def checkpoint(g_done, next_idx):
        stager.flush()      # the file written with a checkpoint holds every record so far
        meta = rec.meta if rec is not None else {}
        release = ({k: meta["release_" + k] for k in ("lon", "lat", "z")}
                   if all(("release_" + k) in meta for k in ("lon", "lat", "z")) else None)
        payload = checkpoint_payload(p.to_host(), g_done, next_idx, sim_time if use_tinterp else None,
                                     release_step=release_step, particle_offset=particle_offset, release=release)
        if checkpoint_sink is not None:
            checkpoint_sink(payload)
        else:
            from .writers import write_restart_file
            write_restart_file(checkpoint_path, **payload)

    prefetch = int(getattr(numerics, "prefetch_size", 1) or 0) > 0
    recon_warned = False
    for step_i in range(n_steps):
        g = start_step_i + step_i
        slots.begin_step()
        # the re-sort permutes the particle arrays, so it runs before the
        # activation writes the step-start status (prev_status is indexed
        # by device slot and is not among the arrays the sort permutes)
        if resort_every > 0 and g % resort_every == 0:
            ctx.sort(p)
        ctx.activate(p, rel_dev, particle_offset, g, prev_status)
        if use_tinterp:
            bracket = resolve_time_bracket(sim_time, file_times, n_time_total, on_exhausted,
                                           direction, snap)
            if bracket is None:
                break
            pay = vel_payload(bracket, direction)
        else:
            file_idx = start_time_idx + step_i * direction
            if on_exhausted == "cycle":
                file_idx %= n_time_total
            elif not 0 <= file_idx < n_time_total:
                break
            nxt = file_idx + direction
            have_next = False
            if file_times is not None:
                if on_exhausted == "cycle" and n_time_total > 1:
                    nxt, have_next = nxt % n_time_total, True
                elif 0 <= nxt < n_time_total:
                    have_next = True
            pay = Payload(file_idx, None, 0.0, None, None)
            if have_next:
                ta, tb = file_times[file_idx], file_times[nxt]
                last_dt = file_times[-1] - file_times[-2] if n_time_total > 1 else 1.0
                if on_exhausted == "cycle" and direction > 0 and tb <= ta:
                    tb = ta + last_dt
                elif direction < 0 and tb >= ta:
                    tb = ta - last_dt
                pay.it, pay.dt_to_next = nxt, abs(tb - ta)
        step_len = dt
        sim_time_end = None
        dtn = pay.dt_to_next
        if rk_snap and dtn is not None and abs(dt - dtn) <= snap:
            step_len = dtn
            sim_time_end = step_end_on_record(sim_time, bracket[4], dtn, direction, file_times,
                                              n_time_total, on_exhausted)
        elif dtn is not None and dtn < dt <= dtn + snap:
            dtn = dt
        stencil = resolve_recon_stencil(bracket, n_time_total) if recon_on else None
        if recon_on and stencil is None and not recon_warned:
            warnings.warn("temporal_reconstruction='pchip': fewer than 4 snapshots at the range "
                          "edge; this step uses linear time interpolation. Further occurrences in "
                          "this realization suppressed.", RuntimeWarning, stacklevel=2)
            recon_warned = True
        if dtn is not None and step_len > dtn:
            pay.it, dtn = None, None           # steady RK for this step
            stencil = None                     # tracking.py:2560: no PCHIP either
        sa = slots.slot(pay.ia)
        sb = slots.slot(pay.ib) if pay.ib is not None else -1
        st = slots.slot(pay.it) if pay.it is not None else -1
        alphas = [0.0] * 4
        if dtn is not None and order != 5:
            alphas = [c * step_len / dtn for c in _C[order]] + [0.0] * (4 - order)
        rk45 = None
        if order == 5:
            # integrators.py:6533-6545: the substep fractions t / dt_to_next
            rk45 = dict(tol_m=numerics.rk45_tol_m, h_min=numerics.rk45_h_min,
                        inv_dtn=0.0 if dtn is None else 1.0 / dtn)
        sigma = None
        if sig3d:
            sigma = _sigma_surfaces(use_tinterp, sim_time, sim_time_end, dt, direction, file_times,
                                    n_time_total, on_exhausted, snap, slots, pay, order, sa, sb,
                                    st, alphas, step_len, dtn, file_idx if not use_tinterp else None)
            sigma.update(Dmin=numerics.Dmin, z0b_min=numerics.z0b_min)
        step_a, step_b, step_t, step_a0 = sa, sb, st, pay.alpha0
        if stencil is not None:
            # temporal_reconstruct.py:298 make_stage_vels_recon: stage c at
            # q + direction c dt, reconstructed on the device from the four
            # resident snapshots; alpha 0 / 1 select the two stage fields
            idx4, q = stencil
            src = [slots.slot(i) for i in idx4]
            t4 = [float(file_times[i]) for i in idx4]
            for dst, c in zip(RECON_SLOTS, _C[2]):
                ctx.reconstruct(dst, src, t4, float(q) + float(direction) * c * step_len)
            step_a, step_b, step_t, step_a0 = RECON_SLOTS[0], -1, RECON_SLOTS[1], 0.0
            alphas = [0.0, 1.0, 0.0, 0.0]
        step_kick = kick
        if az_field:
            # integrators.py:873: the Az of the step payload (the bracket)
            step_kick = dict(kick, az_field=(sa, sb, pay.alpha0))
        ctx.step(p, order, step_a, step_len, float(direction), slot_b=step_b, slot_t=step_t,
                 alpha0=step_a0, alphas=alphas, exact_nearest=exact, w_method=w_method,
                 kick=step_kick, beach=beach, step_i=g, real_id=realization_id, reuse_anchor=True,
                 sigma=sigma, walk=walk, rk45=rk45, **periodic)
        if prefetch and n_time_total and n_time_total > 1:
            # the snapshot after this step's bracket streams in while it runs
            held = [i for i in (pay.ia, pay.ib, pay.it) if i is not None]
            held += list(stencil[0]) if stencil is not None else []
            ahead = max(held) if direction > 0 else min(held)
            nxt = ahead + direction
            if on_exhausted == "cycle":
                nxt %= n_time_total
            if 0 <= nxt < n_time_total:
                slots.prefetch(nxt)
        if use_tinterp:
            sim_time = sim_time_end if sim_time_end is not None else sim_time + dt * direction
        if (g + 1) % max(int(numerics.output_freq), 1) == 0:
            if sample_vars:
                # the scalars at the record's instant (tracking.py:2692-2743)
                ra, rb, ral = sa, sb, pay.alpha0
                if use_tinterp:
                    rbr = resolve_time_bracket(sim_time, file_times, n_time_total, on_exhausted,
                                               direction, snap)
                    if rbr is not None:
                        rp = vel_payload(rbr, direction)
                        ra = slots.slot(rp.ia)
                        rb = slots.slot(rp.ib) if rp.ib is not None else -1
                        ral = rp.alpha0
                else:
                    nxt = file_idx + direction
                    if on_exhausted == "cycle":
                        nxt %= n_time_total
                    if 0 <= nxt < n_time_total:
                        ra, rb, ral = slots.slot(nxt), -1, 0.0
                sampled = ctx.sample_vars(p, range(len(sample_vars)), ra, rb, ral, numerics.sample_interp,
                                          out=stager.sample_rows)
            stager.append(p, prev_status, direction, sampled if sample_vars else None,
                          t_g0 + (g + 1) * dt * direction)
        last_g = g
        next_idx = (start_time_idx + (step_i + 1) * direction) if not use_tinterp else None
        if use_tinterp:
            nb = resolve_time_bracket(sim_time, file_times, n_time_total, on_exhausted, direction, snap)
            next_idx = nb[0] if nb is not None else None
        if ck_on and (g + 1) % ck_freq == 0:
            checkpoint(g, next_idx)
    if ck_on and last_g is not None and (last_g + 1) % ck_freq != 0:
        checkpoint(last_g, next_idx)
    stager.flush()
    return rec


def checkpoint_payload(h: Dict[str, np.ndarray], step_i: int, next_time_idx, sim_time=None, release_step=None,
                       particle_offset: int = 0, release=None) -> dict:
    """The keyword arguments of :func:`~kelp_amd.writers.write_restart_file`
    for the device state ``h`` (as ``DeviceParticles.to_host`` returns it,
    put in particle-id order): the last fully executed global step, the
    record index the next step reads (``restart_file_idx``, the
    ``start_time_idx`` of a resume; -1 when there is none), the wet
    streaks, release schedule and positions, shard offset and, for a
    time-interpolated run, the clock."""
    order = np.argsort(h["pid"], kind="stable")
    take = lambda k: np.asarray(h[k])[order]
    rel = release or {}
    st = ParticleState(lon=take("lon"), lat=take("lat"), z=take("z"), status=take("status").astype(np.int32),
                       nearest_node=take("nn").astype(np.int32), failure_reason=take("fr").astype(np.int8),
                       elem=take("elem").astype(np.int32) if h.get("elem") is not None else None,
                       release_lon=rel.get("lon"), release_lat=rel.get("lat"), release_z=rel.get("z"))
    return dict(state=st, step_i=int(step_i), file_idx=-1 if next_time_idx is None else int(next_time_idx),
                wet_streak=take("ws").astype(np.int32),
                release_step=None if release_step is None else np.asarray(release_step),
                particle_offset=int(particle_offset), sim_time=sim_time)


def write_checkpoint(path: str, h: Dict[str, np.ndarray], step_i: int, next_time_idx,
                     sim_time=None, release_step=None, particle_offset: int = 0, release=None) -> None:
    """writers.py:1420 write_restart_file for the device state: the
    reference's restart variables in a classic NetCDF file, written to a
    temporary sibling, synced and renamed (an interrupted checkpoint never
    replaces a valid one; ``restart_valid`` is 1 only in a complete one)."""
    from .writers import write_restart_file
    write_restart_file(path, **checkpoint_payload(h, step_i, next_time_idx, sim_time, release_step,
                                                  particle_offset, release))


def load_restart(path: str) -> Dict[str, object]:
    """writers.py:1541 load_restart for a restart file or a trajectory file
    with a restart block (this build's or the reference's): the reference's
    keys, plus the keyword arguments of :func:`run_single_realization`
    that continue the run -- ``start_step_i`` (the next global step),
    ``start_time_idx``, ``sim_time_start`` (None for a discrete run),
    ``start_wet_streak`` (zeros when not stored), ``particle_offset``
    (0 when not stored)."""
    from .writers import load_restart as _load
    r = _load(path)
    st = r["state"]
    if st.failure_reason is None:
        st = st.replace(failure_reason=np.zeros(st.n_particles, np.int8))
    ws = r["wet_streak"]
    out = dict(r)
    out.update(state=st, start_step_i=r["step_i"] + 1, start_time_idx=max(r["file_idx"], 0),
               sim_time_start=r["sim_time"],
               start_wet_streak=np.zeros(st.n_particles, np.int32) if ws is None else ws,
               particle_offset=int(r["particle_offset"] or 0))
    return out


def shift_to_midpoints(file_times: np.ndarray) -> np.ndarray:
    """tracking.py:1427 _shift_to_midpoints: end-of-window stamps of
    time-averaged records to the midpoints of their windows (the first one
    extrapolated by half the first interval)."""
    t = np.asarray(file_times, dtype=np.float64)
    if t.size < 2:
        return t.copy()
    out = np.empty_like(t)
    out[1:] = 0.5 * (t[:-1] + t[1:])
    out[0] = t[0] - 0.5 * (t[1] - t[0])
    return out


def resolve_record_times(config: AppConfig, ds, mesh_type: str = "fesom2"):
    """tracking.py:3416-3500: the record times of ``ds`` (seconds since its
    first record), shifted to the window midpoints under
    ``data.time_averaging: mean`` (the FESOM2 and HEALPix default when
    unset, with a warning; FESOM-C and the others default to instant), the
    snapshot interval (``numerics.file_dt`` wins) and the step (``dt``, or
    the snapshot interval)."""
    from .readers import extract_time_info
    file_times, file_dt = extract_time_info(ds)
    avg = config.data.time_averaging
    if avg is None:
        avg = "mean" if ("fesom2" in mesh_type or "healpix" in mesh_type) else "instant"
        warnings.warn(f"data.time_averaging is not set -- assuming {avg!r}. Please set "
                      "data.time_averaging explicitly to 'mean' or 'instant' in your config.",
                      UserWarning, stacklevel=3)
    elif avg not in ("mean", "instant"):
        raise ValueError(f"data.time_averaging must be 'mean' or 'instant', got {avg!r}")
    if avg == "mean" and file_times is not None and len(file_times) > 1:
        file_times = shift_to_midpoints(file_times)
    if config.numerics.file_dt is not None:
        file_dt = float(config.numerics.file_dt)
        if file_times is None:
            file_times = np.arange(ds.time.size, dtype=np.float64) * file_dt
    dt = config.numerics.dt if config.numerics.dt is not None else file_dt
    if dt is None:
        raise ValueError("numerics.dt is not set and the dataset has no time step to take it from")
    return file_times, file_dt, float(dt), avg


def run_tracking(config: AppConfig, *, mesh: Mesh, dataset, reader="fesom2_arrays",
                 init_state: ParticleState, ctx=None) -> List[TrajectoryRecord]:
    """tracking.py:3288 run_tracking for in-memory inputs: the mesh, the
    dataset and the seed state are passed in (the file readers of the
    reference are outside this build); the rest comes from ``config``:
    the record times (:func:`resolve_record_times`), the exact-nearest
    guard (tracking.py:3370), the sampled variables (``mode.sample_vars``)
    and ``time.num_runs`` realizations, each shifted by
    ``time.run_interval`` records.  Every record set carries the release
    positions and time (``release_*``, writers.py:1209).  With
    ``data.output_path`` each realization writes its file, which also holds
    the last checkpoint when ``numerics.checkpoint_freq`` > 0 (rewritten at
    every checkpoint, so :func:`load_restart` on the output file resumes an
    interrupted run)."""
    from .readers import get_reader
    rd = get_reader(reader) if isinstance(reader, str) else reader
    ds = rd.open_dataset(dataset)
    set_exact_nearest(mesh, bool(config.mesh.exact_nearest))
    # tracking.py:3378-3386: an explicit grid_hash_cell_size is pinned,
    # otherwise mesh.grid_hash_cell_rule (occupancy by default) decides the
    # cell, after the exact-nearest flag whose mode it depends on
    from .cellrule import apply_grid_hash_cell_rule
    from .device import DeviceContext
    from .mesh import build_grid_hash
    ctx = ctx or DeviceContext(0)
    if getattr(config.mesh, "grid_hash_cell_size", None) is not None:
        cs = float(config.mesh.grid_hash_cell_size)
        if not (np.isfinite(cs) and 0.0 < cs < 45.0):
            raise ValueError(f"grid-hash cell_size must be a finite value in (0, 45) degrees, got {cs!r}")
        mesh.metadata["grid_hash_cell"] = cs
        mesh.metadata["grid_hash_cell_rule"] = "explicit"
        if mesh.gh_cell_size != cs:
            build_grid_hash(mesh)
    else:
        if ctx.mesh is not mesh:
            ctx.upload_mesh(mesh)
        before = mesh.gh_cell_size
        apply_grid_hash_cell_rule(mesh, ctx, getattr(config.mesh, "grid_hash_cell_rule", None) or "occupancy")
        if mesh.gh_cell_size != before:
            ctx.upload_mesh(mesh)
    start, n_steps = resolve_time_params(config.time)
    file_times, file_dt, dt, avg = resolve_record_times(config, ds, str(config.mesh.type).lower())
    numerics = config.numerics
    if numerics.dt is None:
        import copy
        numerics = copy.copy(numerics)
        numerics.dt = dt
    n = int(ds.time.size)
    from .writers import run_attributes, write_restart_file, write_trajectories
    out = []
    for r in range(int(config.time.num_runs)):
        s0 = start + r * int(config.time.run_interval)
        rec = TrajectoryRecord()
        rec.meta.update(release_lon=np.asarray(init_state.lon, np.float64).copy(),
                        release_lat=np.asarray(init_state.lat, np.float64).copy(),
                        release_z=np.asarray(init_state.z, np.float64).copy(),
                        release_time=np.full(init_state.n_particles,
                                             index_to_sim_time(s0, file_times, n, config.time.on_exhausted,
                                                               dt) if file_times is not None else 0.0),
                        time_averaging=avg)
        path = None
        if config.data.output_path:
            # tracking.py:4015-4023: one file per realization (_r<i> when several)
            import os
            path = str(config.data.output_path)
            if int(config.time.num_runs) > 1:
                stem, ext = os.path.splitext(path)
                path = f"{stem}_r{r}{ext}"
        origin = ds.attrs.get("time_origin") if isinstance(ds.attrs, dict) else None
        attrs = run_attributes(config, s0, n_steps, avg, mesh, numerics=numerics)
        last_ck = {}

        def write_out(restart, path=path, rec=rec, s0=s0, attrs=attrs):
            """The realization's file: the records so far and, with
            checkpoints on, the last checkpoint (tracking.py:2785-2862 keeps
            it in the trajectory file); a checkpoint before the first record
            is a restart-only file."""
            if not rec.rows:
                if restart is not None:
                    write_restart_file(path, **restart)
                return
            arrays = rec.as_arrays()
            if not config.mode.Fdrift3D:
                arrays.pop("vel_w", None)      # no vel_w in 2-D
            write_trajectories(path, arrays, f"seconds since {origin}" if origin else "seconds",
                               "standard", attrs, restart=restart)

        def sink(payload, write_out=write_out, last_ck=last_ck):
            last_ck["payload"] = payload
            write_out(payload)

        ck_on = path is not None and int(numerics.checkpoint_freq or 0) > 0
        out.append(run_single_realization(
            mesh, ds, rd, init_state, config.mode, numerics, start_time_idx=s0, n_steps=n_steps,
            n_time_total=n, file_times=file_times, file_dt=file_dt,
            sample_vars=list(config.mode.sample_vars or []),
            on_exhausted=config.time.on_exhausted, realization_id=r,
            particle_offset=int(config.data.particle_offset or 0), ctx=ctx, writer=rec,
            checkpoint_sink=sink if ck_on else None))
        if path is not None:
            write_out(last_ck.get("payload"))
    return out
