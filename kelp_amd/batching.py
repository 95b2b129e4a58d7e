"""Batched clouds: many independent releases advanced together on one GPU.

The reference runs every release as its own loop (``run_single_realization``,
src/kelp/tracking.py:1805; ``utils/split_seeds.py:89`` shards one release
into several such runs).  :class:`Batch` keeps many such runs in one particle
pool in device memory and advances all of them with one kernel per step
(``kelp_batch_step``), or a whole run of steps as one captured HIP graph
(``kelp_batch_run_graph``).  Each cloud has its own direction, duration,
start step, time on the snapshot axis, 2-D/3-D kind, random-walk diffusivity
and seed; its particles' draws are keyed by (cloud seed, cloud-local id,
cloud-local step), so a cloud's results are the same alone and in any mix.

The pool holds ``capacity`` particle slots; a cloud's particles are appended
at the end of the used prefix (pid = cloud << 40 | local id).  A finished
cloud is collected (its particles copied out by local id) and its slots
freed; a re-sort (every ``resort_every`` steps, by grid-hash cell with the
free slots last) compacts the prefix again.

:func:`autotune` picks the workgroup size, the re-sort interval and the
number of steps per graph launch by timing candidates on the batch itself.
"""
from __future__ import annotations

import ctypes
import itertools
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._native import (CLOUD_3D, CLOUD_KICK, CLOUD_SHIFT, CLOUD_VALID, MAX_CLOUDS, BatchAxis, Cloud,
                      check)
from .device import DeviceContext, DeviceParticles, make_step_args

FREE_PID = MAX_CLOUDS << CLOUD_SHIFT
LOCAL_MASK = (1 << CLOUD_SHIFT) - 1


@dataclass
class CloudSpec:
    """One release: seeds (degrees; z in metres for 3-D clouds), the steps it
    runs, its direction, kind and random walk (``kh`` m^2/s, 0 = none), its
    RNG ``seed`` and ``real_id``, the time of its first step on the snapshot
    axis (``t0``, seconds) and the batch step it starts at."""
    lon: np.ndarray
    lat: np.ndarray
    z: Optional[np.ndarray] = None
    n_steps: int = 1
    direction: float = 1.0
    is3d: bool = True
    kh: float = 0.0
    seed: int = 0
    real_id: int = 0
    t0: float = 0.0
    start_step: int = 0

    @property
    def n(self) -> int:
        return int(np.asarray(self.lon).size)


@dataclass
class TuneChoice:
    """A batch's launch shape: threads per workgroup, re-sort interval in
    steps (0: never) and steps per graph launch (1: eager launches)."""
    block: int = 256
    resort_every: int = 8
    graph_steps: int = 8
    rate: float = 0.0                      # particle-steps/s measured for it
    table: List[dict] = field(default_factory=list)


class Batch:
    """A pool of ``capacity`` particle slots and a table of up to
    ``max_clouds`` clouds on ``ctx`` (whose mesh and snapshots are resident).

    ``axis3`` / ``axis2``: (first slot, number of snapshots) of the 3-D and
    2-D snapshot axes (cycled; one snapshot = steady), ``snap_dt`` their
    spacing.  ``step_kw``: the step-wide arguments of
    :meth:`DeviceContext.step` (order, dt, exact_nearest, w_method, beach,
    sigma, ...)."""

    def __init__(self, ctx: DeviceContext, capacity: int, axis3=(0, 0), axis2=(0, 0), snap_dt: float = 0.0,
                 max_clouds: int = 1024, **step_kw):
        if max_clouds < 1 or max_clouds > MAX_CLOUDS:
            raise ValueError(f"max_clouds must be in [1, {MAX_CLOUDS}]")
        self.ctx = ctx
        self.t = ctx.torch
        self.capacity = int(capacity)
        self.max_clouds = int(max_clouds)
        t, dev = self.t, f"cuda:{ctx.device}"
        P = self.capacity
        f64 = lambda: t.zeros(P, dtype=t.float64, device=dev)
        self.p = DeviceParticles(lon=f64(), lat=f64(), z=f64(), status=t.zeros(P, dtype=t.int32, device=dev),
                                 nn=t.full((P,), -1, dtype=t.int32, device=dev),
                                 fr=t.zeros(P, dtype=t.int8, device=dev), u=f64(), v=f64(), w=f64(),
                                 pid=t.full((P,), FREE_PID, dtype=t.int64, device=dev),
                                 ws=t.zeros(P, dtype=t.int32, device=dev), ctx=ctx)
        self.p.alive_count = t.zeros(1, dtype=t.int64, device=dev)
        self.n_used = 0
        self.n_freed = 0                    # slots freed inside the used prefix since the last sort
        self.axis = BatchAxis(int(axis3[0]), int(axis3[1]), int(axis2[0]), int(axis2[1]), float(snap_dt))
        self.specs: Dict[int, CloudSpec] = {}
        self.free_ids = list(range(self.max_clouds - 1, -1, -1))
        check(ctx.lib.kelp_batch_reset(ctx.h, self.max_clouds), "kelp_batch_reset")
        check(ctx.lib.kelp_batch_set_clouds(ctx.h, 0, 0, None, ctypes.byref(self.axis)), "kelp_batch_set_clouds")
        self.step_kw = dict(step_kw)
        self.step_index = 0
        self.choice = TuneChoice()

    # -- clouds ---------------------------------------------------------------
    def add(self, spec: CloudSpec) -> int:
        """Admit a cloud: its particles go to the end of the used prefix and
        its table entry is written; returns its id.  Raises when the pool or
        the table is full (see :meth:`room`)."""
        n = spec.n
        if n > self.capacity - self.n_used:
            raise MemoryError(f"batch pool full: {n} particles, {self.capacity - self.n_used} slots left")
        if not self.free_ids:
            raise MemoryError("batch cloud table full")
        if spec.is3d and self.axis.n_snap_3d <= 0:
            raise ValueError("a 3-D cloud needs the 3-D snapshot axis")
        if not spec.is3d and self.axis.n_snap_2d <= 0:
            raise ValueError("a 2-D cloud needs the 2-D snapshot axis")
        cid = self.free_ids.pop()
        t, dev = self.t, f"cuda:{self.ctx.device}"
        a, b = self.n_used, self.n_used + n
        self.p.lon[a:b] = t.as_tensor(np.ascontiguousarray(spec.lon, np.float64), device=dev)
        self.p.lat[a:b] = t.as_tensor(np.ascontiguousarray(spec.lat, np.float64), device=dev)
        z = np.zeros(n) if spec.z is None or not spec.is3d else np.ascontiguousarray(spec.z, np.float64)
        self.p.z[a:b] = t.as_tensor(z, device=dev)
        self.p.status[a:b] = 1
        self.p.nn[a:b] = -1
        self.p.fr[a:b] = 0
        self.p.ws[a:b] = 0
        self.p.pid[a:b] = t.arange(n, dtype=t.int64, device=dev) + (cid << CLOUD_SHIFT)
        t.cuda.current_stream().synchronize()
        self.n_used = b
        self.specs[cid] = spec
        self._write_entry(cid, spec)
        return cid

    def _write_entry(self, cid: int, spec: Optional[CloudSpec]) -> None:
        e = Cloud()
        if spec is not None:
            kick = spec.kh > 0.0
            e.flags = CLOUD_VALID | (CLOUD_3D if spec.is3d else 0) | (CLOUD_KICK if kick else 0)
            e.n_steps = int(spec.n_steps)
            e.start_step = int(spec.start_step)
            e.time_dir = float(spec.direction)
            e.t0 = float(spec.t0)
            e.kh_x = e.kh_y = float(spec.kh)
            e.seed = int(spec.seed) & ((1 << 64) - 1)
            e.real_id = int(spec.real_id)
        check(self.ctx.lib.kelp_batch_set_clouds(self.ctx.h, int(cid), 1, ctypes.byref(e), None),
              "kelp_batch_set_clouds")

    def room(self) -> int:
        """Particle slots left at the end of the pool."""
        return self.capacity - self.n_used

    def finished(self) -> List[int]:
        """Clouds whose last step has run."""
        return [c for c, s in self.specs.items() if self.step_index >= s.start_step + s.n_steps]

    def collect(self, cid: int, release: bool = True) -> Dict[str, np.ndarray]:
        """The final state of cloud ``cid`` by local id (lon, lat, z, status,
        fr); with ``release`` its slots and table entry are freed."""
        spec = self.specs[cid]
        t, dev = self.t, f"cuda:{self.ctx.device}"
        n = spec.n
        out = dict(lon=t.empty(n, dtype=t.float64, device=dev), lat=t.empty(n, dtype=t.float64, device=dev),
                   z=t.empty(n, dtype=t.float64, device=dev), status=t.empty(n, dtype=t.int32, device=dev),
                   fr=t.empty(n, dtype=t.int8, device=dev))
        t.cuda.current_stream().synchronize()
        s = self._view(self.n_used).as_struct()
        check(self.ctx.lib.kelp_batch_collect(self.ctx.h, ctypes.byref(s), int(cid), n,
                                              out["lon"].data_ptr(), out["lat"].data_ptr(), out["z"].data_ptr(),
                                              out["status"].data_ptr(), out["fr"].data_ptr(), int(release)),
              "kelp_batch_collect")
        self.ctx.synchronize()
        res = {k: v.cpu().numpy() for k, v in out.items()}
        if release:
            self.n_freed += n
            del self.specs[cid]
            self._write_entry(cid, None)
            self.free_ids.append(cid)
        return res

    # -- stepping -------------------------------------------------------------
    def _view(self, n: int) -> DeviceParticles:
        p = self.p
        return DeviceParticles(lon=p.lon[:n], lat=p.lat[:n], z=p.z[:n], status=p.status[:n], nn=p.nn[:n],
                               fr=p.fr[:n], u=p.u[:n], v=p.v[:n], w=p.w[:n], pid=p.pid[:n], ws=p.ws[:n],
                               alive_count=p.alive_count, ctx=self.ctx)

    def _base(self, block: int):
        kw = dict(self.step_kw)
        order = kw.pop("order")
        dt = kw.pop("dt")
        slot = self.axis.slot0_3d if self.axis.n_snap_3d > 0 else self.axis.slot0_2d
        kw.setdefault("reuse_anchor", True)
        a = make_step_args(order, slot, dt, 1.0, **kw)
        a.block_size = int(block)
        return a

    def sort(self) -> None:
        """Re-sort the used prefix by grid-hash cell (free slots last) and
        shrink it to the live entries."""
        if self.n_used == 0:
            return
        self.ctx.sort(self._view(self.n_used))
        self.ctx.synchronize()
        self.n_used = int((self.p.status[:self.n_used] != 0).sum().item())
        self.n_freed = 0

    def steps(self, n: int, choice: Optional[TuneChoice] = None) -> None:
        """Advance every running cloud by ``n`` batch steps (no wait)."""
        ch = choice or self.choice
        a = self._base(ch.block)
        k = 0
        while k < n:
            if ch.resort_every and self.step_index % ch.resort_every == 0:
                self.sort()
            run = n - k
            if ch.resort_every:
                run = min(run, ch.resort_every - self.step_index % ch.resort_every)
            s = self._view(self.n_used).as_struct()
            if ch.graph_steps > 1:
                while run > 0:
                    m = min(run, ch.graph_steps)
                    check(self.ctx.lib.kelp_batch_run_graph(self.ctx.h, ctypes.byref(s), ctypes.byref(a), m),
                          "kelp_batch_run_graph")
                    run -= m
                    k += m
                    self.step_index += m
            else:
                for _ in range(run):
                    check(self.ctx.lib.kelp_batch_step(self.ctx.h, ctypes.byref(s), ctypes.byref(a)),
                          "kelp_batch_step")
                k += run
                self.step_index += run

    def alive_steps(self, reset: bool = False) -> int:
        """Particle-steps of particles alive at their step start since the
        last reset (counted on the device)."""
        self.ctx.synchronize()
        v = int(self.p.alive_count.item())
        if reset:
            self.p.alive_count.zero_()
            self.t.cuda.current_stream().synchronize()
        return v

    # -- state snapshots (autotuning) ----------------------------------------
    def save(self):
        keys = ("lon", "lat", "z", "status", "nn", "fr", "u", "v", "w", "pid", "ws")
        self.t.cuda.current_stream().synchronize()
        self.ctx.synchronize()
        return ({k: getattr(self.p, k).clone() for k in keys}, self.n_used, self.step_index)

    def restore(self, saved) -> None:
        arrs, n_used, step = saved
        self.ctx.synchronize()
        for k, v in arrs.items():
            getattr(self.p, k).copy_(v)
        self.t.cuda.current_stream().synchronize()
        self.n_used = n_used
        self.step_index = step
        check(self.ctx.lib.kelp_batch_set_step(self.ctx.h, int(step)), "kelp_batch_set_step")


def autotune(batch: Batch, steps: int = 16, blocks: Sequence[int] = (64, 128, 256),
             resort: Sequence[int] = (0, 4, 8, 16), graph: Sequence[int] = (1, 8),
             warmup: int = 2) -> TuneChoice:
    """Time every (workgroup size, re-sort interval, graph length) candidate
    on ``batch`` itself for ``steps`` steps from its current state (restored
    after each) and keep the fastest as ``batch.choice``.  The choice moves
    no result: every candidate computes the same particles bit for bit."""
    saved = batch.save()
    table = []
    best = None
    for blk, rs, gs in itertools.product(blocks, resort, graph):
        ch = TuneChoice(block=blk, resort_every=rs, graph_steps=gs)
        batch.restore(saved)
        batch.steps(warmup, ch)
        batch.ctx.synchronize()
        batch.alive_steps(reset=True)
        t0 = time.perf_counter()
        batch.steps(steps, ch)
        batch.ctx.synchronize()
        el = time.perf_counter() - t0
        ch.rate = batch.alive_steps(reset=True) / el
        table.append(dict(block=blk, resort_every=rs, graph_steps=gs, rate=ch.rate))
        if best is None or ch.rate > best.rate:
            best = ch
    batch.restore(saved)
    best.table = table
    batch.choice = best
    return best


__all__ = ["CloudSpec", "Batch", "TuneChoice", "autotune", "FREE_PID", "LOCAL_MASK"]
