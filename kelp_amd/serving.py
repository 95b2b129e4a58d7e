"""Serving tracking requests against device-resident fields.

A request is one release tracked for a number of steps -- what one
``run_tracking`` call does in the reference (src/kelp/tracking.py:3288,
its step loop :2384-2833; the release itself core/release.py:112).  The
:class:`Server` keeps one or more field sets resident (a :class:`Pool` per
set: a device context holding the mesh and its snapshots, and a particle
pool, see :mod:`kelp_amd.batching`) and runs a loop of continuous batching:

* requests arrive (at their arrival times, against the wall clock) and wait
  in a queue;
* at every step boundary the queue is admitted in arrival order into the
  pool of its field set while the pool has room: particle slots, sized once
  from the device memory free at start-up, and cloud table entries; a
  request that does not fit waits, and the ones behind it that do fit are
  admitted past it;
* every pool with running requests advances all of them with one batched
  kernel per step (captured HIP graphs of ``chunk`` steps), the particle
  state never leaving the device;
* a request whose last step has run is collected (its particles copied out
  by local id, then to the host) and its slots are freed; a re-sort of the
  pool by grid-hash cell compacts the slots and keeps the gathers local.

Each request's draws are keyed by its own seed, its particles' local ids and
its own step count, so its result is the same whatever else shares the GPU.
Reported per request: arrival, admission and completion times; for the
stream: particle-steps/s over the wall time and latency percentiles.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .batching import Batch, CloudSpec, TuneChoice

#: device bytes a pool particle slot takes (state arrays + sort staging)
BYTES_PER_PARTICLE = 3 * 8 + 4 + 4 + 1 + 3 * 8 + 8 + 4 + 24


@dataclass
class Request:
    """One tracking request: seeds, steps, direction, 2-D/3-D, random walk
    (``kh`` m^2/s, 0 = none), RNG seed, the time of its first step on the
    snapshot axis, its field set and its arrival time (seconds after the
    stream starts)."""
    rid: int
    lon: np.ndarray
    lat: np.ndarray
    z: Optional[np.ndarray]
    n_steps: int
    direction: float = 1.0
    is3d: bool = True
    kh: float = 0.0
    seed: int = 0
    t0: float = 0.0
    pool: str = "default"
    arrival: float = 0.0

    @property
    def n(self) -> int:
        return int(np.asarray(self.lon).size)


@dataclass
class Result:
    rid: int
    state: Dict[str, np.ndarray]
    arrival: float
    admitted: float
    done: float
    particle_steps: int

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class Pool:
    """One resident field set: ``ctx`` holds the mesh and the snapshots of
    the axes; ``capacity`` particle slots (default: ``hbm_fraction`` of the
    free device memory, at most ``max_capacity``)."""

    def __init__(self, ctx, axis3=(0, 0), axis2=(0, 0), snap_dt=0.0, capacity: int = 0,
                 max_capacity: int = 8_000_000, hbm_fraction: float = 0.5, max_clouds: int = 4096,
                 choice: Optional[TuneChoice] = None, **step_kw):
        if capacity <= 0:
            free, _ = ctx.torch.cuda.mem_get_info(ctx.device)
            capacity = min(int(free * hbm_fraction) // BYTES_PER_PARTICLE, max_capacity)
        self.batch = Batch(ctx, capacity, axis3=axis3, axis2=axis2, snap_dt=snap_dt, max_clouds=max_clouds,
                           **step_kw)
        self.choice = choice or TuneChoice(256, 8, 8)
        self.live: Dict[int, tuple] = {}

    def fits(self, req: Request) -> bool:
        b = self.batch
        return req.n <= b.room() and len(b.free_ids) > 0

    def could_fit_after_compaction(self, req: Request) -> bool:
        b = self.batch
        return req.n <= b.room() + b.n_freed and len(b.free_ids) > 0


class Server:
    """Continuous batching over the pools (``{name: Pool}``)."""

    def __init__(self, pools: Dict[str, Pool], chunk: int = 4):
        self.pools = pools
        self.chunk = int(chunk)

    def run(self, requests: List[Request], realtime: bool = True, log=None,
            progress_s: float = 0.0) -> List[Result]:
        """Serve ``requests`` (arriving at their ``arrival`` times when
        ``realtime``, else all at once) until every one is done; with
        ``log`` and ``progress_s`` a status line every ``progress_s``
        seconds."""
        arrivals = deque(sorted(requests, key=lambda r: r.arrival))
        queue: List[Request] = []
        results: List[Result] = []
        t_start = time.perf_counter()
        steps_done = 0
        for pl in self.pools.values():
            pl.batch.alive_steps(reset=True)
        next_report = progress_s
        while arrivals or queue or any(pl.live for pl in self.pools.values()):
            now = time.perf_counter() - t_start
            if log and progress_s > 0 and now >= next_report:
                next_report += progress_s
                live = sum(len(pl.live) for pl in self.pools.values())
                used = sum(pl.batch.n_used for pl in self.pools.values())
                log(f"serving t={now:.0f} s: {len(results)} done, {live} running ({used} slots used), "
                    f"{len(queue)} waiting, {len(arrivals)} still to arrive")
            while arrivals and (not realtime or arrivals[0].arrival <= now):
                queue.append(arrivals.popleft())
            # admission: arrival order, smaller requests may pass a waiting one
            waiting = []
            for req in queue:
                pl = self.pools[req.pool]
                if not pl.fits(req) and pl.batch.n_freed and pl.could_fit_after_compaction(req):
                    pl.batch.sort()
                if pl.fits(req):
                    b = pl.batch
                    spec = CloudSpec(lon=req.lon, lat=req.lat, z=req.z, n_steps=req.n_steps,
                                     direction=req.direction, is3d=req.is3d, kh=req.kh, seed=req.seed,
                                     t0=req.t0, start_step=b.step_index)
                    cid = b.add(spec)
                    pl.live[cid] = (req, time.perf_counter() - t_start)
                else:
                    waiting.append(req)
            queue = waiting
            running = [pl for pl in self.pools.values() if pl.live]
            if not running:
                if arrivals and realtime:
                    time.sleep(max(0.0, min(arrivals[0].arrival - (time.perf_counter() - t_start), 0.05)))
                continue
            # one chunk of steps, up to the next completion in each pool
            for pl in running:
                b = pl.batch
                nxt = min(b.specs[c].start_step + b.specs[c].n_steps for c in pl.live) - b.step_index
                b.steps(max(1, min(self.chunk, nxt)), pl.choice)
            steps_done += 1
            for pl in running:
                b = pl.batch
                for cid in b.finished():
                    req, t_adm = pl.live.pop(cid)
                    st = b.collect(cid)
                    results.append(Result(req.rid, st, req.arrival, t_adm, time.perf_counter() - t_start,
                                          req.n * req.n_steps))
        self.wall = time.perf_counter() - t_start
        self.alive_steps = sum(pl.batch.alive_steps() for pl in self.pools.values())
        return results


def latency_summary(results: List[Result]) -> Dict[str, float]:
    lat = np.array([r.latency for r in results])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()), "n": int(lat.size)}


def poisson_stream(rate_hz: float, duration_s: float, seed: int, make_request) -> List[Request]:
    """Requests arriving as a Poisson process of ``rate_hz`` over
    ``duration_s`` seconds; ``make_request(rid, rng, arrival)`` builds each."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t >= duration_s:
            break
        out.append(make_request(len(out), rng, t))
    return out


__all__ = ["Request", "Result", "Pool", "Server", "latency_summary", "poisson_stream", "BYTES_PER_PARTICLE"]
