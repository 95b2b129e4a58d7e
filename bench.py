"""Benchmark of the particle time-stepping hot path on one MI355X.

``python bench.py [--steps K] [--warmup W] [--workload NAME]`` prints one JSON
line: whole-job particle-steps/s (particles alive at each step's start,
counted on the device, over the wall time of the K timed steps) with the
mesh, the snapshots and the particle state already resident in HBM; the
roofline of the step kernel (algorithmic bytes per launch over its average
duration, HIP events on the stream the kernel runs on); and a CPU baseline
(the oracle, one host core, on a bounded sample of the same workload).

Workloads:
  global_3d    (default) 1M-node Fibonacci-sphere Delaunay mesh with a land
               mask, 47 z-layers, 8 six-hourly float32 snapshots, 1M particles,
               RK4, dt 1800 s, linear time blend, Kh = 10 m^2/s random walk,
               beaching (on_failure + stochastic coastal), T and S sampled
               (leastsq) at every step's end, as every step is an output.
  double_gyre  2-D jittered 0.25-degree lattice, steady double gyre, 10,000
               particles, RK2, dt 600 s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(path):
    """HBM bytes per step-kernel launch measured by rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, each scaled by a copy of known size) and recorded in the
    workload's profile summary (line ``traffic_bytes_per_launch: N``); None
    when there is no such record."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


class Global3D:
    """RK4 3-D with time blend, random walk and beaching.

    Algorithmic bytes per particle-step (float32 fields, 2 snapshots blended,
    triangles, barycentric W): state 89 B + 5 evaluation sites x 88 B (anchor
    row 28 B + 3 vertices x 20 B of ids/coordinates) + 4 stages x 104 B (u, v
    at 3 vertices x 2 snapshots x 4 B = 48 B, W at 3 vertices x 2 interfaces x
    2 snapshots x 4 B = 48 B, 8 B node depth) + the kick's anchor and element
    96 B + beaching 8 B = 1049 B.  With bfloat16 snapshots (``--storage
    bf16``) the field values are 2 B: 4 stages x 56 B, 857 B in all.
    """
    name = "global_3d_rk4_kh_beach"
    b_alg = 1049.0
    # rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload's step kernel
    profile = os.path.join(ROOT, "profiles", "step_kernel_rk4_3d_global_1M.md")

    def __init__(self, args):
        from kelp_amd import synth
        self.synth = synth
        t = time.time()
        self.w = synth.global_3d_workload(n_nodes=args.nodes or 1_000_000,
                                          n_particles=args.particles or 1_000_000)
        log(f"workload built in {time.time() - t:.1f} s: {self.w.mesh.n_nodes} nodes, "
            f"{self.w.mesh.n_elements} elements, {self.w.lon.size} particles")
        self.kick = dict(kh_x=self.w.kh, kh_y=self.w.kh, seed=20240501)
        self.beach = dict(criteria=["on_failure", "stochastic"], t_beach=86400.0,
                          depth_min=0.3, seed=20240501)
        self.sample_ts = not args.no_sample
        self.storage = getattr(args, "storage", "f32")
        if self.storage == "bf16":
            self.b_alg = 857.0
            self.profile = None         # the profiled traffic is the float32 workload's
        # snapshot ring: `ring` slots, the next snapshot refilled
        # asynchronously while steps run (or synchronously at the bracket
        # change with --ring-sync); all snapshots resident when 0
        self.ring = int(args.ring)
        self.ring_sync = bool(args.ring_sync)
        if self.ring:
            if self.ring < 3:
                raise SystemExit("--ring needs at least 3 slots (bracket pair + the one refilling)")
            self.sample_ts = False
        self.order = 4
        self.config = {"workload": self.name, "mesh_nodes": self.w.mesh.n_nodes,
                       "mesh_elements": self.w.mesh.n_elements, "z_levels": int(self.w.mesh.sigma.size),
                       "particles": int(self.w.lon.size), "integrator": "rk4", "dt_s": self.w.dt,
                       "snapshots": len(self.w.snaps), "snapshot_dt_s": self.w.snap_dt,
                       "field_dtype": self.storage, "kh_m2_s": self.w.kh, "w_interp": "barycentric",
                       "beaching": self.beach["criteria"], "steps_per_run": self.w.n_steps,
                       "sampled": ["T", "S"] if self.sample_ts else []}
        if self.ring:
            self.config.update(ring_slots=self.ring,
                               ring_refill="sync" if self.ring_sync else "async prefetch")

    def upload(self, ctx):
        ctx.upload_mesh(self.w.mesh)
        if self.ring:
            # the host side of the ring: every snapshot page-locked, as a
            # reader decoding into pinned buffers would leave it
            self.pin = [(ctx.pinned(U), ctx.pinned(V), ctx.pinned(W)) for U, V, W, _, _ in self.w.snaps]
            self.resident = {}
            for e in range(self.ring):
                self._ensure(ctx, e, sync=True)
            return ctx.particles(self.w.lon, self.w.lat, self.w.z)
        for s, (U, V, W, T, S) in enumerate(self.w.snaps):
            ctx.upload_field(s, U, V, W, storage="bf16" if self.storage == "bf16" else None)
            if self.sample_ts:
                ctx.upload_scalar(s, 0, T)
                ctx.upload_scalar(s, 1, S)
        p = ctx.particles(self.w.lon, self.w.lat, self.w.z)
        if self.sample_ts:
            t = ctx.torch
            self.ts_out = t.empty((2, p.n), dtype=t.float64, device=p.lon.device)
        return p

    def _ensure(self, ctx, e, sync):
        """Ring slot e % ring holds the snapshot of epoch e (unwrapped
        snapshot count), refilled now if it does not."""
        s = e % self.ring
        if self.resident.get(s) == e:
            return
        U, V, W = self.pin[e % len(self.w.snaps)]
        st = "bf16" if self.storage == "bf16" else None
        if sync:
            ctx.upload_field(s, U, V, W, storage=st)
        else:
            ctx.upload_field_async(s, U, V, W, storage=st)
        self.resident[s] = e

    def step(self, ctx, p, k, ev=None):
        w = self.w
        t = k * w.dt
        ia, ib, a0, it, dtn = self.synth.snapshot_bracket(t, w.snap_dt, len(w.snaps))
        al = [c * w.dt / dtn for c in (0.0, 0.5, 0.5, 1.0)]
        if self.ring:
            ep = int(np.floor(t / w.snap_dt))
            self._ensure(ctx, ep, sync=True)
            self._ensure(ctx, ep + 1, sync=True)
            ia = ep % self.ring
            it = (ep + 1) % self.ring
            ib = it if ib >= 0 else -1
        if ev is not None:
            ctx.event_record(ev)
        ctx.step(p, 4, ia, w.dt, 1.0, slot_b=ib, slot_t=it, alpha0=a0, alphas=al,
                 exact_nearest=1, w_method=2, kick=self.kick, beach=self.beach, step_i=k,
                 reuse_anchor=True)
        if ev is not None:
            ctx.event_record(ev + 1)
        if self.ring and not self.ring_sync:
            # the snapshot after the bracket streams in behind this step
            self._ensure(ctx, int(np.floor(t / w.snap_dt)) + 2, sync=False)
        if self.sample_ts:
            # T and S at the step's end (the record instant)
            ra, rb, ra0, _, _ = self.synth.snapshot_bracket(t + w.dt, w.snap_dt, len(w.snaps))
            ctx.sample_vars(p, (0, 1), ra, rb, ra0, "leastsq", out=self.ts_out)

    def cpu_baseline(self, budget_s):
        from oracle.oracle import OracleMesh, step_args
        w = self.w
        om = OracleMesh(w.mesh, "glibc")
        n = min(w.lon.size, 20_000)
        st = dict(lon=w.lon[:n], lat=w.lat[:n], z=w.z[:n], status=np.ones(n, np.int32),
                  nn=np.full(n, -1, np.int32))
        steps, alive_steps = 0, 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s and steps < w.n_steps:
            t = steps * w.dt
            ia, ib, a0, it, dtn = self.synth.snapshot_bracket(t, w.snap_dt, len(w.snaps))
            A, T = w.snaps[ia], w.snaps[it]
            f = dict(u_a=A[0], v_a=A[1], w_a=A[2], u_t=T[0], v_t=T[1], w_t=T[2], alpha0=a0)
            if ib >= 0:
                B = w.snaps[ib]
                f.update(u_b=B[0], v_b=B[1], w_b=B[2])
            al = [c * w.dt / dtn for c in (0.0, 0.5, 0.5, 1.0)]
            alive_steps += int(np.sum(st["status"] == 1))
            st = om.step(st, f, step_args(4, w.dt, 1.0, is3d=True, w_method=2, exact=1, alphas=al,
                                          kick=self.kick, beach=self.beach, step_i=steps))
            if self.sample_ts:
                ra, rb, ra0, _, _ = self.synth.snapshot_bracket(t + w.dt, w.snap_dt, len(w.snaps))
                for var in (3, 4):
                    om.sample(w.snaps[ra][var], st["lon"], st["lat"], st["z"], st["status"], st["nn"],
                              values_b=None if rb < 0 else w.snaps[rb][var], alpha=ra0)
            steps += 1
        el = time.perf_counter() - t0
        return {"value": alive_steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                "sample": f"{n} particles x {steps} steps of the same workload (incl. T/S "
                          f"sampling), C oracle (glibc libm), 1 thread, {el:.1f} s"}


class DoubleGyre:
    """2-D RK2, steady f32 double gyre.  Bytes per particle-step: state 65 B +
    3 sites x 88 B + 2 stages x 24 B (u, v at 3 vertices, f32) = 377 B."""
    name = "double_gyre_2d_rk2"
    b_alg = 377.0

    def __init__(self, args):
        from kelp_amd import synth
        self.w = synth.double_gyre_workload(n_particles=args.particles or 10_000)
        self.order = 2
        self.config = {"workload": self.name, "mesh_nodes": self.w.mesh.n_nodes,
                       "mesh_elements": self.w.mesh.n_elements, "particles": int(self.w.lon.size),
                       "integrator": "rk2", "dt_s": self.w.dt, "field_dtype": "f32",
                       "steps_per_run": self.w.n_steps}

    def upload(self, ctx):
        ctx.upload_mesh(self.w.mesh)
        ctx.upload_field(0, self.w.u, self.w.v)
        return ctx.particles(self.w.lon, self.w.lat)

    def step(self, ctx, p, k, ev=None):
        if ev is not None:
            ctx.event_record(ev)
        ctx.step(p, 2, 0, self.w.dt, 1.0, exact_nearest=1, reuse_anchor=True)
        if ev is not None:
            ctx.event_record(ev + 1)

    def cpu_baseline(self, budget_s):
        from oracle.oracle import OracleMesh
        w = self.w
        om = OracleMesh(w.mesh, "glibc")
        n = min(w.lon.size, 10_000)
        st = dict(lon=w.lon[:n].copy(), lat=w.lat[:n].copy(), status=np.ones(n, np.int32),
                  nn=np.full(n, -1, np.int32))
        steps, alive_steps = 0, 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            alive_steps += int(np.sum(st["status"] == 1))
            st = om.step2d(st, 2, w.u, w.v, w.dt, 1.0)
            steps += 1
        el = time.perf_counter() - t0
        return {"value": alive_steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                "sample": f"{n} particles x {steps} steps, C oracle (glibc libm), 1 thread, {el:.1f} s"}


WORKLOADS = {"global_3d": Global3D, "double_gyre": DoubleGyre}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--workload", default="global_3d", choices=sorted(WORKLOADS))
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--nodes", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--resort-every", type=int, default=8,
                    help="re-sort particles by grid-hash cell every N steps (0: never)")
    ap.add_argument("--ring", type=int, default=0,
                    help="global_3d: keep only this many snapshot slots (>= 3) and stream the "
                         "snapshots through them (no T/S sampling)")
    ap.add_argument("--ring-sync", action="store_true",
                    help="with --ring: refill at the bracket change, synchronously")
    ap.add_argument("--storage", default="f32", choices=["f32", "bf16"],
                    help="snapshot storage of the 3-D workload (bf16: rounded on upload)")
    ap.add_argument("--no-sample", action="store_true",
                    help="global_3d: leave out the T/S sampling at each step's end")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy a buffer of this size 3x (calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    from kelp_amd.device import DeviceContext
    wl = WORKLOADS[args.workload](args)
    ctx = DeviceContext(0)
    torch = ctx.torch
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written per copy")
        del src, dst
    p = wl.upload(ctx)
    p.alive_count = torch.zeros(1, dtype=torch.int64, device="cuda:0")
    log("uploaded; warming up")

    def run_step(k, ev=None):
        if args.resort_every > 0 and k % args.resort_every == 0:
            ctx.sort(p)
        wl.step(ctx, p, k, ev)

    if args.steps > 511:
        raise SystemExit("--steps is limited to 511 (one event pair per timed step)")
    for k in range(args.warmup):
        run_step(k)
    ctx.synchronize()
    torch.cuda.synchronize()
    p.alive_count.zero_()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    for j, k in enumerate(range(args.warmup, args.warmup + args.steps)):
        run_step(k, ev=2 + 2 * j)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the step kernel's own launches: HIP events on the context stream
    # around each of them
    kern_ms = sum(ctx.event_elapsed_ms(2 + 2 * j, 3 + 2 * j) for j in range(args.steps))
    alive_steps = int(p.alive_count.item())
    alive_end = int((p.status == 1).sum().item())
    pss = alive_steps / wall
    avg_launch_s = kern_ms / 1e3 / args.steps
    per_launch_units = alive_steps / args.steps
    achieved = wl.b_alg * per_launch_units / avg_launch_s / 1e9
    # measured HBM bytes per launch, valid for the default workload size only
    traffic = None
    if getattr(wl, "profile", None) and not (args.particles or args.nodes):
        traffic = profiled_traffic(wl.profile)
    line = {
        "metric": "particle-steps/s", "value": pss, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(wl.config, alive_end=alive_end, resort_every=args.resort_every),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "step_kernel", "kernel_avg_us": avg_launch_s * 1e6,
                     "b_alg_per_particle_step": wl.b_alg,
                     "particle_steps_per_launch": per_launch_units},
    }
    if not args.no_cpu:
        log("CPU baseline")
        line["cpu_baseline"] = wl.cpu_baseline(args.cpu_budget)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
