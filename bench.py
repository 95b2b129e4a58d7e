"""Benchmark of the particle time-stepping hot path on one MI355X.

``python bench.py [--steps K] [--warmup W] [--workload NAME]`` prints one JSON
line: whole-job particle-steps/s (particles alive at each step's start,
counted on the device, over the wall time of the K timed steps) with the
mesh, the snapshots and the particle state already resident in HBM; the
roofline of the step kernel (algorithmic bytes per launch over its average
duration, HIP events on the stream the kernel runs on); and a CPU baseline
(the C oracle with OpenMP over particles, on every host thread the process
may use and on one, on a bounded sample of the same workload).

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


def host_threads():
    """Host threads the CPU baseline may use: OMP_NUM_THREADS when set (the
    process's CPU share), else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_legs(make_oracle, run, n_total, n_per_thread, budget_s, what):
    """The C oracle (glibc libm, OpenMP over particles) timed on all host
    threads and on one: ``run(oracle, n, budget) -> (particle-steps/s,
    steps)`` over the first ``n`` particles.  The all-thread leg takes
    about two thirds of the budget and ``n_per_thread`` particles per
    thread (up to the workload's count)."""
    from oracle.oracle import set_threads
    om = make_oracle()
    threads = set_threads(host_threads())
    n_all = min(n_total, n_per_thread * threads)
    v_all, steps_all = run(om, n_all, budget_s * 2.0 / 3.0)
    set_threads(1)
    n_one = min(n_total, n_per_thread)
    v_one, steps_one = run(om, n_one, budget_s / 3.0)
    return {"value": v_all, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "value_1_thread": v_one,
            "sample": f"{n_all} particles x {steps_all} steps {what} on {threads} threads; "
                      f"{n_one} particles x {steps_one} steps on 1 thread; C oracle (glibc libm), "
                      f"OpenMP over particles"}


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
        self.cell_rule = args.cell_rule
        self.storage = getattr(args, "storage", "f32")
        self.field_dtype = self.storage
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
        self.cell_report = apply_cell_rule(ctx, self.w.mesh, self.cell_rule)
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

        def run(om, n, budget):
            st = dict(lon=w.lon[:n].copy(), lat=w.lat[:n].copy(), z=w.z[:n].copy(),
                      status=np.ones(n, np.int32), nn=np.full(n, -1, np.int32))
            steps, alive_steps = 0, 0
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < budget and steps < w.n_steps:
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
                        om.sample(w.snaps[ra][var], st["lon"], st["lat"], st["z"], st["status"],
                                  st["nn"], values_b=None if rb < 0 else w.snaps[rb][var], alpha=ra0)
                steps += 1
            return alive_steps / (time.perf_counter() - t0), steps

        return cpu_legs(lambda: OracleMesh(w.mesh, "glibc_omp"), run, w.lon.size, 20_000, budget_s,
                        "of the same workload (incl. T/S sampling)")


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

        def run(om, n, budget):
            st = dict(lon=w.lon[:n].copy(), lat=w.lat[:n].copy(), status=np.ones(n, np.int32),
                      nn=np.full(n, -1, np.int32))
            steps, alive_steps = 0, 0
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < budget:
                alive_steps += int(np.sum(st["status"] == 1))
                st = om.step2d(st, 2, w.u, w.v, w.dt, 1.0)
                steps += 1
            return alive_steps / (time.perf_counter() - t0), steps

        return cpu_legs(lambda: OracleMesh(w.mesh, "glibc_omp"), run, w.lon.size, 10_000, budget_s,
                        "of the same workload")


def apply_cell_rule(ctx, mesh, rule):
    """The grid-hash cell of ``mesh`` by ``rule`` as run_tracking chooses it
    (occupancy, the reference's default, or legacy: the automatic cell),
    the mesh re-uploaded when the cell changes; returns the cell used."""
    from kelp_amd.cellrule import apply_grid_hash_cell_rule
    before = float(mesh.gh_cell_size)
    apply_grid_hash_cell_rule(mesh, ctx, rule)
    if float(mesh.gh_cell_size) != before:
        ctx.upload_mesh(mesh)
    log(f"grid-hash cell ({rule}): {mesh.gh_cell_size:.4f} deg (automatic {before:.4f})")
    return {"rule": rule, "cell_deg": float(mesh.gh_cell_size), "automatic_cell_deg": before}


WORKLOADS = {"global_3d": Global3D, "double_gyre": DoubleGyre}


def _alone_rate(ctx, specs, axis3, axis2, snap_dt, order, dt, budget_s, sigma=None, beach=None, ctx2d=None):
    """The plain path on the same clouds, one cloud at a time (each step one
    kelp_step launch over that cloud), for up to ``budget_s`` seconds:
    particle-steps/s (alive at the step start) and how much was run.  2-D
    clouds run on ``ctx2d`` when given (a sigma mesh's plain step takes 3-D
    fields only)."""
    from kelp_amd import synth
    C = {1: (0.0,), 2: (0.0, 0.5), 4: (0.0, 0.5, 0.5, 1.0)}
    done_steps, n_clouds = 0, 0
    t0 = time.perf_counter()
    for spec in specs:
        if time.perf_counter() - t0 > budget_s:
            break
        cx = ctx2d if (ctx2d is not None and not spec.is3d) else ctx
        t = cx.torch
        p = cx.particles(spec.lon, spec.lat, spec.z if spec.is3d else None, pid=np.arange(spec.n))
        p.alive_count = t.zeros(1, dtype=t.int64, device=p.lon.device)
        slot0, n = axis3 if spec.is3d else axis2
        kick = dict(kh_x=spec.kh, kh_y=spec.kh, seed=spec.seed) if spec.kh > 0 else None
        for k in range(spec.n_steps):
            if n <= 1:
                sa, sb, st, a0, al = slot0, -1, -1, 0.0, [0.0] * 4
            else:
                tt = spec.t0 + spec.direction * (k * dt)
                ia, ib, a0, it, dtn = synth.snapshot_bracket(tt, snap_dt, n, spec.direction)
                sa, sb = slot0 + ia, (slot0 + ib if ib >= 0 else -1)
                st, al = ((-1, [0.0] * 4) if dt > dtn else
                          (slot0 + it, [c * dt / dtn for c in C[order]] + [0.0] * (4 - order)))
            cx.step(p, order, sa, dt, spec.direction, slot_b=sb, slot_t=st, alpha0=a0, alphas=al,
                    exact_nearest=1, w_method=2, kick=kick, step_i=k, reuse_anchor=True,
                    beach=None if beach is None else dict(beach, seed=spec.seed),
                    sigma=sigma if spec.is3d else None)
        cx.synchronize()
        done_steps += int(p.alive_count.item())
        n_clouds += 1
    el = time.perf_counter() - t0
    return done_steps / el, n_clouds, el


def run_batch(args):
    """The batching workload: 256 clouds on the mixed triangle/quad lattice
    with 20 sigma layers, all advanced together until the longest is done."""
    from kelp_amd import synth
    from kelp_amd.batching import Batch, CloudSpec, autotune
    from kelp_amd.device import DeviceContext
    t = time.time()
    mesh = synth.mixed_lattice_mesh()
    u, v, w, eta, u2, v2 = synth.mixed_lattice_fields(mesh)
    clouds = synth.batch_clouds(mesh, n_clouds=args.clouds)
    specs = [CloudSpec(**c) for c in clouds]
    P = sum(s.n for s in specs)
    log(f"batch workload built in {time.time() - t:.1f} s: {mesh.n_nodes} nodes, {mesh.n_elements} "
        f"elements, {len(specs)} clouds, {P} particles")
    ctx = DeviceContext(0)
    torch = ctx.torch
    ctx.upload_mesh(mesh)
    ctx.upload_field(0, u, v, w)
    ctx.upload_eta(0, eta)
    ctx.upload_field(1, u2, v2)
    sigma = dict(start=(0, -1, -1, 0.0, 0.0), end=None, Dmin=0.5, z0b_min=0.1)
    order, dt = 2, synth.DAY
    kw = dict(order=order, dt=dt, exact_nearest=1, w_method=2, sigma=sigma)
    b = Batch(ctx, P, axis3=(0, 1), axis2=(1, 1), max_clouds=len(specs), **kw)
    for s_ in specs:
        b.add(s_)
    choice = autotune(b, steps=8, warmup=2, blocks=(64, 128, 256), resort=(0, 4, 8, 16), graph=(1, 8))
    log(f"autotuner: block {choice.block}, re-sort every {choice.resort_every}, graph of "
        f"{choice.graph_steps} steps ({choice.rate / 1e6:.1f} M particle-steps/s)")
    n_total = max(s_.n_steps for s_ in specs)
    b.alive_steps(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.steps(n_total)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    alive = b.alive_steps()
    pss = alive / wall
    log("plain path, cloud by cloud, for comparison")
    import copy
    m2 = copy.copy(mesh)            # the horizontal mesh as a z-level mesh for the 2-D clouds
    m2.sigma = np.array([], dtype=np.float64)
    m2.vertical_coord = "zlevel"
    ctx2 = DeviceContext(0)
    ctx2.upload_mesh(m2)
    ctx2.upload_field(1, u2, v2)
    plain, n_pl, el_pl = _alone_rate(ctx, specs, (0, 1), (1, 1), 0.0, order, dt, args.cpu_budget, sigma=sigma,
                                     ctx2d=ctx2)
    ctx2.close()
    line = {
        "metric": "particle-steps/s", "value": pss, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": n_total, "warmup": 0, "ms_per_step": wall / n_total * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "field_dtype": "f32", "data": "synthetic",
        "config": {"workload": "batch_256_clouds_mixed_lattice_sigma20", "mesh_nodes": mesh.n_nodes,
                   "mesh_elements": mesh.n_elements, "quads": int((mesh.elem_nodes[:, 3] >= 0).sum()),
                   "sigma_layers": int(mesh.sigma.size - 1), "clouds": len(specs), "particles": P,
                   "integrator": "rk2", "dt_s": dt, "fields": "steady",
                   "clouds_3d": sum(s_.is3d for s_ in specs), "clouds_backward": sum(s_.direction < 0 for s_ in specs),
                   "clouds_kh": sum(s_.kh > 0 for s_ in specs), "alive_particle_steps": alive,
                   "autotune": {"block": choice.block, "resort_every": choice.resort_every,
                                "graph_steps": choice.graph_steps, "candidates": choice.table}},
        "plain_path": {"value": plain, "unit": "particle-steps/s", "clouds": n_pl, "seconds": el_pl,
                       "what": "the same clouds one at a time through kelp_step"},
        "speedup_vs_plain": pss / plain if plain > 0 else None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def _gc_m(lon1, lat1, lon2, lat2):
    """Great-circle distance (m) on the 6371 km sphere (haversine)."""
    p1, p2 = np.radians(lat1), np.radians(lat2)
    dl = np.radians(lon2 - lon1)
    h = np.sin((p2 - p1) / 2) ** 2 + np.cos(p1) * np.cos(p2) * np.sin(dl / 2) ** 2
    return 2 * 6371e3 * np.arcsin(np.sqrt(np.clip(h, 0.0, 1.0)))


def run_precision(args):
    """The precision workload: the HEALPix NEST dual mesh at ``--nside`` with
    69 z-levels over the wet pixels of the 3-D workload's harmonic mask, u/v/W
    from the 3-D workload's generators (four six-hourly snapshots, cycled),
    RK2 with dt 6 h and the Kh = 10 m^2/s random walk, forward for a year.
    Every storage dtype -- float64, float32, bfloat16, float16 with
    per-level scales -- runs the same particles; the error of each against
    the float64-field run (great-circle and depth: median and p99 at 30, 90
    and 365 days) and each one's plain-path throughput on ``--rate-particles``
    particles (24 timed steps) are reported."""
    from kelp_amd import synth
    from kelp_amd.device import DeviceContext
    t = time.time()
    mesh = synth.healpix_c3_mesh(args.nside)
    log(f"HEALPix nside {args.nside} dual mesh in {time.time() - t:.1f} s: {mesh.n_nodes} nodes, "
        f"{mesh.n_elements} elements, {mesh.sigma.size} levels")
    ctx = DeviceContext(0)
    torch = ctx.torch
    ctx.upload_mesh(mesh)
    n_snap, dt, kh = 4, 21600.0, 10.0
    t = time.time()
    snaps64 = [synth.jet_rossby_snapshot_torch(mesh, k * synth.SNAP_DT, torch, "cuda:0", np.float64)
               for k in range(n_snap)]
    torch.cuda.empty_cache()
    log(f"fields generated in {time.time() - t:.1f} s")
    n_err = args.particles or 200_000
    n_rate = args.rate_particles
    lon, lat, z = synth.wet_area_seeds(mesh, max(n_err, n_rate), zmax=1000.0, seed=synth.C2_SEED + 3)
    checkpoints = {30: 120, 90: 360, 365: 1460}
    n_steps = max(checkpoints.values())
    kick = dict(kh_x=kh, kh_y=kh, seed=20241231)
    b_alg = {"f64": 1025.0, "f32": 769.0, "bf16": 641.0, "f16s": 641.0}
    runs, rates = {}, {}

    def bracket(k):
        ia, ib, a0, it, dtn = synth.snapshot_bracket(k * dt, synth.SNAP_DT, n_snap)
        al = [c * dt / dtn for c in (0.0, 0.5)] + [0.0, 0.0]
        return ia, ib, a0, it, al

    for storage in ("f64", "f32", "bf16", "f16s"):
        t = time.time()
        for k, (U, V, W) in enumerate(snaps64):
            if storage == "f64":
                ctx.upload_field(k, U, V, W)
            else:
                ctx.upload_field(k, U.astype(np.float32), V.astype(np.float32), W.astype(np.float32),
                                 storage=None if storage == "f32" else storage)
        log(f"{storage}: snapshots uploaded in {time.time() - t:.1f} s")
        # the error run
        p = ctx.particles(lon[:n_err], lat[:n_err], z[:n_err], pid=np.arange(n_err))
        rec = {}
        t = time.time()
        for k in range(n_steps):
            if args.resort_every and k % args.resort_every == 0:
                ctx.sort(p)
            ia, ib, a0, it, al = bracket(k)
            ctx.step(p, 2, ia, dt, 1.0, slot_b=ib, slot_t=it, alpha0=a0, alphas=al, exact_nearest=1,
                     w_method=2, kick=kick, step_i=k, reuse_anchor=True)
            for day, st in checkpoints.items():
                if k + 1 == st:
                    h = p.to_host()
                    o = np.argsort(h["pid"], kind="stable")
                    rec[day] = {key: h[key][o] for key in ("lon", "lat", "z", "status")}
            if (k + 1) % 365 == 0:
                log(f"{storage}: {k + 1} of {n_steps} steps ({time.time() - t:.1f} s)")
        runs[storage] = rec
        del p
        # the throughput run
        p = ctx.particles(lon[:n_rate], lat[:n_rate], z[:n_rate], pid=np.arange(n_rate))
        p.alive_count = torch.zeros(1, dtype=torch.int64, device="cuda:0")
        n_t = 24
        for k in range(n_t + 2):
            if k == 2:
                ctx.synchronize()
                p.alive_count.zero_()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            if args.resort_every and k % args.resort_every == 0:
                ctx.sort(p)
            ia, ib, a0, it, al = bracket(k)
            if k >= 2:
                ctx.event_record(2 * (k - 2) + 2)
            ctx.step(p, 2, ia, dt, 1.0, slot_b=ib, slot_t=it, alpha0=a0, alphas=al, exact_nearest=1,
                     w_method=2, kick=kick, step_i=k, reuse_anchor=True)
            if k >= 2:
                ctx.event_record(2 * (k - 2) + 3)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        kern = sum(ctx.event_elapsed_ms(2 * j + 2, 2 * j + 3) for j in range(n_t)) / n_t / 1e3
        alive = int(p.alive_count.item())
        ach = b_alg[storage] * alive / n_t / kern / 1e9
        rates[storage] = {"particle_steps_per_s": alive / wall, "ms_per_step": wall / n_t * 1e3,
                          "kernel_avg_us": kern * 1e6, "b_alg_per_particle_step": b_alg[storage],
                          "achieved_GBs": ach, "frac_of_hbm_peak": ach / HBM_PEAK_GBS}
        log(f"{storage}: {alive / wall / 1e6:.1f} M particle-steps/s on {n_rate} particles, step kernel "
            f"{kern * 1e6:.0f} us ({ach:.0f} GB/s)")
        del p
        torch.cuda.empty_cache()
    errors = {}
    ref = runs["f64"]
    for storage in ("f32", "bf16", "f16s"):
        errors[storage] = {}
        for day in checkpoints:
            a, b = runs[storage][day], ref[day]
            ok = (a["status"] == 1) & (b["status"] == 1)
            d = _gc_m(a["lon"][ok], a["lat"][ok], b["lon"][ok], b["lat"][ok])
            dz = np.abs(a["z"][ok] - b["z"][ok])
            errors[storage][f"{day}d"] = {
                "median_m": float(np.median(d)), "p99_m": float(np.percentile(d, 99)),
                "depth_median_m": float(np.median(dz)), "depth_p99_m": float(np.percentile(dz, 99)),
                "alive_both": int(ok.sum()),
                "status_differs": int((a["status"] != b["status"]).sum())}
    f32 = rates["f32"]
    line = {
        "metric": "particle-steps/s", "value": f32["particle_steps_per_s"], "unit": "particle-steps/s",
        "n_gpus": 1, "steps": 24, "warmup": 2, "ms_per_step": f32["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "field_dtype": "f32", "data": "synthetic",
        "config": {"workload": f"precision_healpix_nest{args.nside}_69lev", "mesh_nodes": mesh.n_nodes,
                   "mesh_elements": mesh.n_elements, "z_levels": int(mesh.sigma.size), "snapshots": n_snap,
                   "integrator": "rk2", "dt_s": dt, "kh_m2_s": kh, "particles_rate": n_rate,
                   "particles_error": n_err, "steps_error_run": n_steps, "resort_every": args.resort_every},
        "roofline": {"bound": "hbm", "achieved": f32["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": f32["frac_of_hbm_peak"], "traffic": None, "kernel": "step_kernel",
                     "kernel_avg_us": f32["kernel_avg_us"], "b_alg_per_particle_step": f32["b_alg_per_particle_step"]},
        "by_storage": rates,
        "error_vs_f64": errors,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def run_serving(args):
    """The serving workload on the 3-D global mesh's fields (resident, cycled
    six-hourly snapshots): Poisson request arrivals against the wall clock,
    continuous batching, latency percentiles."""
    from kelp_amd import synth
    from kelp_amd.batching import CloudSpec, TuneChoice
    from kelp_amd.device import DeviceContext
    from kelp_amd.serving import Pool, Request, Server, latency_summary, poisson_stream
    t = time.time()
    w = synth.global_3d_workload(n_nodes=args.nodes or 1_000_000, n_particles=1)
    log(f"fields built in {time.time() - t:.1f} s")
    ctx = DeviceContext(0)
    ctx.upload_mesh(w.mesh)
    st = "bf16" if args.storage == "bf16" else None
    for s_, (U, V, W, _, _) in enumerate(w.snaps):
        ctx.upload_field(s_, U, V, W, storage=st)
    ctx.upload_field(len(w.snaps), w.snaps[0][0][0].copy(), w.snaps[0][1][0].copy(), storage=st)
    mesh = w.mesh
    n_snap = len(w.snaps)
    dt = 3600.0

    def make(rid, rng, arrival):
        n = int(min(500_000, max(1, round(5000.0 * np.exp(args.size_sigma * rng.standard_normal())))))
        # a random spherical cap of radius 1-10 degrees, seeds on wet elements
        # inside it; a cap without enough ocean is drawn again
        while True:
            lon, lat = _cap_seeds(rng, n)
            if lon is not None:
                break
        days = rng.uniform(1.0, 90.0)
        return Request(rid=rid, lon=lon, lat=lat, z=rng.uniform(0.0, 1000.0, lon.size),
                       n_steps=int(round(days * 24)), direction=float(rng.choice([1.0, -1.0])),
                       is3d=bool(rng.random() < 0.5), kh=float(rng.choice([0.0, 10.0])),
                       seed=int(rng.integers(1, 2 ** 62)), t0=float(rng.uniform(0, n_snap * w.snap_dt)),
                       arrival=arrival)

    def _cap_seeds(rng, n):
        rad = np.radians(rng.uniform(1.0, 10.0))
        cz = rng.uniform(-1.0, 1.0)
        clon = rng.uniform(-180.0, 180.0)
        c = np.array([np.sqrt(1 - cz * cz) * np.cos(np.radians(clon)),
                      np.sqrt(1 - cz * cz) * np.sin(np.radians(clon)), cz])
        lon_out, lat_out = [], []
        need = n
        for _attempt in range(8):
            if need <= 0:
                break
            m = max(4 * need, 256)
            zc = rng.uniform(np.cos(rad), 1.0, m)
            ph = rng.uniform(0, 2 * np.pi, m)
            e1 = np.cross(c, [0.0, 0.0, 1.0] if abs(c[2]) < 0.9 else [1.0, 0.0, 0.0])
            e1 /= np.linalg.norm(e1)
            e2 = np.cross(c, e1)
            sz = np.sqrt(1 - zc * zc)
            p = zc[:, None] * c + (sz * np.cos(ph))[:, None] * e1 + (sz * np.sin(ph))[:, None] * e2
            lo = np.degrees(np.arctan2(p[:, 1], p[:, 0]))
            la = np.degrees(np.arcsin(np.clip(p[:, 2], -1, 1)))
            nn = ctx.nearest(lo, la, exact=False)
            inside = ctx.containing(lo, la, nn) >= 0
            lon_out.append(lo[inside][:need])
            lat_out.append(la[inside][:need])
            need -= int(min(inside.sum(), need))
        if need > 0:
            return None, None
        return np.concatenate(lon_out), np.concatenate(lat_out)

    t = time.time()
    reqs = poisson_stream(args.rate, args.serve_seconds, 5050, make)
    log(f"{len(reqs)} requests generated in {time.time() - t:.1f} s "
        f"({sum(r.n for r in reqs)} particles, {sum(r.n * r.n_steps for r in reqs) / 1e9:.2f} G particle-steps)")
    pool = Pool(ctx, axis3=(0, n_snap), axis2=(n_snap, 1), snap_dt=w.snap_dt, max_capacity=args.pool,
                max_clouds=8192, choice=TuneChoice(256, 8, 4), order=2, dt=dt, exact_nearest=1, w_method=2)
    srv = Server({"c2": pool}, chunk=4)
    for r in reqs:
        r.pool = "c2"
    res = srv.run(reqs, realtime=True, log=log, progress_s=10.0)
    lat = latency_summary(res)
    pss = srv.alive_steps / srv.wall
    log(f"served {len(res)} requests in {srv.wall:.1f} s: {pss / 1e6:.1f} M particle-steps/s, "
        f"p50 {lat['p50_s']:.2f} s, p99 {lat['p99_s']:.2f} s")
    # the plain path on the first requests, one at a time
    specs = [CloudSpec(lon=r.lon, lat=r.lat, z=r.z, n_steps=r.n_steps, direction=r.direction, is3d=r.is3d,
                       kh=r.kh, seed=r.seed, t0=r.t0) for r in sorted(reqs, key=lambda q: q.rid)]
    plain, n_pl, el_pl = _alone_rate(ctx, specs, (0, n_snap), (n_snap, 1), w.snap_dt, 2, dt, args.cpu_budget)
    line = {
        "metric": "particle-steps/s", "value": pss, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "field_dtype": args.storage, "data": "synthetic",
        "config": {"workload": "serving_poisson_global_3d", "mesh_nodes": mesh.n_nodes, "rate_hz": args.rate,
                   "arrival_window_s": args.serve_seconds, "requests": len(reqs),
                   "size": "lognormal median 5000, sigma %.2f, cap 500000" % args.size_sigma,
                   "duration": "uniform 1-90 days at 1 h", "integrator": "rk2", "dt_s": dt,
                   "pool_capacity": pool.batch.capacity},
        "latency": lat,
        "wall_s": srv.wall,
        "plain_path": {"value": plain, "unit": "particle-steps/s", "requests": n_pl, "seconds": el_pl,
                       "what": "the same requests one at a time through kelp_step"},
        "speedup_vs_plain": pss / plain if plain > 0 else None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--workload", default="global_3d", choices=sorted(WORKLOADS) + ["batch", "serving", "precision"])
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--nodes", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cell-rule", default="occupancy", choices=["occupancy", "legacy"],
                    help="global_3d: the grid-hash cell rule (run_tracking's default: occupancy)")
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
    ap.add_argument("--clouds", type=int, default=256, help="batch: number of release clouds")
    ap.add_argument("--nside", type=int, default=1024, help="precision: HEALPix nside")
    ap.add_argument("--rate-particles", type=int, default=6_000_000,
                    help="precision: particles of the throughput runs")
    ap.add_argument("--rate", type=float, default=50.0, help="serving: request arrivals per second")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="serving: arrival window (s)")
    ap.add_argument("--size-sigma", type=float, default=1.0, help="serving: lognormal sigma of cloud sizes")
    ap.add_argument("--pool", type=int, default=8_000_000, help="serving: particle slots at most")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    from kelp_amd.device import DeviceContext
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "serving":
        return run_serving(args)
    if args.workload == "precision":
        return run_precision(args)
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
    # measured HBM bytes per launch: the profile summary's figure, valid only
    # for the configuration it was measured on (default sizes, f32 storage,
    # T/S sampling, re-sort every 8 steps, all snapshots resident)
    traffic, traffic_src = None, None
    profiled_cfg = (not (args.particles or args.nodes) and args.resort_every == 8
                    and not args.no_sample and not args.ring and args.storage == "f32"
                    and getattr(args, "cell_rule", "occupancy") == "occupancy")
    if getattr(wl, "profile", None) and profiled_cfg:
        traffic = profiled_traffic(wl.profile)
        if traffic is not None:
            traffic_src = os.path.relpath(wl.profile, ROOT)
    line = {
        "metric": "particle-steps/s", "value": pss, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "field_dtype": getattr(wl, "field_dtype", "f32"),
        "data": "synthetic",
        "config": dict(wl.config, alive_end=alive_end, resort_every=args.resort_every,
                       **({"grid_hash": wl.cell_report} if getattr(wl, "cell_report", None) else {})),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
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
