"""Benchmark of the particle time-stepping hot path on one MI355X.

``python bench.py [--steps K] [--warmup W] [--workload NAME]`` prints one JSON
line: whole-job particle-steps/s with the mesh, snapshots and particle state
already resident in HBM, the roofline of the step kernel (algorithmic bytes
per launch over its measured average duration, HIP events on the stream the
kernel runs on) and a CPU baseline (the oracle, one host core, on a bounded
sample of the same workload).
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


def workload_double_gyre(args):
    """2-D double gyre on the jittered 0.25-degree lattice, 10,000 particles, RK2,
    dt 600 s.  Algorithmic bytes per particle-step: 377 (state 65 B + 3 sites x
    88 B of anchor row + element connectivity/coordinates + 2 stages x 24 B of
    f32 u,v at 3 vertices)."""
    from kelp_amd import synth
    w = synth.double_gyre_workload(n_particles=args.particles or 10_000)
    return dict(name="double_gyre_2d_rk2", w=w, b_alg=377.0,
                config={"workload": "double_gyre_2d_rk2", "mesh_nodes": w.mesh.n_nodes,
                        "mesh_elements": w.mesh.n_elements, "particles": int(w.lon.size),
                        "integrator": "rk2", "dt_s": w.dt, "field_dtype": "f32",
                        "steps_per_run": w.n_steps})


WORKLOADS = {"double_gyre": workload_double_gyre}


def cpu_baseline(w, order, budget_s):
    """The oracle (glibc libm, one core) on the first particles of the same
    workload for as many steps as fit in ``budget_s``."""
    from oracle.oracle import OracleMesh
    om = OracleMesh(w.mesh, "glibc")
    n = min(w.lon.size, 10_000)
    st = dict(lon=w.lon[:n].copy(), lat=w.lat[:n].copy(), status=np.ones(n, np.int32),
              nn=np.full(n, -1, np.int32))
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        st = om.step2d(st, order, w.u, w.v, w.dt, 1.0)
        steps += 1
    el = time.perf_counter() - t0
    return {"value": n * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} particles x {steps} steps, C oracle (glibc libm), 1 thread, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="double_gyre", choices=sorted(WORKLOADS))
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from kelp_amd.device import DeviceContext
    spec = WORKLOADS[args.workload](args)
    w = spec["w"]
    ctx = DeviceContext(0)
    ctx.upload_mesh(w.mesh)
    ctx.upload_field(0, w.u, w.v)
    p = ctx.particles(w.lon, w.lat)
    order = w.order
    for _ in range(args.warmup):
        ctx.step(p, order, 0, w.dt, 1.0)
    ctx.synchronize()
    torch = ctx.torch
    torch.cuda.synchronize()
    alive_steps = 0
    t0 = time.perf_counter()
    ctx.event_record(0)
    for _ in range(args.steps):
        ctx.step(p, order, 0, w.dt, 1.0)
    ctx.event_record(1)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ctx.event_elapsed_ms(0, 1)
    alive = int((p.status == 1).sum().item())
    n = int(w.lon.size)
    # every timed step advanced the particles alive at its start; the count
    # alive at the end is a lower bound, the start count an upper bound
    alive_steps = n * args.steps
    pss = alive_steps / wall
    avg_launch_s = kern_ms / 1e3 / args.steps
    achieved = spec["b_alg"] * n / avg_launch_s / 1e9
    line = {
        "metric": "particle-steps/s", "value": pss, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(spec["config"], alive_end=alive),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel_avg_us": avg_launch_s * 1e6, "b_alg_per_particle_step": spec["b_alg"]},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, order, args.cpu_budget)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
