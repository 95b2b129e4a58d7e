"""Build the HIP library (and the CPU oracle used by the tests).

``python -m kelp_amd.build`` compiles the sources under kelp_amd/csrc for
gfx950 into kelp_amd/libkelp_hip.so with hipcc.  The step kernel is
instantiated per (3-D, field dtype, u/v layout) in its own object file
(kelp_step_variant.hip compiled with -DKV_*), so the objects compile in
parallel and link into one library.  Contraction is disabled
(-ffp-contract=off) so every product and sum rounds like the reference's
Numba kernels; packed-FP32 instructions are disabled as a precaution.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libkelp_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

#: (object name, source, extra defines)
UNITS = [("kelp_hip", "kelp_hip.hip", []), ("kelp_sort", "kelp_sort.hip", []),
         ("kelp_sample", "kelp_sample.hip", [])] + [
    (f"kelp_sigma_{'f32' if f32 else 'f64'}_{'elem' if el else 'node'}", "kelp_step_sigma.hip",
     [f"-DKS_F32={f32}", f"-DKS_ELEM={el}"]) for f32 in (0, 1) for el in (0, 1)] + [
    (f"kelp_step_{'3d' if d3 else '2d'}_{'f32' if f32 else 'f64'}_{'elem' if el else 'node'}",
     "kelp_step_variant.hip", [f"-DKV_IS3D={d3}", f"-DKV_F32={f32}", f"-DKV_ELEM={el}"])
    for d3 in (0, 1) for f32 in (0, 1) for el in (0, 1)
] + [
    (f"kelp_step_{'3d' if d3 else '2d'}_bf16_node", "kelp_step_variant.hip",
     [f"-DKV_IS3D={d3}", "-DKV_F32=0", "-DKV_BF16=1", "-DKV_ELEM=0"]) for d3 in (0, 1)
] + [
    (f"kelp_step_{'3d' if d3 else '2d'}_f16s_node", "kelp_step_variant.hip",
     [f"-DKV_IS3D={d3}", "-DKV_F32=0", "-DKV_F16S=1", "-DKV_ELEM=0"]) for d3 in (0, 1)
] + [("kelp_batch_host", "kelp_batch_host.hip", [])] + [
    (f"kelp_batch_{dt}_{'sigma' if sg else 'zlevel'}", "kelp_batch.hip",
     [f"-DKB_F32={int(dt == 'f32')}", f"-DKB_BF16={int(dt == 'bf16')}", f"-DKB_F16S={int(dt == 'f16s')}",
      f"-DKB_SIGMA={sg}"])
    for dt, sg in (("f32", 0), ("f32", 1), ("bf16", 0), ("f16s", 0))
]


# the slowest units first, so the parallel build does not end on a long tail
_SLOW_FIRST = ("kelp_batch_", "kelp_step_3d", "kelp_sigma_")
UNITS.sort(key=lambda u: next((i for i, p in enumerate(_SLOW_FIRST) if u[0].startswith(p)), len(_SLOW_FIRST)))


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{' '.join(cmd)}\n{r.stderr[-4000:]}")
    return r.stderr


def _up_to_date(obj: str, dep: str, cmd) -> bool:
    """True when ``obj`` is newer than every file its dependency list names
    and was built by the same command line."""
    try:
        with open(dep) as fh:
            text = fh.read()
        with open(obj + ".cmd") as fh:
            if fh.read() != " ".join(cmd):
                return False
        t = os.path.getmtime(obj)
    except OSError:
        return False
    files = text.replace("\\\n", " ").split(":", 1)[-1].split()
    return all(os.path.exists(f) and os.path.getmtime(f) <= t for f in files)


def build_hip(verbose: bool = False, jobs: int = 0, incremental: bool = False) -> str:
    """Compile every unit and link.  ``incremental``: skip units whose
    object is newer than the sources and headers it was built from."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(len(UNITS), max(1, os.cpu_count() or 1))
    cmds, objs = [], []
    for name, src, defs in UNITS:
        obj = os.path.join(OBJ_DIR, name + ".o")
        dep = os.path.join(OBJ_DIR, name + ".d")
        objs.append(obj)
        cmd = [hipcc(), *HIPCC_FLAGS, *defs, "-c", "-o", obj, os.path.join(CSRC, src)]
        if incremental and _up_to_date(obj, dep, cmd):
            continue
        cmds.append(cmd + ["-MD", "-MF", dep])

    def _compile(cmd):
        out = _run(cmd)
        obj = cmd[cmd.index("-o") + 1]
        with open(obj + ".cmd", "w") as fh:
            fh.write(" ".join(cmd[:-3]))
        return out

    with ThreadPoolExecutor(jobs) as ex:
        logs = list(ex.map(_compile, cmds))
    tmp = OUT + ".tmp"
    logs.append(_run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]))
    if verbose:
        for lg in logs:
            if lg:
                print(lg[-2000:], file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def build(verbose: bool = False, incremental: bool = False) -> str:
    out = build_hip(verbose, incremental=incremental)
    build_oracle()
    return out


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
