"""Build the HIP library (and the CPU oracle used by the tests).

``python -m kelp_amd.build`` compiles kelp_amd/csrc/kelp_hip.hip for gfx950
into kelp_amd/libkelp_hip.so with hipcc.  Contraction is disabled
(-ffp-contract=off) so every product and sum rounds like the reference's
Numba kernels; packed-FP32 instructions are disabled as a precaution.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRCS = [os.path.join(HERE, "csrc", f) for f in ("kelp_hip.hip", "kelp_sort.hip")]
OUT = os.path.join(HERE, "libkelp_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build_hip(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *SRCS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{' '.join(cmd)}\n{r.stderr[-4000:]}")
    if verbose and r.stderr:
        print(r.stderr[-2000:], file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def build(verbose: bool = False) -> str:
    out = build_hip(verbose)
    build_oracle()
    return out


if __name__ == "__main__":
    print(build(verbose=True))
