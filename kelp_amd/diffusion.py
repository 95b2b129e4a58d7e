"""Per-step diffusion arguments, mirroring src/kelp/numerics/diffusion.py:84.

The counter-based draws themselves run on the device (the step kernel's
split uniform-Kh kick); this module only carries the per-step key and
diffusivities the tracking layer hands to an integrator.
"""
from __future__ import annotations

from typing import NamedTuple


class DiffusionStepArgs(NamedTuple):
    scheme: str
    kh_x: float
    kh_y: float
    seed: int
    real_id: int
    step_i: int
    composition: str = "split"
    dres_m: float = 0.0
    kh_eval_position: str = "end"
    az_enabled: bool = False
    az_const: float = 0.0
    az_requires_field: bool = False
    pid_offset: int = 0
