"""Run configuration, mirroring src/kelp/core/config.py.

The YAML layout and the field names and defaults of the sections the
device path reads are the reference's (``mode``, ``numerics`` with its
``diffusion`` and ``beaching`` blocks, ``mesh``, ``data``, ``time``;
core/config.py:17-801), so a reference config file loads unchanged.  As in
the reference, unknown keys are ignored and every section must be present
(an empty one as ``section: {}``).  Values the device path does not
implement are accepted here and refused by the integrators with
``NotImplementedError``.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field, fields
from typing import Any, List, Optional, Tuple

import yaml


@dataclass
class ModeConfig:
    Fdrift2D: bool = True
    Fdrift3D: bool = False
    time_dir: float = 1.0
    Vpuls: bool = False
    var_u: Optional[str] = None
    var_v: Optional[str] = None
    var_w: Optional[str] = None
    var_eta: Optional[str] = None
    sigma_layer: Optional[int] = None
    sigma_mean: bool = False
    max_levels: Optional[int] = None
    sample_vars: List[str] = field(default_factory=list)
    track_failure_reason: bool = False
    w_positive: str = "auto"
    uv_vertical: str = "layer"


@dataclass
class DiffusionConfig:
    enabled: bool = False
    scheme: str = "uniform"
    kh_zonal: float = 0.0
    kh_meridional: float = 0.0
    var_kh_zonal: Optional[str] = None
    var_kh_meridional: Optional[str] = None
    dres_m: Optional[float] = None
    seed: int = 0
    composition: str = "split"
    kh_eval_position: str = "end"
    az_enabled: bool = False
    az_const: float = 0.0
    var_az: Optional[str] = None


@dataclass
class BeachingConfig:
    enabled: bool = False
    criteria: List[str] = field(default_factory=lambda: ["depth"])
    depth_min: float = 0.3
    use_eta: bool = False
    dry_mask_var: Optional[str] = None
    zone_depth: Optional[float] = None
    t_beach: Optional[float] = 86400.0
    resuspension: str = "never"
    rewet_hysteresis: float = 0.1
    min_wet_steps: int = 1
    t_resuspend: Optional[float] = None
    seed: Optional[int] = None


@dataclass
class NumericsConfig:
    dt: Optional[float] = None
    file_dt: Optional[float] = None
    output_freq: int = 1
    output_flush_every: int = 0
    Dmin: float = 0.5
    z0b_min: float = 0.1
    Dch: float = 1.0
    integrator: str = "euler"
    rk45_tol_m: float = 1.0
    rk45_h_min: Optional[float] = None
    uv_interp: str = "leastsq"
    edge_leastsq_bc: str = "freeslip"
    elem_search: str = "local"
    w_interp: str = "leastsq"
    sample_interp: str = "leastsq"
    temporal_interp: bool = False
    temporal_reconstruction: str = "linear"
    velocity_float32: bool = False
    prefetch_size: int = 1
    checkpoint_freq: int = 0
    periodic_lon: Optional[Tuple[float, float]] = None
    periodic_lat: Optional[Tuple[float, float]] = None
    diffusion: DiffusionConfig = field(default_factory=DiffusionConfig)
    beaching: BeachingConfig = field(default_factory=BeachingConfig)


@dataclass
class MeshConfig:
    type: str = "fesomc"
    meshpath: str = ""
    r_earth: Optional[float] = None
    grid_hash_cell_size: Optional[float] = None
    grid_hash_cell_rule: str = "occupancy"
    exact_nearest: bool = True


@dataclass
class DataConfig:
    ncinput: Any = ""
    ncinput_v: Any = ""
    ncinput_w: Any = ""
    inicloud: str = ""
    output_path: str = ""
    time_averaging: Optional[str] = None
    particle_offset: Optional[int] = None


@dataclass
class TimeConfig:
    start_time_idx: int = 0
    end_time_idx: Optional[int] = None
    n_steps: Optional[int] = None
    on_exhausted: str = "stop"
    num_runs: int = 1
    run_interval: int = 0
    n_pulses: int = 1
    pulse_interval: float = 0.0


@dataclass
class AppConfig:
    mode: ModeConfig
    numerics: NumericsConfig
    mesh: MeshConfig
    data: DataConfig
    time: TimeConfig


_SECTIONS = {"mode": ModeConfig, "numerics": NumericsConfig, "mesh": MeshConfig,
             "data": DataConfig, "time": TimeConfig}


def _build(cls, data: dict):
    """core/config.py:839 _build_section: known keys only (unknown keys are
    ignored, as in the reference)."""
    if data is None:
        data = {}
    if not isinstance(data, dict):
        raise ValueError(f"section {cls.__name__} must be a mapping, got {type(data).__name__}")
    names = {f.name for f in fields(cls)}
    kw = {k: v for k, v in data.items() if k in names}
    if cls is NumericsConfig:
        if isinstance(kw.get("diffusion"), dict):
            kw["diffusion"] = _build(DiffusionConfig, kw["diffusion"])
        if isinstance(kw.get("beaching"), dict):
            kw["beaching"] = _build(BeachingConfig, kw["beaching"])
        for key in ("periodic_lon", "periodic_lat"):
            if kw.get(key) is not None:
                lo, hi = (float(x) for x in kw[key])
                if not lo < hi:
                    raise ValueError(f"numerics.{key} needs lo < hi, got {kw[key]}")
                kw[key] = (lo, hi)
    return cls(**kw)


def config_from_dict(data: dict, base_dir: str = ".") -> AppConfig:
    """core/config.py:830-880: every section present; relative paths
    resolve against ``base_dir``; ``~`` expands."""
    missing = [s for s in _SECTIONS if s not in data]
    if missing:
        raise ValueError(f"config: missing section(s) {missing} (an empty section is `name: {{}}`)")
    cfg = AppConfig(**{s: _build(cls, data[s]) for s, cls in _SECTIONS.items()})
    for obj, key in ((cfg.mesh, "meshpath"), (cfg.data, "inicloud"), (cfg.data, "output_path"),
                     (cfg.data, "ncinput"), (cfg.data, "ncinput_v"), (cfg.data, "ncinput_w")):
        v = getattr(obj, key)
        if isinstance(v, str) and v and "://" not in v:
            v = os.path.expanduser(v)
            if not os.path.isabs(v):
                v = os.path.normpath(os.path.join(base_dir, v))
            setattr(obj, key, v)
    if cfg.numerics.integrator not in ("euler", "rk2", "rk4", "rk45"):
        raise ValueError(f"numerics.integrator: unknown {cfg.numerics.integrator!r}")
    if cfg.time.on_exhausted not in ("stop", "cycle"):
        raise ValueError(f"time.on_exhausted must be stop or cycle, got {cfg.time.on_exhausted!r}")
    return cfg


def load_config(path: str) -> AppConfig:
    """core/config.py:883 load_config (safe YAML)."""
    with open(path) as fh:
        data = yaml.safe_load(fh) or {}
    if not isinstance(data, dict):
        raise ValueError(f"{path}: the config must be a mapping of sections")
    return config_from_dict(data, os.path.dirname(os.path.abspath(path)))


def resolve_time_params(time_cfg: TimeConfig) -> Tuple[int, int]:
    """core/config.py:928: (start_time_idx, n_steps), n_steps first, else
    end_time_idx - start_time_idx."""
    if time_cfg.on_exhausted not in ("stop", "cycle"):
        raise ValueError(f"on_exhausted must be 'stop' or 'cycle', got '{time_cfg.on_exhausted}'")
    start = time_cfg.start_time_idx
    if time_cfg.n_steps is not None:
        return start, time_cfg.n_steps
    if time_cfg.end_time_idx is not None:
        return start, time_cfg.end_time_idx - start
    raise ValueError("TimeConfig must specify either 'n_steps' or 'end_time_idx'")
