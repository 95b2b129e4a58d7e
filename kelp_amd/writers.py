"""Trajectory output format 2 as a file (writers.py:1209 TrajectoryWriter).

The reference writes CF-1.8 trajectory files with netCDF4 (HDF5); that
library is not part of this build, so the same layout goes into a classic
NetCDF file through ``scipy.io.netcdf_file``: dimensions ``trajectory`` and
``time`` (the records, written at the end of the realization); ``trajectory`` (int32 ids), ``time`` (seconds at the
end of each recorded step), ``lon``/``lat``/``z``/``vel_u``/``vel_v``
[/``vel_w``] and sampled fields as float64 (trajectory, time) with NaN fill,
``status`` int32 and ``failure_reason`` int8 with their flag tables, the
``release_*`` positions and instants, and the run-describing global
attributes (``integrator``, ``dt``, ``time_averaging``, ...).  Readers of
the reference's files (xarray, netCDF4) read these too.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

#: the output format written (writers.py OUTPUT_FORMAT)
OUTPUT_FORMAT = 2

_COORDS = "lat lon z"
_VEL_NOTE = ("flow velocity (the same sign in forward and backward runs) that the integrator "
             "evaluated for the step ending at this record, at the step's start position and "
             "time (rk45: at its end); NaN unless the particle was ALIVE through the step")
_SAMPLE_NOTE = ("forcing field at the particle position and record time (discrete mode: the "
                "forcing record at the end of the step; interpolated mode: blended to the "
                "record time; at the end of the forcing, both modes: the start-of-step field)")

# per-variable CF metadata of format 2 (the reference's tables, restated)
_VAR_ATTRS = {
    "lon": dict(standard_name="longitude", long_name="longitude", units="degrees_east"),
    "lat": dict(standard_name="latitude", long_name="latitude", units="degrees_north"),
    "z": dict(standard_name="depth", long_name="depth below the surface", units="m", positive="down"),
    "vel_u": dict(long_name="eastward flow velocity at the particle (geographic)", units="m s-1",
                  comment=_VEL_NOTE),
    "vel_v": dict(long_name="northward flow velocity at the particle (geographic)", units="m s-1",
                  comment=_VEL_NOTE),
    "vel_w": dict(long_name="vertical flow velocity at the particle, positive down "
                            "(sigma meshes: the dia-sigma flux)", units="m s-1", comment=_VEL_NOTE),
    "release_lon": dict(long_name="release longitude", units="degrees_east"),
    "release_lat": dict(long_name="release latitude", units="degrees_north"),
    "release_z": dict(long_name="release depth, positive down", units="m"),
}
# ParticleStatus ALIVE, BEACHED, OUTSIDE, NOT_RELEASED; FailureReason codes
_STATUS_FLAGS = np.array([1, 2, 3, 4], np.int32)
_STATUS_MEANINGS = "alive beached outside not_released"
_FAILURE_FLAGS = np.array([0, 1, 2, 3, 4], np.int8)
_FAILURE_MEANINGS = "none interpolation_failed out_of_bounds surface_hit bottom_hit"


def run_attributes(config, start_time_idx: int, n_steps: int, time_averaging: str,
                   mesh=None, numerics=None) -> Dict[str, object]:
    """The run-describing global attributes (tracking.py:4024-4112 and
    :2290-2299) of one realization's file; ``numerics`` with the resolved
    step when ``config.numerics.dt`` was left to the data."""
    num, mode = numerics or config.numerics, config.mode
    is3d = bool(getattr(mode, "Fdrift3D", False))
    meta: Dict[str, object] = {
        "integrator": num.integrator, "dt": float(num.dt), "mesh_type": str(config.mesh.type),
        "time_dir": float(mode.time_dir), "start_time_idx": int(start_time_idx),
        "n_steps": int(n_steps), "mode_3d": int(is3d),
        "temporal_interp": int(bool(num.temporal_interp)),
        "temporal_reconstruction": str(num.temporal_reconstruction),
        "time_averaging": str(time_averaging),
        "kelp_version": "kelp_amd (MI355X)",
    }
    if is3d:
        meta["w_interp"] = str(num.w_interp)
        meta["w_positive"] = str(getattr(mode, "w_positive", "auto"))
    if mode.sample_vars:
        meta["sample_interp"] = str(num.sample_interp)
    meta["uv_interp"] = str(num.uv_interp)
    meta["exact_nearest"] = int(bool(config.mesh.exact_nearest))
    if mesh is not None and getattr(mesh, "gh_cell_size", 0.0) > 0.0:
        meta["grid_hash_cell"] = float(mesh.gh_cell_size)
    meta["grid_hash_cell_rule"] = str(config.mesh.grid_hash_cell_rule)
    meta["prefetch_size"] = int(getattr(num, "prefetch_size", 1))
    meta["particle_offset"] = int(config.data.particle_offset or 0)
    d, b = num.diffusion, num.beaching
    draws = bool(d.enabled) or (bool(b.enabled) and ("stochastic" in (b.criteria or ())
                                                     or b.resuspension in ("stochastic",
                                                                           "auto_or_stochastic")))
    meta["random_draws"] = int(draws)
    if d.enabled:
        meta.update(diffusion_enabled=1, diffusion_scheme=str(d.scheme), kh_zonal=float(d.kh_zonal),
                    kh_meridional=float(d.kh_meridional), diffusion_seed=int(d.seed))
    return meta


def write_trajectories(path: str, arrays: Dict[str, np.ndarray], time_units: str = "seconds",
                       calendar: Optional[str] = "standard", attrs: Optional[Dict[str, object]] = None,
                       sample_attrs: Optional[Dict[str, dict]] = None) -> None:
    """One realization's records (``TrajectoryRecord.as_arrays()``: lon,
    lat, z, status, failure_reason, vel_u/vel_v[/vel_w], ``sample__<name>``,
    time, release_*) into a format-2 file at ``path``; ``time_units`` such
    as ``"seconds since 1958-01-01 23:15:00"`` (the first record's stamp)."""
    from scipy.io import netcdf_file
    P, T = np.asarray(arrays["lon"]).shape
    with netcdf_file(path, "w", version=2) as nc:
        nc.createDimension("trajectory", P)
        # written once per realization: classic NetCDF allows only the
        # first dimension to be unlimited, so time is a fixed dimension here
        nc.createDimension("time", T)
        tv = nc.createVariable("trajectory", "i4", ("trajectory",))
        tv[:] = np.arange(P, dtype=np.int32)
        tv.cf_role = "trajectory_id"
        tv.long_name = "particle ID"
        time = nc.createVariable("time", "f8", ("time",))
        time[:] = np.asarray(arrays["time"], np.float64)
        time.long_name = "time at the end of the recorded step"
        time.units = time_units
        if " since " in time_units:
            time.standard_name = "time"
            time.axis = "T"
            if calendar:
                time.calendar = calendar

        def field(name, data, extra):
            v = nc.createVariable(name, "f8", ("trajectory", "time"))
            v._FillValue = np.nan
            v[:] = np.asarray(data, np.float64)
            for k, val in extra.items():
                setattr(v, k, val)

        for name in ("lon", "lat", "z"):
            if name in arrays:
                field(name, arrays[name], _VAR_ATTRS[name])
        sv = nc.createVariable("status", "i4", ("trajectory", "time"))
        sv[:] = np.asarray(arrays["status"], np.int32)
        sv.long_name = "particle status"
        sv.flag_values = _STATUS_FLAGS
        sv.flag_meanings = _STATUS_MEANINGS
        sv.coordinates = _COORDS
        for name in ("release_lon", "release_lat", "release_z"):
            if name in arrays:
                v = nc.createVariable(name, "f8", ("trajectory",))
                v._FillValue = np.nan
                v[:] = np.asarray(arrays[name], np.float64)
                for k, val in _VAR_ATTRS[name].items():
                    setattr(v, k, val)
        if "release_time" in arrays:
            v = nc.createVariable("release_time", "f8", ("trajectory",))
            v._FillValue = np.nan
            v[:] = np.asarray(arrays["release_time"], np.float64)
            v.long_name = "instant the particle starts moving"
            v.units = time_units
            if calendar and " since " in time_units:
                v.calendar = calendar
        for name in ("vel_u", "vel_v", "vel_w"):
            if name in arrays:
                field(name, arrays[name], _VAR_ATTRS[name])
        if "failure_reason" in arrays:
            fv = nc.createVariable("failure_reason", "b", ("trajectory", "time"))
            fv[:] = np.asarray(arrays["failure_reason"], np.int8)
            fv.long_name = "particle failure reason"
            fv.flag_values = _FAILURE_FLAGS
            fv.flag_meanings = _FAILURE_MEANINGS
            fv.coordinates = _COORDS
        for key in sorted(k for k in arrays if k.startswith("sample__")):
            name = key[len("sample__"):]
            extra = {k: v for k, v in ((sample_attrs or {}).get(name) or {}).items() if v is not None}
            extra.setdefault("comment", _SAMPLE_NOTE)
            extra["coordinates"] = _COORDS
            field(name, arrays[key], extra)
        nc.Conventions = "CF-1.8"
        nc.featureType = "trajectory"
        nc.kelp_output_format = np.int32(OUTPUT_FORMAT)
        nc.history = "Created by kelp_amd write_trajectories"
        for k, v in (attrs or {}).items():
            setattr(nc, str(k), v)


def read_trajectories(path: str) -> Dict[str, np.ndarray]:
    """Every variable of a format-2 file (writers.py load_trajectories)."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        return {k: np.array(v[:]) for k, v in nc.variables.items()}
