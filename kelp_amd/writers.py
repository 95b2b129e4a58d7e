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

import os
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
    # tracking.py:3380-3386: "explicit" when the cell was given
    meta["grid_hash_cell_rule"] = ("explicit" if getattr(config.mesh, "grid_hash_cell_size", None) is not None
                                   else str(config.mesh.grid_hash_cell_rule))
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
                       sample_attrs: Optional[Dict[str, dict]] = None,
                       restart: Optional[dict] = None) -> None:
    """One realization's records (``TrajectoryRecord.as_arrays()``: lon,
    lat, z, status, failure_reason, vel_u/vel_v[/vel_w], ``sample__<name>``,
    time, release_*) into a format-2 file at ``path``; ``time_units`` such
    as ``"seconds since 1958-01-01 23:15:00"`` (the first record's stamp).
    ``restart``: the keyword arguments of :func:`write_restart_file` (less
    the path) for a checkpoint block inside the file, as the reference's
    writer keeps one (writers.py:567 write_checkpoint), so
    :func:`load_restart` resumes from the trajectory file itself.  The file
    is written to a temporary sibling and renamed into place."""
    from scipy.io import netcdf_file
    P, T = np.asarray(arrays["lon"]).shape
    with _atomic(path) as tmp, netcdf_file(tmp, "w", version=2) as nc:
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
        if restart is not None:
            _restart_block(nc, P, **restart)


def read_trajectories(path: str) -> Dict[str, np.ndarray]:
    """Every variable of a format-2 file but the checkpoint's ``restart_*``
    (writers.py load_trajectories; :func:`load_restart` reads those)."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        return {k: np.array(v[...]) for k, v in nc.variables.items() if not k.startswith("restart_")}


# ---------------------------------------------------------------------------
# Checkpoints in the reference's restart contract
# ---------------------------------------------------------------------------
# The variables, their types and the attributes are the ones the reference
# writes and reads (writers.py:581-664 the block inside a trajectory file,
# :1420 write_restart_file the standalone file, :1541 load_restart), so a
# file from either side resumes the other's run: f8 / i4 / i1 arrays over
# ``trajectory``, the i4 scalars ``restart_step_i``, ``restart_file_idx``
# and ``restart_valid`` (1 only on a complete checkpoint), the attribute
# ``restart_optional`` naming the optional arrays present, and the frame /
# shard / vertical-mode attributes.  ``restart_sim_time`` (the clock of a
# time-interpolated run) is an attribute of this build's own; the
# reference's reader ignores it.  Files are classic NetCDF (CDF-2).

_RESTART_ARRAYS = (
    # name, type, ParticleState field (None: a separate argument)
    ("restart_lon", "f8", "lon"), ("restart_lat", "f8", "lat"), ("restart_z", "f8", "z"),
    ("restart_status", "i4", "status"), ("restart_nearest_node", "i4", "nearest_node"),
    ("restart_failure_reason", "b", "failure_reason"), ("restart_release_step", "i4", None),
    ("restart_elem", "i4", "elem"), ("restart_release_lon", "f8", "release_lon"),
    ("restart_release_lat", "f8", "release_lat"), ("restart_release_z", "f8", "release_z"),
    ("restart_wet_streak", "i4", None),
)
_REQUIRED = ("restart_lon", "restart_lat", "restart_z", "restart_status", "restart_nearest_node")
_ROTATION = ("restart_rotation_alpha", "restart_rotation_beta", "restart_rotation_gamma")
_NP_TYPE = {"f8": np.float64, "i4": np.int32, "b": np.int8}


class _atomic:
    """A temporary sibling of ``path``, renamed onto it after a clean exit
    (synced first) and removed after a failed one: an interrupted write
    never replaces or damages the file already at ``path``."""

    def __init__(self, path: str):
        self.path = str(path)
        self.tmp = f"{self.path}.tmp{os.getpid()}"

    def __enter__(self) -> str:
        return self.tmp

    def __exit__(self, exc_type, exc, tb):
        if exc_type is None:
            fd = os.open(self.tmp, os.O_RDONLY)
            try:
                os.fsync(fd)
            finally:
                os.close(fd)
            os.replace(self.tmp, self.path)
        else:
            try:
                os.remove(self.tmp)
            except OSError:
                pass
        return False


def _restart_payload(state, release_step, wet_streak) -> Dict[str, np.ndarray]:
    """The arrays of one checkpoint by restart name, required first; a
    length that differs from ``state.lon``'s is an error (writers.py
    _check_length), and release positions come as all three or none."""
    P = int(np.asarray(state.lon).shape[0])
    rel = [getattr(state, k, None) for k in ("release_lon", "release_lat", "release_z")]
    if any(a is None for a in rel) and not all(a is None for a in rel):
        raise ValueError("restart: release_lon / release_lat / release_z must be given together")
    if release_step is None:
        release_step = getattr(state, "release_step", None)
    out: Dict[str, np.ndarray] = {}
    for name, typ, attr in _RESTART_ARRAYS:
        if attr is not None:
            val = getattr(state, attr, None)
        else:
            val = release_step if name == "restart_release_step" else wet_streak
        if val is None:
            if name in _REQUIRED:
                raise ValueError(f"restart: {name} is required")
            continue
        a = np.asarray(val)
        if a.shape != (P,):
            raise ValueError(f"restart: {name} has shape {a.shape}, expected ({P},)")
        out[name] = a.astype(_NP_TYPE[typ])
    return out


def _restart_block(nc, P: int, state, step_i: int, file_idx: int, wet_streak=None, rotation_angles=None,
                   release_step=None, particle_offset=None, uv_vertical=None, sim_time=None) -> None:
    """The restart variables and attributes into an open classic file that
    has (or gets) the ``trajectory`` dimension of length ``P``: the flag
    is written 0 first and 1 last."""
    payload = _restart_payload(state, release_step, wet_streak)
    if int(np.asarray(state.lon).shape[0]) != P:
        raise ValueError(f"restart: {np.asarray(state.lon).shape[0]} particles, the file has {P}")
    if "trajectory" not in nc.dimensions:
        nc.createDimension("trajectory", P)
    flag = nc.createVariable("restart_valid", "i4", ())
    flag[...] = 0
    types = {name: typ for name, typ, _ in _RESTART_ARRAYS}
    for name, a in payload.items():
        nc.createVariable(name, types[name], ("trajectory",))[:] = a
    for name, val in (("restart_step_i", step_i), ("restart_file_idx", file_idx)):
        nc.createVariable(name, "i4", ())[...] = int(val)
    if rotation_angles is not None:
        for k, v in zip(_ROTATION, rotation_angles):
            setattr(nc, k, float(v))
    if particle_offset is not None:
        # classic files have no 64-bit integer attribute: int32, or a double
        # (exact below 2^53) for offsets beyond it
        off = int(particle_offset)
        nc.restart_particle_offset = np.int32(off) if -2**31 <= off < 2**31 else np.float64(off)
    if uv_vertical is not None and str(uv_vertical) != "layer":
        nc.restart_uv_vertical = str(uv_vertical)
    if sim_time is not None:
        nc.restart_sim_time = float(sim_time)
    nc.restart_optional = " ".join(n for n in payload if n not in _REQUIRED)
    flag[...] = 1


def write_restart_file(path: str, state, step_i: int, file_idx: int, wet_streak=None, rotation_angles=None,
                       release_step=None, particle_offset=None, uv_vertical=None, sim_time=None) -> None:
    """writers.py:1420 write_restart_file: one self-contained checkpoint
    (``state`` a :class:`~kelp_amd.particles.ParticleState`, the global
    ``step_i`` of the last step that ran, ``file_idx`` the record index of
    the next step; optional wet streaks, release steps, rotated-frame
    angles, shard offset, ``uv_vertical`` and -- for a time-interpolated
    run -- the clock ``sim_time``).  Nothing is written, and a file already
    at ``path`` is kept, when an array's length is wrong."""
    from scipy.io import netcdf_file
    P = int(np.asarray(state.lon).shape[0])
    _restart_payload(state, release_step, wet_streak)
    with _atomic(path) as tmp, netcdf_file(tmp, "w", version=2) as nc:
        _restart_block(nc, P, state, step_i, file_idx, wet_streak=wet_streak, rotation_angles=rotation_angles,
                       release_step=release_step, particle_offset=particle_offset, uv_vertical=uv_vertical,
                       sim_time=sim_time)
        nc.history = "Created by kelp_amd write_restart_file"


def _file_kind(path: str) -> str:
    if os.path.isdir(path):
        return "zarr"
    with open(path, "rb") as fh:
        magic = fh.read(4)
    if magic[:3] == b"CDF":
        return "classic"
    if magic == b"\x89HDF":
        return "hdf5"
    raise ValueError(f"{path}: neither a NetCDF nor a Zarr store")


def load_restart(path: str) -> dict:
    """writers.py:1541 load_restart: the checkpoint in ``path`` (a
    standalone restart file or a trajectory file with a restart block) as
    the reference returns it -- ``state`` (ParticleState; ``release_step``,
    ``elem`` and the release positions attached when stored), ``step_i``,
    ``file_idx``, ``wet_streak``, ``rotation_angles``, ``release_step``,
    ``particle_offset``, ``uv_vertical`` -- plus ``sim_time`` (None when
    not stored).  ValueError when there is no complete checkpoint (no flag,
    or a flag that is not 1).  Classic NetCDF files are read here; the
    reference's NETCDF4 (HDF5) and Zarr files need the ``netCDF4`` /
    ``zarr`` packages, and without them the error names the package."""
    kind = _file_kind(path)
    if kind == "classic":
        from scipy.io import netcdf_file
        with netcdf_file(path, "r", mmap=False) as nc:
            attrs = {k: _attr(v) for k, v in nc._attributes.items()}
            return _assemble(path, nc.variables, lambda n: np.asarray(nc.variables[n][...]), attrs)
    pkg = "netCDF4" if kind == "hdf5" else "zarr"
    try:
        mod = __import__(pkg)
    except ImportError:
        raise RuntimeError(f"{path} is a {kind} file; reading it needs the {pkg} package") from None
    if pkg == "zarr":
        g = mod.open_group(path, mode="r")
        return _assemble(path, g, lambda n: np.asarray(g[n][:]), {k: _attr(v) for k, v in g.attrs.items()})
    with mod.Dataset(path, "r") as ds:
        attrs = {k: _attr(ds.getncattr(k)) for k in ds.ncattrs()}
        # a flag created but never written reads masked: 0
        return _assemble(path, ds.variables, lambda n: np.ma.filled(ds.variables[n][:], 0), attrs)


def _attr(v):
    if isinstance(v, bytes):
        return v.decode()
    a = np.asarray(v)
    if a.dtype.kind in "iuf" and a.size == 1:
        return a.ravel()[0].item()
    return v


def _assemble(path: str, names, get, attrs: dict) -> dict:
    from .particles import ParticleState
    if "restart_valid" not in names or int(np.asarray(get("restart_valid")).ravel()[0]) != 1:
        raise ValueError(f"No valid restart checkpoint in {path}")
    listed = set(str(attrs["restart_optional"]).split()) if "restart_optional" in attrs else None
    got = {}
    for name, typ, _ in _RESTART_ARRAYS:
        ok = name in names and (name in _REQUIRED or listed is None or name in listed)
        got[name] = np.array(get(name), dtype=_NP_TYPE[typ]) if ok else None
    rot = tuple(float(attrs[k]) for k in _ROTATION) if _ROTATION[0] in attrs else None
    off = next((int(attrs[k]) for k in ("restart_particle_offset", "particle_offset") if k in attrs), None)
    uvv = next((str(attrs[k]) for k in ("restart_uv_vertical", "uv_vertical") if k in attrs), None)
    sim = float(attrs["restart_sim_time"]) if "restart_sim_time" in attrs else None
    state = ParticleState(lon=got["restart_lon"], lat=got["restart_lat"], z=got["restart_z"],
                          status=got["restart_status"], nearest_node=got["restart_nearest_node"],
                          failure_reason=got["restart_failure_reason"], release_step=got["restart_release_step"],
                          elem=got["restart_elem"], release_lon=got["restart_release_lon"],
                          release_lat=got["restart_release_lat"], release_z=got["restart_release_z"])
    return {"state": state, "step_i": int(np.asarray(get("restart_step_i")).ravel()[0]),
            "file_idx": int(np.asarray(get("restart_file_idx")).ravel()[0]),
            "wet_streak": got["restart_wet_streak"], "rotation_angles": rot,
            "release_step": got["restart_release_step"], "particle_offset": off, "uv_vertical": uvv,
            "sim_time": sim}
