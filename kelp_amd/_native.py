"""ctypes binding of libkelp_hip.so (the C ABI declared in include/kelp_hip.h).

The library is built in-tree by :func:`kelp_amd.build.build`.  Loading fails
loudly when it is missing: there is no CPU fallback for the step.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64, c_void_p

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libkelp_hip.so")

KELP_OK = 0
KELP_DT_F64 = 0
KELP_DT_F32 = 1
KELP_DT_BF16 = 2
KELP_DT_F16S = 3


class KelpError(RuntimeError):
    """A C-ABI call returned an error code."""


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("n_nodes", c_int64), ("n_elems", c_int64), ("max_v", c_int64), ("max_conn", c_int64),
        ("host_nodes_lon", c_void_p), ("host_nodes_lat", c_void_p), ("host_depth", c_void_p),
        ("host_elem_nodes", c_void_p), ("host_node_to_elems", c_void_p),
        ("host_node_to_elems_count", c_void_p), ("host_gh_offsets", c_void_p),
        ("host_gh_node_ids", c_void_p),
        ("gh_cell", c_double), ("gh_lon_min", c_double), ("gh_lat_min", c_double),
        ("gh_nx", c_int64), ("gh_ny", c_int64), ("gh_wrap", c_int64), ("gh_far_rows", c_int64),
        ("r_earth", c_double),
        ("n_levels", c_int64), ("host_z_levels", c_void_p),
        ("host_elem_coastal", c_void_p),
        ("host_elem_center_lon", c_void_p), ("host_elem_center_lat", c_void_p),
        ("vertical_coord", c_int64), ("host_elem_neighbors", c_void_p),
    ]


class Particles(ctypes.Structure):
    _fields_ = [
        ("n", c_int64),
        ("lon", c_void_p), ("lat", c_void_p), ("z", c_void_p),
        ("status", c_void_p), ("nn", c_void_p), ("fr", c_void_p),
        ("u", c_void_p), ("v", c_void_p), ("w", c_void_p),
        ("pid", c_void_p), ("wet_streak", c_void_p), ("alive_count", c_void_p),
        ("elem", c_void_p),
    ]


class StepArgs(ctypes.Structure):
    _fields_ = [
        ("order", c_int32), ("slot_a", c_int32), ("slot_b", c_int32), ("slot_t", c_int32),
        ("alpha0", c_double), ("alphas", c_double * 4),
        ("dt", c_double), ("time_dir", c_double),
        ("exact_nearest", c_int32), ("w_method", c_int32),
        ("kick", c_int32), ("beach", c_int32),
        ("kh_x", c_double), ("kh_y", c_double),
        ("seed", ctypes.c_uint64), ("real_id", ctypes.c_uint64), ("step_i", ctypes.c_uint64),
        ("use_depth", c_int32), ("use_coastal", c_int32), ("use_stochastic", c_int32),
        ("on_failure", c_int32), ("resuspension", c_int32), ("min_wet_steps", c_int32),
        ("depth_min", c_double), ("zone_depth", c_double), ("t_beach", c_double),
        ("rewet_hysteresis", c_double), ("t_resuspend", c_double), ("beach_dt", c_double),
        ("beach_seed", ctypes.c_uint64),
        ("az", c_int32), ("az_const", c_double), ("use_eta", c_int32),
        ("periodic_lon", c_int32), ("periodic_lat", c_int32),
        ("px0", c_double), ("px1", c_double), ("py0", c_double), ("py1", c_double),
        ("reuse_anchor", c_int32),
        ("sig_start_a", c_int32), ("sig_start_b", c_int32), ("sig_start_t", c_int32),
        ("sig_start_alpha0", c_double), ("sig_start_alpha", c_double),
        ("sig_end_a", c_int32), ("sig_end_b", c_int32), ("sig_end_t", c_int32),
        ("sig_end_alpha0", c_double), ("sig_end_alpha", c_double),
        ("Dmin", c_double), ("z0b_min", c_double), ("walk", c_int32),
        ("az_field", c_int32), ("az_slot_a", c_int32), ("az_slot_b", c_int32), ("az_alpha0", c_double),
        ("rk45_tol_m", c_double), ("rk45_h_min", c_double), ("rk45_inv_dtn", c_double),
        ("block_size", c_int32),
    ]


class Cloud(ctypes.Structure):
    """kelp_cloud: one cloud of a batch."""
    _fields_ = [
        ("flags", c_int32), ("n_steps", c_int32), ("start_step", c_int64),
        ("time_dir", c_double), ("t0", c_double), ("kh_x", c_double), ("kh_y", c_double),
        ("seed", ctypes.c_uint64), ("real_id", ctypes.c_uint64),
    ]


class BatchAxis(ctypes.Structure):
    """kelp_batch_axis: the snapshot axes the clouds of a batch read."""
    _fields_ = [
        ("slot0_3d", c_int32), ("n_snap_3d", c_int32), ("slot0_2d", c_int32), ("n_snap_2d", c_int32),
        ("snap_dt", c_double),
    ]


class RecordRows(ctypes.Structure):
    """kelp_record_rows: device rows of one staged record."""
    _fields_ = [
        ("lon", c_void_p), ("lat", c_void_p), ("z", c_void_p), ("u", c_void_p), ("v", c_void_p),
        ("w", c_void_p), ("status", c_void_p), ("fr", c_void_p), ("samples", c_void_p),
    ]


CLOUD_SHIFT = 40
CLOUD_VALID, CLOUD_3D, CLOUD_KICK = 1, 2, 4
MAX_CLOUDS = 65536


#: every symbol include/kelp_hip.h declares, with its (restype, argtypes)
SIGNATURES = {
    "kelp_last_error": (ctypes.c_char_p, []),
    "kelp_version": (c_int, []),
    "kelp_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
    "kelp_ctx_destroy": (c_int, [c_void_p]),
    "kelp_synchronize": (c_int, [c_void_p]),
    "kelp_ctx_stream": (c_void_p, [c_void_p]),
    "kelp_mesh_upload": (c_int, [c_void_p, POINTER(MeshDesc)]),
    "kelp_mesh_node_xyz": (c_int, [c_void_p, c_void_p]),
    "kelp_nodes_to_mesh_ids": (c_int, [c_void_p, c_int64, c_void_p]),
    "kelp_elems_to_mesh_ids": (c_int, [c_void_p, c_int64, c_void_p]),
    "kelp_elems_from_mesh_ids": (c_int, [c_void_p, c_int64, c_void_p]),
    "kelp_nodes_from_mesh_ids": (c_int, [c_void_p, c_int64, c_void_p]),
    "kelp_field_upload": (c_int, [c_void_p, c_int, c_int, c_int64, c_void_p, c_void_p,
                                  c_void_p, c_int64]),
    "kelp_field_upload_elem": (c_int, [c_void_p, c_int, c_int, c_int64, c_void_p, c_void_p,
                                       c_void_p, c_int64]),
    "kelp_field_upload_eta": (c_int, [c_void_p, c_int, c_void_p]),
    "kelp_sigma_surface_upload": (c_int, [c_void_p, c_int, c_void_p]),
    "kelp_field_clear": (c_int, [c_void_p, c_int]),
    "kelp_scalar_upload": (c_int, [c_void_p, c_int, c_int, c_int64, c_void_p]),
    "kelp_sample": (c_int, [c_void_p, POINTER(Particles), c_int, c_int, c_int, c_double, c_int,
                            c_void_p]),
    "kelp_activate": (c_int, [c_void_p, POINTER(Particles), c_void_p, c_int64, c_int64, c_int64,
                              c_void_p]),
    "kelp_record_stage": (c_int, [c_void_p, POINTER(Particles), c_void_p, c_int64, c_int64, c_double, c_int,
                                  c_void_p, POINTER(RecordRows)]),
    "kelp_field_upload_async": (c_int, [c_void_p, c_int, c_int, c_int, c_int64, c_void_p, c_void_p,
                                        c_void_p, c_int64, c_void_p]),
    "kelp_field_upload_wait": (c_int, [c_void_p, c_int]),
    "kelp_field_reconstruct": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_double]),
    "kelp_field_upload_az": (c_int, [c_void_p, c_int, c_int64, c_void_p]),
    "kelp_sample_vars": (c_int, [c_void_p, POINTER(Particles), c_int, c_void_p, c_int, c_int,
                                 c_double, c_int, c_void_p]),
    "kelp_step": (c_int, [c_void_p, POINTER(Particles), POINTER(StepArgs)]),
    "kelp_sort_particles": (c_int, [c_void_p, POINTER(Particles)]),
    "kelp_sort_particles_noalloc": (c_int, [c_void_p, POINTER(Particles)]),
    "kelp_nearest": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int, c_void_p]),
    "kelp_containing": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "kelp_malloc": (c_int, [c_void_p, c_int64, POINTER(c_void_p)]),
    "kelp_free": (c_int, [c_void_p, c_void_p]),
    "kelp_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_void_p, c_int64]),
    "kelp_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_void_p, c_int64]),
    "kelp_event_record": (c_int, [c_void_p, c_int]),
    "kelp_event_elapsed": (c_int, [c_void_p, c_int, c_int, POINTER(c_float)]),
    "kelp_math_probe": (c_int, [c_void_p, c_int, c_int64, c_void_p, c_void_p, c_void_p]),
    "kelp_batch_set_clouds": (c_int, [c_void_p, c_int64, c_int64, c_void_p, POINTER(BatchAxis)]),
    "kelp_gh_sample_stats": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                     c_double, c_double, c_int64, c_int64, c_int, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "kelp_batch_reset": (c_int, [c_void_p, c_int64]),
    "kelp_batch_set_step": (c_int, [c_void_p, c_int64]),
    "kelp_batch_get_step": (c_int, [c_void_p, POINTER(c_int64)]),
    "kelp_batch_step": (c_int, [c_void_p, POINTER(Particles), POINTER(StepArgs)]),
    "kelp_batch_run_graph": (c_int, [c_void_p, POINTER(Particles), POINTER(StepArgs), c_int]),
    "kelp_batch_collect": (c_int, [c_void_p, POINTER(Particles), c_int64, c_int64, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_int]),
}

_LIB = None


def load(path: str = "") -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it has not been built.
    ``KELP_HIP_LIB`` overrides the in-tree path (alternative builds)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or os.environ.get("KELP_HIP_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()' or python -m kelp_amd.build)")
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64 /
    # libhsa-runtime64 (sonames libamdhip64.so.7 / libhsa-runtime64.so.1).
    # Loading torch first lets this library's NEEDED entries resolve to
    # those already-loaded copies instead of a second runtime from
    # /opt/rocm, which would fail to see the device once torch holds it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != KELP_OK:
        msg = load().kelp_last_error().decode(errors="replace")
        raise KelpError(f"{what or 'kelp call'} failed ({rc}): {msg}")
