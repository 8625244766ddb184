"""ctypes binding of libfluidsim_hip.so (the C-ABI declared in include/fluidsim.h).

The library is required: importing the solvers works without it, but any call
that needs the GPU path raises ``RuntimeError`` when the library is missing or
no HIP device is present -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int32, c_int64, c_size_t, c_uint32, c_uint64, c_void_p

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfluidsim_hip.so")

FS_OK = 0
FS_ERR_ARG = -1
FS_ERR_CAPACITY = -2
FS_ERR_HIP = -3
FS_ERR_WORKSPACE = -4


class FsGrid(ctypes.Structure):
    _fields_ = [
        ("dim_x", c_int32), ("dim_y", c_int32), ("dim_z", c_int32),
        ("round_mode", c_int32), ("capacity", c_int32), ("num_points", c_int32),
        ("cell_width", c_float), ("cell_width_inv", c_float),
        ("point_ids", c_void_p), ("point_cells", c_void_p), ("cell_of", c_void_p),
        ("cell_start", c_void_p), ("cell_count", c_void_p), ("rank", c_void_p),
        ("scatter_ids", c_void_p), ("scan_state", c_void_p), ("sync", c_void_p),
    ]


class FsModel(ctypes.Structure):
    _fields_ = [
        ("n", c_int32), ("num_worlds", c_int32),
        ("flags", c_void_p), ("mask", c_void_p), ("world", c_void_p),
        ("mass", c_void_p), ("gravity", c_void_p), ("rest_density", c_void_p),
        ("surface_tension", c_void_p), ("viscosity", c_void_p), ("stiffness", c_void_p), ("exponent", c_void_p),
    ]


FS_CONTROL_LIST_NEED = 7   # fs_work::control word: longest overflowed list needed


class FsWork(ctypes.Structure):
    _fields_ = [
        ("capacity", c_int32), ("list_capacity", c_int32),
        ("pos", c_void_p), ("pos_alt", c_void_p), ("vel", c_void_p), ("sca", c_void_p), ("wvol", c_void_p),
        ("flags", c_void_p), ("x_pred", c_void_p), ("v_pred", c_void_p), ("nbr", c_void_p),
        ("nbr_count", c_void_p), ("partials", c_void_p), ("control", c_void_p), ("sync", c_void_p),
        ("packed", c_void_p),
    ]


class FsWcsphParams(ctypes.Structure):
    _fields_ = [
        ("smoothing_length", c_float), ("particle_radius", c_float), ("bound", c_float * 3),
        ("boundary_damping", c_float), ("dt", c_float), ("clamp_negative_pressure", c_int32),
        ("dt_device", c_void_p),
    ]


class FsWcsphIO(ctypes.Structure):
    _fields_ = [
        ("x_in", c_void_p), ("v_in", c_void_p), ("x_out", c_void_p), ("v_out", c_void_p),
        ("rho", c_void_p), ("pressure", c_void_p), ("accel", c_void_p), ("rho_in", c_void_p),
        ("pressure_in", c_void_p),
    ]


class FsWcsphAdjointIO(ctypes.Structure):
    _fields_ = [
        ("forward", FsWcsphIO), ("xbar_out", c_void_p), ("vbar_out", c_void_p), ("xbar_in", c_void_p),
        ("vbar_in", c_void_p), ("grad_stiffness", c_void_p), ("grad_viscosity", c_void_p), ("scratch", c_void_p),
    ]


class FsDfsphParams(ctypes.Structure):
    _fields_ = [
        ("eta_density", c_double), ("eta_divergence", c_double),
        ("smoothing_length", c_float), ("particle_radius", c_float), ("bound", c_float * 3),
        ("boundary_damping", c_float), ("dt", c_float), ("max_kappa", c_float), ("max_kappa_v", c_float),
        ("max_rho_iterations", c_int32), ("max_vel_iterations", c_int32), ("warm_start", c_int32),
        ("min_neighbors_for_factor", c_int32), ("exact_residual", c_int32), ("dt_device", c_void_p),
    ]


class FsDfsphIO(ctypes.Structure):
    _fields_ = [
        ("x_in", c_void_p), ("v_in", c_void_p), ("kappa_in", c_void_p), ("kappa_v_in", c_void_p),
        ("x_out", c_void_p), ("v_out", c_void_p), ("rho", c_void_p), ("factor", c_void_p),
        ("rho_star", c_void_p), ("d_rho_dt", c_void_p), ("kappa", c_void_p), ("kappa_v", c_void_p),
        ("neighbor_count", c_void_p), ("v_star", c_void_p), ("accel", c_void_p), ("iterations", c_void_p),
        ("rho_in", c_void_p), ("factor_in", c_void_p), ("rho_star_in", c_void_p), ("d_rho_dt_in", c_void_p),
        ("neighbor_count_in", c_void_p), ("v_star_in", c_void_p),
    ]


class FsPbfParams(ctypes.Structure):
    _fields_ = [
        ("smoothing_length", c_float), ("particle_radius", c_float), ("bound", c_float * 3),
        ("boundary_damping", c_float), ("boundary_epsilon", c_float), ("dt", c_float),
        ("lambda_epsilon", c_float), ("scorr_k", c_float), ("scorr_n", c_float), ("scorr_w_delta_q", c_float),
        ("max_delta_position", c_float), ("xsph_c", c_float),
        ("max_iterations", c_int32), ("min_neighbors_for_lambda", c_int32), ("rebuild_grid_each_iteration", c_int32),
    ]


class FsPbfIO(ctypes.Structure):
    _fields_ = [
        ("x_in", c_void_p), ("v_in", c_void_p), ("x_out", c_void_p), ("v_out", c_void_p),
        ("rho", c_void_p), ("pbf_lambda", c_void_p), ("neighbor_count", c_void_p), ("delta_p", c_void_p),
        ("accel", c_void_p),
    ]


class FsSeparationParams(ctypes.Structure):
    _fields_ = [
        ("rest_dist", c_float), ("query_radius", c_float), ("strength", c_float), ("wall_band", c_float),
        ("wall_margin", c_float), ("bound_lo", c_float * 3), ("bound_hi", c_float * 3), ("iterations", c_int32),
    ]


class FsScene(ctypes.Structure):
    _fields_ = [
        ("particle_offset", c_int32), ("particle_count", c_int32), ("bucket_base", c_int32), ("dims", c_int32 * 3),
        ("smoothing_length", c_float), ("cell_width_inv", c_float), ("particle_radius", c_float),
        ("bound", c_float * 3), ("boundary_damping", c_float), ("dt", c_float),
        ("clamp_negative_pressure", c_int32), ("pad_", c_int32), ("cfl_smoothing_length", c_double),
    ]


class FsBatchIO(ctypes.Structure):
    _fields_ = [
        ("x_in", c_void_p), ("v_in", c_void_p), ("x_out", c_void_p), ("v_out", c_void_p),
        ("rho", c_void_p), ("pressure", c_void_p), ("accel", c_void_p), ("scene_of", c_void_p),
        ("scene_sorted", c_void_p),
    ]


# name -> (restype, argtypes); every symbol declared in include/fluidsim.h
SIGNATURES = {
    "fs_last_error": (ctypes.c_char_p, []),
    "fs_version": (c_int32, []),
    "fs_profile_enable": (None, [c_int32]),
    "fs_profile_reset": (None, []),
    "fs_profile_report": (c_int32, [ctypes.c_char_p, c_size_t]),
    "fs_tune_variant_count": (c_int32, [c_int32]),
    "fs_tune_describe": (c_int32, [c_int32, c_int32, POINTER(c_int32), POINTER(c_int32)]),
    "fs_tune_describe_ex": (c_int32, [c_int32, c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32),
                                      POINTER(c_int32), POINTER(c_int32)]),
    "fs_tune_variant_coop": (c_int32, [c_int32, c_int32]),
    "fs_tune_select": (c_int32, [c_int32, c_int32]),
    "fs_tune_selected": (c_int32, [c_int32]),
    "fs_check_fast_math": (c_int32, [c_int32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                     POINTER(ctypes.c_uint64), c_void_p]),
    "fs_grid_workspace_bytes": (c_size_t, [c_int32, c_int32, c_int32, c_int32]),
    "fs_grid_init": (c_int32, [POINTER(FsGrid), c_void_p, c_size_t, c_int32, c_int32, c_int32, c_int32, c_int32,
                               c_void_p]),
    "fs_grid_build": (c_int32, [POINTER(FsGrid), c_void_p, c_int32, c_float, c_void_p]),
    "fs_grid_export_tables": (c_int32, [POINTER(FsGrid), c_void_p, c_void_p, c_void_p]),
    "fs_grid_count_candidates": (c_int32, [POINTER(FsGrid), c_void_p, c_float, c_int32, c_void_p, c_void_p]),
    "fs_grid_list_candidates": (c_int32, [POINTER(FsGrid), c_void_p, c_float, c_int32, c_void_p, c_void_p,
                                          c_void_p]),
    "fs_work_bytes": (c_size_t, [c_int32, c_int32]),
    "fs_work_init": (c_int32, [POINTER(FsWork), c_void_p, c_size_t, c_int32, c_int32, c_void_p]),
    "fs_wcsph_step": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsWcsphParams),
                                POINTER(FsWcsphIO), c_void_p]),
    "fs_wcsph_adjoint_bytes": (c_size_t, [c_int32]),
    "fs_wcsph_adjoint_step": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsWcsphParams),
                                        POINTER(FsWcsphAdjointIO), c_size_t, c_void_p]),
    "fs_dfsph_step": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsDfsphParams),
                                POINTER(FsDfsphIO), c_void_p]),
    "fs_pbf_step": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsPbfParams),
                              POINTER(FsPbfIO), c_void_p]),
    "fs_wcsph_density_eos": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsWcsphParams),
                                       POINTER(FsWcsphIO), c_void_p]),
    "fs_wcsph_force_integrate": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel),
                                           POINTER(FsWcsphParams), POINTER(FsWcsphIO), c_void_p]),
    "fs_dfsph_density_factor": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsDfsphParams),
                                          POINTER(FsDfsphIO), c_int32, c_void_p]),
    "fs_dfsph_nonpressure_predict": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel),
                                               POINTER(FsDfsphParams), POINTER(FsDfsphIO), c_void_p]),
    "fs_dfsph_rate_kappa": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsDfsphParams),
                                      POINTER(FsDfsphIO), c_int32, c_void_p]),
    "fs_dfsph_residual": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsDfsphParams),
                                    POINTER(FsDfsphIO), c_int32, c_int32, c_void_p]),
    "fs_dfsph_correction": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsDfsphParams),
                                      POINTER(FsDfsphIO), c_int32, c_int32, c_void_p]),
    "fs_dfsph_integrate": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsDfsphParams),
                                     POINTER(FsDfsphIO), c_void_p]),
    "fs_dfsph_finalize": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsDfsphParams),
                                    POINTER(FsDfsphIO), c_void_p]),
    "fs_pbf_predict": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsPbfParams),
                                 POINTER(FsPbfIO), c_void_p]),
    "fs_pbf_density": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsPbfParams),
                                 POINTER(FsPbfIO), c_void_p]),
    "fs_pbf_nonpressure": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsPbfParams),
                                     POINTER(FsPbfIO), c_void_p]),
    "fs_pbf_lambda": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsPbfParams),
                                POINTER(FsPbfIO), c_int32, c_void_p]),
    "fs_pbf_delta": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsPbfParams),
                               POINTER(FsPbfIO), c_int32, c_void_p]),
    "fs_pbf_final_density": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsPbfParams),
                                       POINTER(FsPbfIO), c_void_p]),
    "fs_pbf_xsph_finalize": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), POINTER(FsPbfParams),
                                       POINTER(FsPbfIO), c_void_p]),
    "fs_particle_separation": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsSeparationParams), c_void_p,
                                         c_void_p, c_void_p, c_int32, c_void_p]),
    "fs_batch_wcsph_step": (c_int32, [POINTER(FsGrid), POINTER(FsWork), POINTER(FsModel), c_void_p, c_int32,
                                      POINTER(FsBatchIO), c_void_p]),
    "fs_batch_cfl_dt": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_int32, c_double, c_double,
                                  c_double, c_double, c_double, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fs_dfsph_cfl_dt": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_double, c_double, c_int32, c_void_p,
                                  c_void_p, c_void_p, c_void_p]),
    "fs_batch_cfl_dt_noalloc": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_int32, c_double,
                                          c_double, c_double, c_double, c_double, c_int32, c_void_p, c_void_p,
                                          c_void_p]),
    "fs_norm_maxima": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p]),
    "fs_check_finite": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p]),
}

_lib = None


def load(path: str | None = None):
    """Load the shared library once (raises RuntimeError if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m fluid_simulation_amd.build` "
                           "(the fluid solvers have no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != FS_OK:
        msg = load().fs_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def require_device(device) -> None:
    """The HIP path needs a GPU tensor device; fail loudly otherwise."""
    import torch

    if not torch.cuda.is_available() or str(device).startswith("cpu"):
        raise RuntimeError(f"fluid_simulation_amd solvers run on a HIP device only (got device={device!r}); "
                           "there is no CPU fallback")
    load()


def profile_report() -> dict:
    """{kernel: (launches, total_ms)} accumulated while profiling was enabled."""
    lib = load()
    n = lib.fs_profile_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib.fs_profile_report(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        name, count, ms = line.rsplit(" ", 2)
        out[name] = (int(count), float(ms))
    return out


def stream_handle(device=None) -> int:
    import torch

    return torch.cuda.current_stream(device).cuda_stream
