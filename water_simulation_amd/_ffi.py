"""ctypes declarations of the libwsim C ABI (include/wsim.h).

The library is loaded from water_simulation_amd/lib/libwsim.so.  There is no
fallback: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libwsim.so"

F3 = C.c_float * 3
I3 = C.c_int32 * 3

WSIM_OK = 0
WSIM_ERR_RUNTIME = 1
WSIM_ERR_OUT_OF_RANGE = 2
WSIM_ERR_INVALID_ARGUMENT = 3
WSIM_ERR_DEVICE = 4


class DFSPHConfigC(C.Structure):
    _fields_ = [
        ("rest_density", C.c_float),
        ("particle_radius", C.c_float),
        ("smoothing_length", C.c_float),
        ("viscosity", C.c_float),
        ("damping", C.c_float),
        ("surface_tension", C.c_float),
        ("gravity", C.c_float * 3),
        ("max_density_iters", C.c_uint32),
        ("max_divergence_iters", C.c_uint32),
        ("eta_density", C.c_float),
        ("eta_divergence", C.c_float),
        ("domain_min", C.c_float * 3),
        ("domain_max", C.c_float * 3),
        ("enable_divergence_solve", C.c_uint32),
        ("pressure_form", C.c_uint32),
        ("enable_surface_tension", C.c_uint32),
        ("fast_math", C.c_uint32),
    ]


class StepStatsC(C.Structure):
    _fields_ = [
        ("density_iters", C.c_uint32),
        ("divergence_iters", C.c_uint32),
        ("mean_density_error", C.c_double),
        ("mean_divergence_error", C.c_double),
        ("max_velocity", C.c_float),
        ("nan_count", C.c_uint32),
        ("substeps", C.c_uint64),
        ("escapees", C.c_uint32),
        ("reserved", C.c_uint32),
        ("total_mass", C.c_double),
        ("momentum", C.c_double * 3),
        ("kinetic_energy", C.c_double),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            if name == "reserved":
                continue
            v = getattr(self, name)
            out[name] = list(v) if name == "momentum" else v
        return out


class KernelShapesC(C.Structure):
    _fields_ = [
        ("density_alpha_nt", C.c_uint32),
        ("xsph_nt", C.c_uint32),
        ("change_nt", C.c_uint32),
        ("change_pu", C.c_uint32),
        ("apply_nt", C.c_uint32),
        ("apply_pu", C.c_uint32),
        ("cell_cpt", C.c_uint32),
        ("remap", C.c_uint32),
    ]


class TimeStepConfigC(C.Structure):
    _fields_ = [
        ("frame_dt", C.c_float),
        ("dt_max", C.c_float),
        ("cfl_lambda", C.c_float),
        ("max_substeps", C.c_uint32),
    ]


class SceneC(C.Structure):
    _fields_ = [
        ("name", C.c_char * 256),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("fps", C.c_float),
        ("frame_start", C.c_int32),
        ("frame_end", C.c_int32),
        ("format", C.c_char * 32),
        ("video", C.c_char * 32),
        ("cam_position", C.c_float * 3),
        ("cam_look_at", C.c_float * 3),
        ("cam_up", C.c_float * 3),
        ("cam_fov_y_deg", C.c_float),
        ("solver", C.c_char * 32),
        ("particle_radius", C.c_float),
        ("rest_density", C.c_float),
        ("viscosity", C.c_float),
        ("surface_tension", C.c_float),
        ("gravity", C.c_float * 3),
        ("initial_block_min", C.c_float * 3),
        ("initial_block_max", C.c_float * 3),
    ]


vp = C.c_void_p
st = C.c_int
sz = C.c_size_t


class MeshParamsC(C.Structure):
    _fields_ = [
        ("particle_radius", C.c_double),
        ("grid_voxel_factor", C.c_double),
        ("smooth_sigma", C.c_double),
        ("iso_factor", C.c_double),
        ("swap_yz", C.c_uint32),
        ("reserved", C.c_uint32),
        ("weights", C.c_void_p),
        ("smooth_pre_iters", C.c_int32),
        ("k_neighbors", C.c_int32),
        ("smooth_post_iters", C.c_int32),
        ("reserved2", C.c_int32),
    ]


class MeshInfoC(C.Structure):
    _fields_ = [
        ("dims", C.c_int32 * 3),
        ("reserved", C.c_int32),
        ("lo", C.c_double * 3),
        ("extent", C.c_double * 3),
        ("spacing", C.c_double * 3),
        ("max_density", C.c_float),
        ("iso", C.c_float),
        ("n_vertices", C.c_uint64),
        ("n_triangles", C.c_uint64),
    ]
u32p = C.POINTER(C.c_uint32)

# name -> (restype, argtypes); every symbol here is declared in include/wsim.h
SIGNATURES = {
    "wsim_last_error": (C.c_char_p, []),
    "wsim_abi_version": (C.c_int, []),
    "wsim_device_count": (st, [C.POINTER(C.c_int)]),
    "wsim_device_alloc": (st, [sz, C.POINTER(vp)]),
    "wsim_device_free": (st, [vp]),
    "wsim_memcpy_h2d": (st, [vp, vp, sz, vp]),
    "wsim_memcpy_d2h": (st, [vp, vp, sz, vp]),
    "wsim_memcpy_d2d": (st, [vp, vp, sz, vp]),
    "wsim_memset": (st, [vp, C.c_int, sz, vp]),
    "wsim_device_copy": (st, [vp, vp, sz, vp]),
    "wsim_pair_ops_check": (st, [C.c_float, C.POINTER(C.c_uint64)]),
    "wsim_mesh_params_default": (None, [C.POINTER(MeshParamsC)]),
    "wsim_mesher_create": (st, [vp, C.POINTER(vp)]),
    "wsim_mesher_destroy": (None, [vp]),
    "wsim_mesher_run": (st, [vp, vp, sz, C.POINTER(MeshParamsC)]),
    "wsim_mesher_info": (st, [vp, C.POINTER(MeshInfoC)]),
    "wsim_mesher_grid": (st, [vp, C.POINTER(vp)]),
    "wsim_mesher_vertices": (st, [vp, C.POINTER(vp)]),
    "wsim_mesher_triangles": (st, [vp, C.POINTER(vp)]),
    "wsim_mc_table": (st, [vp, vp]),
    "wsim_stream_create": (st, [C.POINTER(vp)]),
    "wsim_stream_destroy": (st, [vp]),
    "wsim_stream_synchronize": (st, [vp]),
    "wsim_device_synchronize": (st, []),
    "wsim_store_create": (st, [sz, vp, C.POINTER(vp)]),
    "wsim_store_destroy": (None, [vp]),
    "wsim_store_resize": (st, [vp, sz]),
    "wsim_store_count": (sz, [vp]),
    "wsim_store_capacity": (sz, [vp]),
    "wsim_store_register_attribute": (st, [vp, C.c_char_p, C.c_int, C.POINTER(sz)]),
    "wsim_store_has_attribute": (C.c_int, [vp, C.c_char_p]),
    "wsim_store_find_attribute": (st, [vp, C.c_char_p, C.POINTER(sz)]),
    "wsim_store_attribute_data": (st, [vp, sz, C.POINTER(vp)]),
    "wsim_store_positions": (st, [vp, C.POINTER(vp)]),
    "wsim_store_velocities": (st, [vp, C.POINTER(vp)]),
    "wsim_store_position_handle": (sz, [vp]),
    "wsim_store_velocity_handle": (sz, [vp]),
    "wsim_grid_create": (st, [F3, I3, C.c_float, vp, C.POINTER(vp)]),
    "wsim_grid_destroy": (None, [vp]),
    "wsim_grid_reserve": (st, [vp, sz]),
    "wsim_grid_build": (st, [vp, vp, sz]),
    "wsim_cell_sort_noalloc": (st, [vp, sz, F3, I3, C.c_float, vp, vp, vp, vp]),
    "wsim_grid_cell_start": (vp, [vp]),
    "wsim_grid_sorted_indices": (vp, [vp]),
    "wsim_grid_sorted_keys": (vp, [vp]),
    "wsim_grid_origin": (None, [vp, F3]),
    "wsim_grid_cells_per_axis": (None, [vp, I3]),
    "wsim_grid_cell_length": (C.c_float, [vp]),
    "wsim_grid_total_cells": (C.c_uint32, [vp]),
    "wsim_grid_count": (sz, [vp]),
    "wsim_accel_create": (st, [C.POINTER(vp)]),
    "wsim_accel_destroy": (None, [vp]),
    "wsim_accel_build": (st, [vp, vp, sz, F3, F3, vp]),
    "wsim_accel_leaf_count": (sz, [vp]),
    "wsim_accel_sorted_indices": (vp, [vp]),
    "wsim_accel_morton_codes": (vp, [vp]),
    "wsim_accel_internal_left": (vp, [vp]),
    "wsim_accel_internal_right": (vp, [vp]),
    "wsim_accel_internal_parent": (vp, [vp]),
    "wsim_accel_node_aabbs": (vp, [vp]),
    "wsim_dfsph_config_default": (None, [C.POINTER(DFSPHConfigC)]),
    "wsim_solver_create": (st, [vp, vp, C.POINTER(DFSPHConfigC), vp, C.POINTER(vp)]),
    "wsim_solver_destroy": (None, [vp]),
    "wsim_solver_step": (st, [vp, C.c_float]),
    "wsim_solver_max_velocity": (C.c_float, [vp]),
    "wsim_solver_particle_mass": (C.c_float, [vp]),
    "wsim_solver_last_density_iters": (C.c_uint32, [vp]),
    "wsim_solver_density_at": (st, [vp, sz, C.POINTER(C.c_float)]),
    "wsim_solver_alpha_at": (st, [vp, sz, C.POINTER(C.c_float)]),
    "wsim_solver_boundary_mass_at": (st, [vp, sz, C.POINTER(C.c_float)]),
    "wsim_solver_stats": (st, [vp, C.POINTER(StepStatsC)]),
    "wsim_solver_neighbor_counts": (st, [vp, vp, vp, vp, vp]),
    "wsim_solver_fluid_grid": (vp, [vp]),
    "wsim_solver_boundary_grid": (vp, [vp]),
    "wsim_solver_set_timing": (st, [vp, C.c_int]),
    "wsim_solver_reset_timing": (st, [vp]),
    "wsim_solver_kernel_kinds": (C.c_int, []),
    "wsim_solver_kernel_name": (C.c_char_p, [C.c_int]),
    "wsim_solver_kernel_time": (st, [vp, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "wsim_solver_get_shapes": (st, [vp, C.POINTER(KernelShapesC)]),
    "wsim_solver_set_shapes": (st, [vp, C.POINTER(KernelShapesC)]),
    "wsim_solver_autotune": (st, [vp, C.c_int]),
    "wsim_solver_tune_count": (sz, [vp]),
    "wsim_solver_tune_entry": (st, [vp, sz, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "wsim_solver_set_graphs": (st, [vp, C.c_int]),
    "wsim_solver_set_viscosity": (st, [vp, C.c_float]),
    "wsim_adjoint_step": (st, [vp, C.c_float, vp, vp, vp, vp, C.POINTER(C.c_double)]),
    "wsim_rollout_create": (st, [vp, C.c_uint32, C.POINTER(vp)]),
    "wsim_rollout_destroy": (None, [vp]),
    "wsim_rollout_forward": (st, [vp, vp, vp, C.POINTER(C.c_float), C.c_uint32]),
    "wsim_rollout_backward": (st, [vp, vp, vp, vp, vp, C.POINTER(C.c_double)]),
    "wsim_batch_create": (st, [vp, vp, C.POINTER(DFSPHConfigC), C.c_uint32,
                               C.POINTER(C.c_uint32), C.POINTER(C.c_float), vp, C.POINTER(vp)]),
    "wsim_batch_step": (st, [vp, C.POINTER(C.c_float)]),
    "wsim_batch_scene_count": (C.c_uint32, [vp]),
    "wsim_batch_scene_offsets": (st, [vp, C.POINTER(C.c_uint32)]),
    "wsim_batch_scene_stats": (st, [vp, C.c_uint32, C.POINTER(StepStatsC)]),
    "wsim_sample_aabb_boundary": (st, [F3, F3, C.c_float, C.c_int, vp, sz, C.POINTER(sz)]),
    "wsim_timestep_config_default": (None, [C.POINTER(TimeStepConfigC)]),
    "wsim_solver_advance": (st, [vp, C.POINTER(TimeStepConfigC), C.c_float, C.c_float,
                                 C.c_uint32, C.POINTER(C.c_float), C.POINTER(StepStatsC),
                                 C.POINTER(C.c_uint32)]),
    "wsim_timestep_next_dt": (C.c_float, [C.POINTER(TimeStepConfigC), C.c_float, C.c_float,
                                          C.c_float]),
    "wsim_lattice_block": (st, [F3, F3, C.c_float, vp, sz, C.POINTER(sz)]),
    "wsim_write_frame": (st, [C.c_char_p, vp, sz]),
    "wsim_scene_load": (st, [C.c_char_p, C.POINTER(SceneC)]),
}

_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """Load libwsim once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (there is no non-HIP fallback)")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class WsimError(RuntimeError):
    """A libwsim call failed (reference: std::runtime_error)."""


class OutOfRange(IndexError):
    """Reference: std::out_of_range."""


class InvalidArgument(ValueError):
    """Invalid argument rejected by libwsim."""


class DeviceError(WsimError):
    """HIP runtime failure (reference: CUDA error -> std::runtime_error)."""


def check(status: int) -> None:
    if status == WSIM_OK:
        return
    msg = lib().wsim_last_error().decode(errors="replace")
    if status == WSIM_ERR_OUT_OF_RANGE:
        raise OutOfRange(msg)
    if status == WSIM_ERR_INVALID_ARGUMENT:
        raise InvalidArgument(msg)
    if status == WSIM_ERR_DEVICE:
        raise DeviceError(msg)
    raise WsimError(msg)
