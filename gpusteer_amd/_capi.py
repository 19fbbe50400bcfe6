"""ctypes binding of libgpusteer's C-ABI (include/gpusteer.h).

The library is loaded from gpusteer_amd/lib/libgpusteer.so (built in-tree by
gpusteer_amd/build.py).  There is no fallback: if the library is missing the
import of this module raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libgpusteer.so"

GS_OK = 0
GS_ERR_INVALID_ARGUMENT = 1
GS_ERR_INVALID_HANDLE = 2
GS_ERR_SIZE_MISMATCH = 3
GS_ERR_UNSUPPORTED = 4
GS_ERR_NOT_BINNED = 5
GS_ERR_OUT_OF_MEMORY = 6
GS_ERR_DEVICE = 7
GS_ERR_NONFINITE = 8
GS_ERR_FRAME = 9
GS_SHAPE_KNN = 0
GS_SHAPE_FLOCK = 1
GS_MAX_K = 32
GS_MAX_CELLS = 1 << 28
GS_MAX_SPLIT_CELLS = 1 << 24
# Limits of gs_find_knn_noalloc (its fixed device array).
GS_NOALLOC_MAX_POINTS = 65536
GS_NOALLOC_MAX_CELLS = 131072
GS_NOALLOC_MAX_K = 16

GS_MODE_OPENSTEER = 0
GS_MODE_GPUREF = 1
GS_KIND_AGENTS, GS_KIND_OBSTACLES, GS_KIND_WALLS = 0, 1, 2
GS_WRAP_NONE, GS_WRAP_BOX, GS_WRAP_SPHERE = -1, 0, 1
GS_KNN_DEFAULT, GS_KNN_NO_SEED, GS_KNN_BINNED = 0, 1, 2
GS_CLOSEST_CLAMPED, GS_CLOSEST_AS_REFERENCE = 0, 1

# Behaviour bits (include/OpenSteer/CUDA/CUDAKernelGlobals.cuh:12-31).
BIT_PURSUE = 1 << 0
BIT_EVADE = 1 << 1
BIT_SEEK = 1 << 2
BIT_FLEE = 1 << 3
BIT_SEPARATION = 1 << 4
BIT_ALIGNMENT = 1 << 5
BIT_COHESION = 1 << 6
BIT_AVOID_OBSTACLES = 1 << 7
BIT_AVOID_WALLS = 1 << 8
BIT_AVOID_NEIGHBORS = 1 << 9
BIT_AVOID_CLOSE_NEIGHBORS = 1 << 10
BIT_FOLLOW_PATH = 1 << 11
BIT_ANTI_PENETRATION_WALL = 1 << 12
BIT_ANTI_PENETRATION_AGENT = 1 << 13

# Order of the float arrays in gs_agents.
AGENT_FIELDS = (
    "px", "py", "pz", "fx", "fy", "fz", "speed", "ax", "ay", "az",
    "sx", "sy", "sz", "ux", "uy", "uz", "gx", "gy", "gz",
    "radius", "max_speed", "max_force", "mass",
)
WALL_FIELDS = ("sx", "sy", "sz", "ex", "ey", "ez", "nx", "ny", "nz")

_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)


class GsAgents(C.Structure):
    _fields_ = [(f, _fp) for f in AGENT_FIELDS] + [("bits", _up)]


class GsWalls(C.Structure):
    _fields_ = [(f, _fp) for f in WALL_FIELDS]


class GsFlockTerm(C.Structure):
    _fields_ = [("min_d", C.c_float), ("max_d", C.c_float), ("cos_max", C.c_float),
                ("weight", C.c_float), ("mask", C.c_uint32)]


class GsBoidsParams(C.Structure):
    _fields_ = [
        ("wrap_shape", C.c_int),
        ("wrap_center", C.c_float * 3),
        ("wrap_size", C.c_float * 3),
        ("search_radius", C.c_uint32),
        ("knn_flags", C.c_uint32),
        ("separation", GsFlockTerm),
        ("alignment", GsFlockTerm),
        ("cohesion", GsFlockTerm),
        ("dt", C.c_float),
        ("mode", C.c_int),
    ]


class GsBoidsTiming(C.Structure):
    _fields_ = [("bin_build_ms", C.c_float), ("knn_ms", C.c_float), ("steer_integrate_ms", C.c_float),
                ("step_ms", C.c_float)]


GS_SCENARIO_BOIDS, GS_SCENARIO_CHOKE = 0, 1
GS_ENSEMBLE_GRAPH, GS_ENSEMBLE_EAGER = 0, 1


class GsScenario(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("n", C.c_uint32 * 2),
        ("origin", C.c_float * 3),
        ("size", C.c_float * 3),
        ("cells", C.c_uint32 * 3),
        ("dt", C.c_float),
        ("mode", C.c_int),
        ("wrap_shape", C.c_int),
        ("wrap_center", C.c_float * 3),
        ("wrap_size", C.c_float * 3),
        ("separation", GsFlockTerm),
        ("alignment", GsFlockTerm),
        ("cohesion", GsFlockTerm),
        ("choke_separation", GsFlockTerm),
        ("goal", (C.c_float * 3) * 2),
        ("seek_weight", C.c_float),
        ("seek_mask", C.c_uint32),
        ("walls_min_time", C.c_float),
        ("walls_weight", C.c_float),
        ("obstacles_min_time", C.c_float),
        ("obstacles_weight", C.c_float),
        ("k_walls", C.c_uint32),
        ("k_obstacles", C.c_uint32),
        ("other_max_distance", C.c_float),
    ]


class GsCalibParams(C.Structure):
    _fields_ = [(f, C.c_float * 2) for f in ("w_sep", "w_ali", "w_coh", "w_obs", "w_wall", "radius", "max_speed")]


class GsCalibConfig(C.Structure):
    _fields_ = [
        ("separation", GsFlockTerm),
        ("alignment", GsFlockTerm),
        ("cohesion", GsFlockTerm),
        ("seek_weight", C.c_float),
        ("seek_mask", C.c_uint32),
        ("walls_min_time", C.c_float),
        ("obstacles_min_time", C.c_float),
        ("max_force", C.c_float),
        ("mass", C.c_float),
        ("dt", C.c_float),
        ("penetration_weight", C.c_float),
        ("steps", C.c_uint32),
        ("checkpoint_every", C.c_uint32),
    ]


GS_PARAMS_MAX_GROUPS = 32
GS_PARAMS_MAX_PATH = 64


class GsGroupParams(C.Structure):
    _fields_ = [
        ("num_agents", C.c_uint32),
        ("max_speed", C.c_float),
        ("max_force", C.c_float),
        ("start_position", C.c_float * 3),
        ("min_start_radius", C.c_float),
        ("max_start_radius", C.c_float),
        ("path_points", C.c_uint32),
        ("path", (C.c_float * 3) * GS_PARAMS_MAX_PATH),
        ("path_radius", C.c_float),
        ("path_cyclic", C.c_uint32),
        ("body_color", C.c_float * 3),
    ]


class GsWorldParams(C.Structure):
    _fields_ = [
        ("dimensions", C.c_float * 3),
        ("cells", C.c_uint32 * 3),
        ("obstacle_count", C.c_uint32),
        ("map_filename", C.c_char * 256),
    ]


SIM_FLOAT_FIELDS = ("max_pursuit_prediction_time", "min_time_to_collision", "min_time_to_obstacle",
                    "min_time_to_wall", "path_prediction_time", "min_separation_distance",
                    "max_separation_distance", "min_flocking_distance", "max_flocking_distance",
                    "cos_max_flocking_angle")
SIM_WEIGHT_FIELDS = ("weight_alignment", "weight_cohesion", "weight_separation", "weight_pursuit", "weight_seek",
                     "weight_flee", "weight_evade", "weight_follow_path", "weight_avoid_obstacles",
                     "weight_avoid_walls", "weight_avoid_neighbors")
SIM_MASK_FIELDS = ("mask_alignment", "mask_anti_penetration_agents", "mask_cohesion", "mask_separation",
                   "mask_pursuit", "mask_seek", "mask_flee", "mask_evade", "mask_follow_path",
                   "mask_avoid_obstacles", "mask_avoid_walls", "mask_avoid_neighbors")


class GsSimParams(C.Structure):
    _fields_ = ([("world", GsWorldParams), ("n_groups", C.c_uint32),
                 ("groups", GsGroupParams * GS_PARAMS_MAX_GROUPS)]
                + [(f, C.c_uint32) for f in ("seed", "knn", "kno", "knw", "search_radius")]
                + [(f, C.c_float) for f in SIM_FLOAT_FIELDS]
                + [("avoid_close_neighbors", C.c_uint32)]
                + [(f, C.c_float) for f in SIM_WEIGHT_FIELDS]
                + [(f, C.c_uint32) for f in SIM_MASK_FIELDS])


class GpuSteerError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        self.status = status
        super().__init__(f"{where}: status {status} ({detail})")


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(or python gpusteer_amd/build.py)")
    return C.CDLL(str(LIB_PATH))


lib = _load()
_h = C.c_void_p
_f3 = C.c_float * 3
_u3 = C.c_uint32 * 3

_SIGS = {
    "gs_status_string": (C.c_char_p, [C.c_int]),
    "gs_last_error": (C.c_char_p, []),
    "gs_version": (C.c_int, []),
    "gs_debug_fail_alloc_after": (C.c_int, [C.c_int64]),
    "gs_context_create": (C.c_int, [C.c_int, C.c_void_p, C.POINTER(_h)]),
    "gs_context_destroy": (C.c_int, [_h]),
    "gs_context_synchronize": (C.c_int, [_h]),
    "gs_context_stream": (C.c_int, [_h, C.POINTER(C.c_void_p)]),
    "gs_bins_create": (C.c_int, [_h, _fp, _fp, _up, C.POINTER(_h)]),
    "gs_bins_create_centered": (C.c_int, [_h, _up, _fp, C.POINTER(_h)]),
    "gs_bins_destroy": (C.c_int, [_h]),
    "gs_bins_geometry": (C.c_int, [_h, _fp, _fp, _up]),
    "gs_group_create": (C.c_int, [_h, C.c_int, C.POINTER(_h)]),
    "gs_group_destroy": (C.c_int, [_h]),
    "gs_group_size": (C.c_int, [_h, _up]),
    "gs_group_upload_agents": (C.c_int, [_h, C.c_uint32, C.POINTER(GsAgents)]),
    "gs_group_download_agents": (C.c_int, [_h, C.POINTER(GsAgents)]),
    "gs_group_device_agents": (C.c_int, [_h, C.POINTER(GsAgents)]),
    "gs_group_device_order": (C.c_int, [_h, C.POINTER(_up), C.POINTER(_up)]),
    "gs_group_upload_obstacles": (C.c_int, [_h, C.c_uint32, _fp, _fp, _fp, _fp]),
    "gs_group_upload_walls": (C.c_int, [_h, C.c_uint32, C.POINTER(GsWalls)]),
    "gs_group_download_bins": (C.c_int, [_h, _up, _up, _up, _fp, _fp, _fp]),
    "gs_group_upload_walls_mid": (C.c_int, [_h, C.c_uint32, C.POINTER(GsWalls), _fp, _fp, _fp]),
    "gs_group_upload_wall_set": (C.c_int, [_h, _h]),
    "gs_wall_set_load_map": (C.c_int, [C.c_char_p, C.POINTER(_h)]),
    "gs_wall_set_create": (C.c_int, [C.c_uint32, C.POINTER(GsWalls), C.POINTER(_h)]),
    "gs_wall_set_split": (C.c_int, [_h, _fp, _up]),
    "gs_wall_set_split_cells": (C.c_int, [_h, C.c_uint32, _fp, _fp]),
    "gs_wall_set_size": (C.c_int, [_h, _up]),
    "gs_wall_set_set_mid": (C.c_int, [_h, _fp, _fp, _fp]),
    "gs_wall_set_get": (C.c_int, [_h, C.POINTER(GsWalls), _fp, _fp, _fp]),
    "gs_wall_set_destroy": (C.c_int, [_h]),
    "gs_sim_params_load": (C.c_int, [C.c_char_p, C.POINTER(GsSimParams)]),
    "gs_group_check_finite": (C.c_int, [_h, _up]),
    "gs_knn_create": (C.c_int, [_h, C.c_uint32, C.c_uint32, C.POINTER(_h)]),
    "gs_knn_destroy": (C.c_int, [_h]),
    "gs_knn_shape": (C.c_int, [_h, _up, _up]),
    "gs_knn_download": (C.c_int, [_h, _up, _fp]),
    "gs_knn_device": (C.c_int, [_h, C.POINTER(_up), C.POINTER(_fp)]),
    "gs_update_knn_database": (C.c_int, [_h, _h]),
    "gs_find_knn": (C.c_int, [_h, _h, _h, _h, C.c_uint32, C.c_uint32]),
    "gs_find_knn_within": (C.c_int, [_h, _h, _h, _h, C.c_uint32, C.c_uint32, C.c_float]),
    "gs_find_knn_brute": (C.c_int, [_h, _h, _h]),
    "gs_group_bin_geometry": (C.c_int, [_h, _fp, _fp, _up]),
    # device arrays: px, py, pz (n floats), idx, dist (k * n)
    "gs_find_knn_noalloc": (C.c_int, [_h, _h, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.c_void_p, C.c_void_p]),
    "gs_steer_separation": (C.c_int, [_h, _h, _h, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_int]),
    "gs_steer_alignment": (C.c_int, [_h, _h, _h, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_int]),
    "gs_steer_cohesion": (C.c_int, [_h, _h, _h, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_int]),
    "gs_steer_flocking": (C.c_int, [_h, _h, _h, C.POINTER(GsFlockTerm), C.POINTER(GsFlockTerm),
                                    C.POINTER(GsFlockTerm), C.c_int]),
    "gs_steer_seek": (C.c_int, [_h, _fp, C.c_float, C.c_uint32, C.c_int]),
    "gs_steer_avoid_obstacles": (C.c_int, [_h, _h, _h, C.c_float, C.c_float, C.c_uint32]),
    "gs_steer_avoid_walls": (C.c_int, [_h, _h, _h, C.c_float, C.c_float, C.c_uint32]),
    "gs_update_group": (C.c_int, [_h, C.c_float, C.c_int]),
    "gs_update_group_anti_penetration": (C.c_int, [_h, _h, _h, C.c_float, C.c_int]),
    "gs_anti_penetration_agents": (C.c_int, [_h, _h, _h, C.c_uint32]),
    "gs_path_create": (C.c_int, [_h, C.c_uint32, _fp, C.c_float, C.c_int, C.POINTER(_h)]),
    "gs_path_info": (C.c_int, [_h, _up, _fp]),
    "gs_path_get": (C.c_int, [_h, _fp, _fp, _fp]),
    "gs_path_destroy": (C.c_int, [_h]),
    "gs_steer_follow_path": (C.c_int, [_h, _h, C.c_float, C.c_float, C.c_uint32]),
    "gs_group_cell_density": (C.c_int, [_h, _up]),
    "gs_group_cell_velocity": (C.c_int, [_h, _fp]),
    "gs_write_cell_density": (C.c_int, [_h, C.c_char_p, _up, C.c_uint32, C.c_char_p, C.c_int]),
    "gs_write_cell_velocity": (C.c_int, [_h, C.c_char_p, _up, C.c_uint32, C.c_char_p, C.c_int]),
    "gs_steer_avoid_neighbors": (C.c_int, [_h, _h, _h, C.c_float, C.c_float, C.c_int, C.c_float, C.c_uint32,
                                           C.c_int]),
    "gs_wrap_world": (C.c_int, [_h, C.c_int, _fp, _fp]),
    "gs_clear_bits": (C.c_int, [_h]),
    "gs_step_boids": (C.c_int, [_h, _h, _h, C.POINTER(GsBoidsParams), C.c_uint32]),
    # A0, A1, walls, obstacles, bins, knn[8], scenario (crowd fields), steps
    "gs_step_choke": (C.c_int, [_h, _h, _h, _h, _h, C.POINTER(_h), C.POINTER(GsScenario), C.c_uint32]),
    "gs_context_set_shape": (C.c_int, [_h, C.c_int, C.c_uint32, C.c_uint32]),
    "gs_context_get_shape": (C.c_int, [_h, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "gs_frame_begin": (C.c_int, [_h]),
    "gs_frame_end": (C.c_int, [_h, C.POINTER(_h)]),
    "gs_frame_launch": (C.c_int, [_h, C.c_uint32]),
    "gs_frame_destroy": (C.c_int, [_h]),
    "gs_step_boids_timed": (C.c_int, [_h, _h, _h, C.POINTER(GsBoidsParams), C.c_uint32, C.POINTER(GsBoidsTiming)]),
    "gs_ensemble_create": (C.c_int, [_h, C.c_uint32, C.POINTER(GsScenario), C.c_uint32, C.POINTER(_h)]),
    "gs_ensemble_destroy": (C.c_int, [_h]),
    "gs_ensemble_slots": (C.c_int, [_h, _up]),
    "gs_ensemble_upload_agents": (C.c_int, [_h, C.c_uint32, C.c_uint32, C.POINTER(GsAgents)]),
    "gs_ensemble_download_agents": (C.c_int, [_h, C.c_uint32, C.c_uint32, C.POINTER(GsAgents)]),
    "gs_ensemble_upload_walls": (C.c_int, [_h, C.c_uint32, C.c_uint32, C.POINTER(GsWalls)]),
    "gs_ensemble_upload_walls_mid": (C.c_int, [_h, C.c_uint32, C.c_uint32, C.POINTER(GsWalls), _fp, _fp, _fp]),
    "gs_ensemble_upload_obstacles": (C.c_int, [_h, C.c_uint32, C.c_uint32, _fp, _fp, _fp, _fp]),
    "gs_ensemble_download_knn": (C.c_int, [_h, C.c_uint32, C.c_uint32, _up]),
    "gs_ensemble_step": (C.c_int, [_h, C.c_uint32, C.c_uint32]),
    "gs_calib_create": (C.c_int, [_h, C.POINTER(GsCalibConfig), C.POINTER(_h)]),
    "gs_calib_destroy": (C.c_int, [_h]),
    "gs_calib_forward": (C.c_int, [_h, C.POINTER(GsCalibParams), C.POINTER(C.c_double)]),
    "gs_calib_backward": (C.c_int, [_h, C.POINTER(GsCalibParams)]),
    "gs_calib_download_state": (C.c_int, [_h, C.c_uint32, C.c_uint32, C.POINTER(GsAgents)]),
    "gs_calib_timing": (C.c_int, [_h, _fp, _fp]),
}
for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_SIGS)


def last_error() -> str:
    return (lib.gs_last_error() or b"").decode()


def check(status: int, where: str) -> None:
    if status != GS_OK:
        raise GpuSteerError(status, where, last_error() or lib.gs_status_string(status).decode())


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def uptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_up)


def f3(v) -> "C.Array":
    return _f3(*[float(x) for x in v])


def u3(v) -> "C.Array":
    return _u3(*[int(x) for x in v])
