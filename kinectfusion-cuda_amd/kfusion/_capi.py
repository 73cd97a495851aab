"""ctypes declarations of the C ABI in include/kfusion.h (libkfusion.so).

The library is loaded from the package's own lib/ directory.  There is no
fallback: if it is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_ROOT, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libkfusion.so")
SYNTH_PATH = os.path.join(LIB_DIR, "libkfsynth.so")

KF_OK = 0
KF_ERR_DIMS = 1
KF_ERR_NONFINITE = 2
KF_ERR_NOT_PD = 3
KF_ERR_OOM = 4
KF_ERR_HIP = 5
KF_ERR_BUDGET = 6
KF_ERR_ARG = 7
KF_ERR_IO = 8

KF_VOXEL_F32 = 0
KF_VOXEL_SNORM16_U8 = 1

KF_STAGE_COUNT = 4
STAGE_NAMES = ("preprocess", "raycast", "icp", "fuse")


class kf_sensor(C.Structure):
    _fields_ = [
        ("pose", C.c_float * 16),
        ("inv_pose", C.c_float * 16),
        ("intr", (C.c_float * 9) * 3),
        ("inv_intr", (C.c_float * 9) * 3),
    ]


class kf_icp_config(C.Structure):
    _fields_ = [
        ("iters_per_layer", C.c_int * 3),
        ("levels", C.c_int),
        ("width", C.c_int),
        ("height", C.c_int),
        ("distance_thresh", C.c_float),
        ("angle_thresh", C.c_float),
        ("iteration_stop_thresh_angle", C.c_float),
        ("iteration_stop_thresh_distance", C.c_float),
        ("compensated", C.c_int),
    ]


class kf_pipeline_config(C.Structure):
    _fields_ = [
        ("frame_width", C.c_int),
        ("frame_height", C.c_int),
        ("depth_scale", C.c_float),
        ("fx", C.c_float),
        ("fy", C.c_float),
        ("cx", C.c_float),
        ("cy", C.c_float),
        ("unmirror", C.c_int),
        ("volume_n", C.c_int),
        ("volume_size", C.c_float),
        ("mue", C.c_float),
        ("voxel_format", C.c_int),
        ("initial_pose", C.c_float * 16),
        ("icp", kf_icp_config),
        ("vis_volume_n", C.c_int),
        ("vis_volume_size", C.c_float),
        ("vis_mue", C.c_float),
    ]


class kf_session_config(C.Structure):
    _fields_ = [
        ("frame_width", C.c_int),
        ("frame_height", C.c_int),
        ("depth_scale", C.c_float),
        ("fx", C.c_float),
        ("fy", C.c_float),
        ("cx", C.c_float),
        ("cy", C.c_float),
        ("volume_n", C.c_int),
        ("volume_size", C.c_float),
        ("mue", C.c_float),
        ("voxel_format", C.c_int),
        ("initial_pose", C.c_float * 16),
    ]


class kf_batch_shape(C.Structure):
    _fields_ = [
        ("preprocess_blocks", C.c_uint),
        ("raycast_blocks", C.c_uint),
        ("icp_blocks", C.c_uint),
        ("cull_blocks", C.c_uint),
        ("fuse_blocks", C.c_uint),
    ]


class kf_batch_config(C.Structure):
    _fields_ = [
        ("max_sessions", C.c_int),
        ("icp", kf_icp_config),
        ("use_graph", C.c_int),
        ("autotune", C.c_int),
        ("shape", kf_batch_shape),
    ]


class kf_server_config(C.Structure):
    _fields_ = [
        ("batch", kf_batch_config),
        ("hbm_budget_bytes", C.c_size_t),
        ("queue_depth", C.c_int),
        ("launch_fill", C.c_float),
        ("max_wait_ms", C.c_float),
    ]


class kf_server_stats(C.Structure):
    _fields_ = [
        ("frames", C.c_int64),
        ("ticks", C.c_int64),
        ("admitted", C.c_int64),
        ("queued_admissions", C.c_int64),
        ("failed_frames", C.c_int64),
        ("wall_s", C.c_double),
        ("frames_per_s", C.c_double),
        ("mean_frames_per_tick", C.c_double),
        ("offered_frames_per_s", C.c_double),
        ("latency_p50_ms", C.c_double),
        ("latency_p99_ms", C.c_double),
        ("latency_max_ms", C.c_double),
        ("latency_mean_ms", C.c_double),
        ("peak_resident_bytes", C.c_size_t),
    ]


class kf_poisson_load(C.Structure):
    _fields_ = [
        ("sessions", C.c_int),
        ("configs", C.c_void_p),
        ("banks", C.c_void_p),
        ("bank_len", C.c_void_p),
        ("frames_per_session", C.c_void_p),
        ("opens_s", C.c_void_p),
        ("rate_hz", C.c_double),
        ("jitter", C.c_double),
        ("seed", C.c_uint64),
        ("keep_open", C.c_int),
    ]


P = C.c_void_p
I = C.c_int
F = C.c_float
FP = C.POINTER(C.c_float)
SZ = C.c_size_t
I64 = C.c_int64

# name -> (restype, argtypes); every entry point returns an int status
SIGNATURES = {
    "kf_status_string": (C.c_char_p, [I]),
    "kf_last_error": (C.c_char_p, []),
    "kf_device_init": (I, [I]),
    "kf_device_synchronize": (I, []),
    "kf_malloc": (I, [C.POINTER(P), SZ]),
    "kf_free": (I, [P]),
    "kf_memcpy_h2d": (I, [P, P, SZ, P]),
    "kf_memcpy_d2h": (I, [P, P, SZ, P]),
    "kf_memcpy_d2d": (I, [P, P, SZ, P]),
    "kf_stream_create": (I, [C.POINTER(P)]),
    "kf_stream_destroy": (I, [P]),
    "kf_stream_synchronize": (I, [P]),
    "kf_event_create": (I, [C.POINTER(P)]),
    "kf_event_destroy": (I, [P]),
    "kf_event_record": (I, [P, P]),
    "kf_event_elapsed_ms": (I, [P, P, FP]),
    "kf_sensor_init": (I, [C.POINTER(kf_sensor), F, F, F, F]),
    "kf_sensor_set_pose": (I, [C.POINTER(kf_sensor), P]),
    "kf_convert_to_depth_meters": (I, [P, P, I, I, F, I, P, FP]),
    "kf_apply_bilateral_filter": (I, [P, P, I, I, P, FP]),
    "kf_down_sample": (I, [P, I, I, P, I, I, P, FP]),
    "kf_create_vertex_map": (I, [P, P, I, I, P, P, FP]),
    "kf_create_normal_map": (I, [P, P, I, I, P, FP]),
    "kf_preprocess": (I, [P, I, I, F, I, I, P, P, P, P, P, P, FP]),
    "kf_volume_create": (I, [I, F, F, I, C.POINTER(P)]),
    "kf_volume_create_scaled": (I, [I, F, F, I, C.POINTER(P)]),
    "kf_volume_destroy": (I, [P]),
    "kf_volume_info": (I, [P, C.POINTER(I), FP, FP, FP, C.POINTER(I)]),
    "kf_volume_bytes": (I, [P, C.POINTER(SZ)]),
    "kf_volume_touched_total": (I, [P, C.POINTER(I64)]),
    "kf_initialize_grid": (I, [P, F, F, P, FP]),
    "kf_fuse": (I, [P, I, I, P, C.POINTER(kf_sensor), P, FP]),
    "kf_raycast": (I, [P, C.POINTER(kf_sensor), P, P, I, I, I, P, FP]),
    "kf_volume_download": (I, [P, P]),
    "kf_volume_upload": (I, [P, P]),
    "kf_volume_save": (I, [P, C.c_char_p]),
    "kf_volume_load": (I, [P, C.c_char_p]),
    "kf_construct_icp_residuals": (I, [P, P, P, I, I, P, P, P, P, P, F, F, P, P, P, FP]),
    "kf_solve_normal_equations": (I, [P, P, C.c_uint, P, P, FP]),
    "kf_solve_normal_equations_noalloc": (I, [P, P, C.c_uint, P, P, P]),
    "kf_icp_reduce": (I, [P, P, P, I, I, P, P, P, P, P, F, F, I, P, C.POINTER(I64), P, FP]),
    "kf_icp_create": (I, [C.POINTER(kf_icp_config), C.POINTER(P)]),
    "kf_icp_destroy": (I, [P]),
    "kf_icp_compute_pose": (I, [P, P, P, P, C.POINTER(kf_sensor), P, P, FP]),
    "kf_icp_execution_times": (I, [P, P]),
    "kf_icp_iterations": (I, [P, C.POINTER(I)]),
    "kf_pipeline_create": (I, [C.POINTER(kf_pipeline_config), P, C.POINTER(P)]),
    "kf_pipeline_destroy": (I, [P]),
    "kf_pipeline_frame": (I, [P, P]),
    "kf_pipeline_sync": (I, [P, C.POINTER(I)]),
    "kf_pipeline_pose": (I, [P, P]),
    "kf_pipeline_volume": (I, [P, C.POINTER(P)]),
    "kf_pipeline_map": (I, [P, I, I, C.POINTER(P), C.POINTER(I), C.POINTER(I)]),
    "kf_pipeline_set_timing": (I, [P, I]),
    "kf_pipeline_stage_times": (I, [P, P, C.POINTER(I64)]),
    "kf_pipeline_fuse_kernel_ms": (I, [P, FP, C.POINTER(I64)]),
    "kf_pipeline_render": (I, [P, P, I, P]),
    "kf_render_shading": (I, [P, I, I, C.POINTER(kf_sensor), P, P, FP]),
    "kf_render_normals": (I, [P, I, I, P, P, P, FP]),
    "kf_tum_open": (I, [C.c_char_p, C.POINTER(P)]),
    "kf_tum_close": (I, [P]),
    "kf_tum_size": (I, [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
    "kf_tum_frame": (I, [P, I, I, P, P]),
    "kf_pose_error": (I, [P, P, FP, FP]),
    "kf_session_bytes": (I, [C.POINTER(kf_session_config), C.POINTER(SZ)]),
    "kf_batch_create": (I, [C.POINTER(kf_batch_config), P, C.POINTER(P)]),
    "kf_batch_destroy": (I, [P]),
    "kf_batch_add_session": (I, [P, C.POINTER(kf_session_config), C.POINTER(I)]),
    "kf_batch_remove_session": (I, [P, I]),
    "kf_batch_tick": (I, [P, I, P, P]),
    "kf_batch_sync": (I, [P]),
    "kf_batch_pose": (I, [P, I, P]),
    "kf_batch_session_status": (I, [P, I, C.POINTER(I)]),
    "kf_batch_volume": (I, [P, I, C.POINTER(P)]),
    "kf_batch_map": (I, [P, I, I, I, C.POINTER(P), C.POINTER(I), C.POINTER(I)]),
    "kf_batch_shape_get": (I, [P, C.POINTER(kf_batch_shape), C.POINTER(I)]),
    "kf_batch_tick_times": (I, [P, C.POINTER(C.c_double), C.POINTER(I64), FP]),
    "kf_batch_touched_total": (I, [P, C.POINTER(C.c_ulonglong)]),
    "kf_batch_set_stage_timing": (I, [P, I]),
    "kf_batch_stage_times": (I, [P, P, C.POINTER(I64)]),
    "kf_batch_session_save": (I, [P, I, C.c_char_p]),
    "kf_batch_session_restore": (I, [P, I, C.c_char_p]),
    "kf_server_checkpoint": (I, [P, I, C.c_char_p]),
    "kf_server_restore": (I, [P, I, C.c_char_p]),
    "kf_server_create": (I, [C.POINTER(kf_server_config), C.POINTER(P)]),
    "kf_server_destroy": (I, [P]),
    "kf_server_now": (C.c_double, []),
    "kf_server_open": (I, [P, C.POINTER(kf_session_config), C.POINTER(I)]),
    "kf_server_close": (I, [P, I]),
    "kf_server_submit": (I, [P, I, P, C.c_double]),
    "kf_server_poll": (I, [P, C.POINTER(I)]),
    "kf_server_drain": (I, [P]),
    "kf_server_stats_get": (I, [P, C.POINTER(kf_server_stats)]),
    "kf_server_latencies": (I, [P, C.POINTER(C.c_double), SZ, C.POINTER(SZ)]),
    "kf_server_state": (I, [P, I, C.POINTER(I)]),
    "kf_server_pose": (I, [P, I, P]),
    "kf_server_volume": (I, [P, I, C.POINTER(P)]),
    "kf_server_run_load": (I, [P, C.POINTER(kf_poisson_load), C.POINTER(I), C.POINTER(kf_server_stats)]),
}

_lib = None


def load():
    """Loads libkfusion.so and declares every entry point; raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libkfusion.so is not built ({LIB_PATH}); run `make -C {PKG_ROOT}` — there is no CPU fallback")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return list(SIGNATURES)
