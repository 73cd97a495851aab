"""Python surface of the MI355X-native KinectFusion per-frame path.

Thin wrappers over the C ABI (include/kfusion.h, libkfusion.so) that mirror the
reference's class surface (isikmustafa/kinectfusion-cuda, project/src):
``GridMap`` (CudaGridMap), ``GridMapPyramid``, ``DepthMap``, ``VoxelGrid``,
``Sensor``, ``RigidTransform3D``, ``ICP``, ``LinearLeastSquares`` and the
``kernel`` functions, each returning elapsed milliseconds like the reference's
wrappers.  Every call goes to HIP kernels; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _capi
from ._capi import (KF_ERR_ARG, KF_ERR_BUDGET, KF_ERR_DIMS, KF_ERR_IO, KF_ERR_NONFINITE, KF_ERR_NOT_PD, KF_OK,
                    KF_VOXEL_F32, KF_VOXEL_SNORM16_U8, STAGE_NAMES, kf_batch_config, kf_batch_shape,
                    kf_icp_config, kf_pipeline_config, kf_poisson_load, kf_sensor, kf_server_config,
                    kf_server_stats, kf_session_config)

__all__ = [
    "KfError", "lib", "init", "DeviceArray", "GridMap", "DepthMap", "GridMapPyramid", "Sensor",
    "RigidTransform3D", "VoxelGrid", "ICP", "IcpConfig", "LinearLeastSquares", "kernel", "Pipeline",
    "Batch", "SessionConfig", "session_bytes", "Server", "TumSequence", "pose_error", "KF_VOXEL_F32", "KF_VOXEL_SNORM16_U8", "STAGE_NAMES",
]


class KfError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


def lib():
    return _capi.load()


def check(status, what=""):
    if status != KF_OK:
        l = lib()
        detail = l.kf_last_error().decode()
        name = l.kf_status_string(status).decode()
        raise KfError(status, f"{what}: {name}" + (f" ({detail})" if detail else ""))


_inited = False


def init(device=0):
    global _inited
    if not _inited:
        check(lib().kf_device_init(device), "kf_device_init")
        _inited = True


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _ms():
    return C.c_float(0.0)


# ----------------------------------------------------------------- memory
class DeviceArray:
    """A device allocation with a numpy shape/dtype (kf_malloc / kf_free)."""

    def __init__(self, shape, dtype=np.float32):
        init()
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        ptr = C.c_void_p()
        check(lib().kf_malloc(C.byref(ptr), self.nbytes), "kf_malloc")
        self.ptr = ptr

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(lib().kf_memcpy_h2d(self.ptr, _p(a), self.nbytes, None), "kf_memcpy_h2d")

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().kf_memcpy_d2h(_p(out), self.ptr, self.nbytes, None), "kf_memcpy_d2h")
        return out

    def free(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            lib().kf_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():  # the HIP runtime may already be torn down; the process frees everything
            return
        try:
            self.free()
        except Exception:
            pass


class GridMap(DeviceArray):
    """CudaGridMap (cuda_grid_map.h:13-29): a w x h map of 1 (depth), 4 (vertex/normal)
    float channels or raw uint16 codes."""

    def __init__(self, width, height, channels=1, dtype=np.float32):
        shape = (height, width) if channels == 1 else (height, width, channels)
        super().__init__(shape, dtype)
        self.width, self.height, self.channels = width, height, channels

    def get_grid_dims(self):
        return (self.width, self.height)


class DepthMap(GridMap):
    """DepthMap (depth_map.h:8-15): raw uint16 depth codes with host upload."""

    def __init__(self, width, height):
        super().__init__(width, height, 1, np.uint16)

    def update(self, data):  # depth_map.cpp:30-35
        self.upload(np.asarray(data, np.uint16).reshape(self.height, self.width))


class GridMapPyramid:
    """GridMapPyramid<T> (grid_map_pyramid.h:10-73)."""

    def __init__(self, width, height, n_layers, channels=1, dtype=np.float32):
        e = 1 << (n_layers - 1)
        if width % e or height % e:  # grid_map_pyramid.h:63-72
            raise KfError(KF_ERR_DIMS, "Can't construct pyramid. Dimensions must be divisible by 2^(m_n_layers - 1).")
        self.layers = [GridMap(width >> i, height >> i, channels, dtype) for i in range(n_layers)]

    def __getitem__(self, i):
        return self.layers[i]

    def __len__(self):
        return len(self.layers)


# ----------------------------------------------------------------- sensor / pose
class Sensor:
    """Sensor (sensor.h:7-53) with generalised intrinsics (fx, fy, cx, cy)."""

    def __init__(self, fx=525.0, fy=None, cx=319.5, cy=239.5):
        self.s = kf_sensor()
        check(lib().kf_sensor_init(C.byref(self.s), fx, fx if fy is None else fy, cx, cy), "kf_sensor_init")

    def set_pose(self, pose):
        p = np.ascontiguousarray(np.asarray(pose, np.float32).reshape(-1))
        check(lib().kf_sensor_set_pose(C.byref(self.s), _p(p)), "kf_sensor_set_pose")

    def get_pose(self):
        return np.array(self.s.pose[:], np.float32)

    def get_inverse_pose(self):
        return np.array(self.s.inv_pose[:], np.float32)

    def get_intr(self, level):
        return np.array(self.s.intr[level][:], np.float32)

    def get_inverse_intr(self, level):
        return np.array(self.s.inv_intr[level][:], np.float32)


class RigidTransform3D:
    """RigidTransform3D (rigid_transform_3d.h:5-23): glm column-major R (9) and t (3)."""

    def __init__(self, transformation=None):
        m = np.eye(4, dtype=np.float32).T.reshape(-1) if transformation is None else \
            np.asarray(transformation, np.float32).reshape(-1)
        self.rot_mat = np.array([m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]], np.float32)
        self.transl_vec = np.array(m[12:15], np.float32)

    def get_transformation(self):
        P = np.zeros(16, np.float32)
        R = self.rot_mat
        P[0:3], P[4:7], P[8:11] = R[0:3], R[3:6], R[6:9]
        P[12:15] = self.transl_vec
        P[15] = 1.0
        return P


# ----------------------------------------------------------------- kernels
class kernel:
    """The reference's ``kernel::`` wrappers; each returns elapsed ms."""

    @staticmethod
    def convert_to_depth_meters(raw: GridMap, out: GridMap, scale, unmirror=False):
        if raw.get_grid_dims() != out.get_grid_dims():  # measurement.cu:144-147
            raise KfError(KF_ERR_DIMS, "convertToDepthMeters: input and output surface objects are not of same size!")
        ms = _ms()
        check(lib().kf_convert_to_depth_meters(raw.ptr, out.ptr, raw.width, raw.height, scale, int(unmirror), None,
                                               C.byref(ms)), "convertToDepthMeters")
        return ms.value

    @staticmethod
    def apply_bilateral_filter(inp: GridMap, out: GridMap):
        if inp.get_grid_dims() != out.get_grid_dims():
            raise KfError(KF_ERR_DIMS, "applyBilateralFilter: input and output surface objects are not of same size!")
        ms = _ms()
        check(lib().kf_apply_bilateral_filter(inp.ptr, out.ptr, inp.width, inp.height, None, C.byref(ms)),
              "applyBilateralFilter")
        return ms.value

    @staticmethod
    def down_sample(inp: GridMap, out: GridMap):
        ms = _ms()
        check(lib().kf_down_sample(inp.ptr, inp.width, inp.height, out.ptr, out.width, out.height, None,
                                   C.byref(ms)), "downSample")
        return ms.value

    @staticmethod
    def create_vertex_map(depth: GridMap, vmap: GridMap, inv_k):
        if depth.get_grid_dims() != vmap.get_grid_dims():
            raise KfError(KF_ERR_DIMS, "createVertexMap: input and output surface objects are not of same size!")
        ki = np.ascontiguousarray(np.asarray(inv_k, np.float32).reshape(9))
        ms = _ms()
        check(lib().kf_create_vertex_map(depth.ptr, vmap.ptr, depth.width, depth.height, _p(ki), None, C.byref(ms)),
              "createVertexMap")
        return ms.value

    @staticmethod
    def create_normal_map(vmap: GridMap, nmap: GridMap):
        if vmap.get_grid_dims() != nmap.get_grid_dims():
            raise KfError(KF_ERR_DIMS, "computeNormalMap: input and output surface objects are not of same size!")
        ms = _ms()
        check(lib().kf_create_normal_map(vmap.ptr, nmap.ptr, vmap.width, vmap.height, None, C.byref(ms)),
              "createNormalMap")
        return ms.value

    @staticmethod
    def initialize_grid(grid: "VoxelGrid", f=2.0, w=0.0):
        ms = _ms()
        check(lib().kf_initialize_grid(grid.h, f, w, None, C.byref(ms)), "initializeGrid")
        return ms.value

    @staticmethod
    def fuse(depth_m: GridMap, grid: "VoxelGrid", sensor: Sensor):
        ms = _ms()
        check(lib().kf_fuse(depth_m.ptr, depth_m.width, depth_m.height, grid.h, C.byref(sensor.s), None,
                            C.byref(ms)), "fuse")
        return ms.value

    @staticmethod
    def raycast(grid: "VoxelGrid", sensor: Sensor, vmap: GridMap, nmap: GridMap, level):
        if vmap.get_grid_dims() != nmap.get_grid_dims():  # raycast.cu:155-158
            raise KfError(KF_ERR_DIMS, "raycast: vertex and normal surface objects are not of same size!")
        ms = _ms()
        check(lib().kf_raycast(grid.h, C.byref(sensor.s), vmap.ptr, nmap.ptr, vmap.width, vmap.height, level, None,
                               C.byref(ms)), "raycast")
        return ms.value

    @staticmethod
    def construct_icp_residuals(vmap, tvmap, tnmap, prev_rot, prev_transl, cur_rot, cur_transl, intrinsics,
                                distance_thresh, angle_thresh, mat_a: DeviceArray, vec_b: DeviceArray):
        f = lambda a, n: np.ascontiguousarray(np.asarray(a, np.float32).reshape(n))
        args = [f(prev_rot, 9), f(prev_transl, 3), f(cur_rot, 9), f(cur_transl, 3), f(intrinsics, 9)]
        ms = _ms()
        check(lib().kf_construct_icp_residuals(vmap.ptr, tvmap.ptr, tnmap.ptr, vmap.width, vmap.height,
                                               *[_p(a) for a in args], distance_thresh, angle_thresh,
                                               mat_a.ptr, vec_b.ptr, None, C.byref(ms)), "constructIcpResiduals")
        return ms.value

    @staticmethod
    def icp_reduce(vmap, tvmap, tnmap, prev_rot, prev_transl, cur_rot, cur_transl, intrinsics, distance_thresh,
                   angle_thresh, compensated=False):
        f = lambda a, n: np.ascontiguousarray(np.asarray(a, np.float32).reshape(n))
        args = [f(prev_rot, 9), f(prev_transl, 3), f(cur_rot, 9), f(cur_transl, 3), f(intrinsics, 9)]
        sums = np.zeros(27, np.float64)
        inl = C.c_int64(0)
        check(lib().kf_icp_reduce(vmap.ptr, tvmap.ptr, tnmap.ptr, vmap.width, vmap.height, *[_p(a) for a in args],
                                  distance_thresh, angle_thresh, int(compensated), _p(sums), C.byref(inl), None,
                                  None), "kf_icp_reduce")
        return sums, inl.value


class LinearLeastSquares:
    """LinearLeastSquares (linear_least_squares.h:7-35)."""

    def solve(self, mat_a: DeviceArray, vec_b: DeviceArray, n_equations):
        x = np.zeros(6, np.float32)
        check(lib().kf_solve_normal_equations(mat_a.ptr, vec_b.ptr, n_equations, _p(x), None, None),
              "LinearLeastSquares::solve")
        return x


# ----------------------------------------------------------------- volume
class VoxelGrid:
    """VoxelGrid (voxel_grid.h:8-61): VoxelGrid(width_m, n, mue_scale) as the reference,
    or with an explicit truncation distance ``mue=`` in metres."""

    def __init__(self, total_width_in_meters, n, mue_scale=None, mue=None, voxel_format=KF_VOXEL_F32):
        init()
        h = C.c_void_p()
        if mue is not None:
            check(lib().kf_volume_create(n, total_width_in_meters, mue, voxel_format, C.byref(h)), "kf_volume_create")
        else:
            check(lib().kf_volume_create_scaled(n, total_width_in_meters, 3.0 if mue_scale is None else mue_scale,
                                                voxel_format, C.byref(h)), "kf_volume_create")
        self.h = h
        n_, w_, r_, m_, f_ = C.c_int(), C.c_float(), C.c_float(), C.c_float(), C.c_int()
        check(lib().kf_volume_info(h, C.byref(n_), C.byref(w_), C.byref(r_), C.byref(m_), C.byref(f_)), "info")
        self.n, self.total_width_in_meters, self.resolution, self.mue, self.format = \
            n_.value, w_.value, r_.value, m_.value, f_.value

    def download(self):
        """(F, W) as linear x-fastest arrays of n^3 floats (voxel_grid.cpp:43-60 order)."""
        fw = np.empty(self.n ** 3 * 2, np.float32)
        check(lib().kf_volume_download(self.h, _p(fw)), "kf_volume_download")
        return fw[0::2].copy(), fw[1::2].copy()

    def upload(self, F, W):
        fw = np.empty(self.n ** 3 * 2, np.float32)
        fw[0::2], fw[1::2] = F, W
        check(lib().kf_volume_upload(self.h, _p(fw)), "kf_volume_upload")

    def save_voxel_grid(self, path):
        check(lib().kf_volume_save(self.h, str(path).encode()), "saveVoxelGrid")

    def load_voxel_grid(self, path):
        check(lib().kf_volume_load(self.h, str(path).encode()), "loadVoxelGrid")

    def touched_total(self):
        c = C.c_int64()
        check(lib().kf_volume_touched_total(self.h, C.byref(c)), "touched")
        return c.value

    def nbytes(self):
        b = C.c_size_t()
        check(lib().kf_volume_bytes(self.h, C.byref(b)), "bytes")
        return b.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().kf_volume_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():  # the HIP runtime may already be torn down; the process frees everything
            return
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- ICP
def IcpConfig(iters_per_layer=(10, 5, 4), width=640, height=480, distance_thresh=0.1,
              angle_thresh=np.deg2rad(20.0), iteration_stop_thresh_angle=0.0,
              iteration_stop_thresh_distance=0.0, compensated=False):
    """IcpConfig (icp.h:11-20)."""
    c = kf_icp_config()
    its = list(iters_per_layer)
    for i, v in enumerate(its):
        c.iters_per_layer[i] = int(v)
    c.levels = len(its)
    c.width, c.height = width, height
    c.distance_thresh = distance_thresh
    c.angle_thresh = angle_thresh
    c.iteration_stop_thresh_angle = iteration_stop_thresh_angle
    c.iteration_stop_thresh_distance = iteration_stop_thresh_distance
    c.compensated = int(compensated)
    return c


class ICP:
    """ICP (icp.h:29-53), device-resident coarse-to-fine loop."""

    def __init__(self, config):
        init()
        self.cfg = config
        h = C.c_void_p()
        check(lib().kf_icp_create(C.byref(config), C.byref(h)), "kf_icp_create")
        self.h = h

    def compute_pose(self, vertex_pyramid, target_vertex_pyramid, target_normal_pyramid, sensor: Sensor):
        L = self.cfg.levels
        arr = C.c_void_p * L
        vp = arr(*[vertex_pyramid[i].ptr.value for i in range(L)])
        tv = arr(*[target_vertex_pyramid[i].ptr.value for i in range(L)])
        tn = arr(*[target_normal_pyramid[i].ptr.value for i in range(L)])
        out = np.zeros(16, np.float32)
        st = lib().kf_icp_compute_pose(self.h, vp, tv, tn, C.byref(sensor.s), _p(out), None, None)
        if st in (KF_ERR_NONFINITE, KF_ERR_NOT_PD):  # icp.cpp:67-73 throws
            raise KfError(st, "Error: an element in the result vector is not finite or NAN.")
        check(st, "ICP::computePose")
        return RigidTransform3D(out)

    def get_execution_times(self):
        t = np.zeros(2, np.float32)
        check(lib().kf_icp_execution_times(self.h, _p(t)), "getExecutionTimes")
        return t

    def iterations(self):
        i = C.c_int()
        check(lib().kf_icp_iterations(self.h, C.byref(i)), "iterations")
        return i.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().kf_icp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():  # the HIP runtime may already be torn down; the process frees everything
            return
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- pipeline
class Pipeline:
    """The per-frame loop (kinect_fusion.cpp:44-183) on one volume."""

    def __init__(self, width=640, height=480, fx=525.0, fy=525.0, cx=319.5, cy=239.5, depth_scale=1e-3,
                 volume_n=256, volume_size=3.0, mue=0.03, voxel_format=KF_VOXEL_F32, initial_pose=None,
                 icp=None, unmirror=False, vis_volume=None):
        init()
        c = kf_pipeline_config()
        c.frame_width, c.frame_height = width, height
        c.depth_scale = depth_scale
        c.fx, c.fy, c.cx, c.cy = fx, fy, cx, cy
        c.unmirror = int(unmirror)
        c.volume_n, c.volume_size, c.mue, c.voxel_format = volume_n, volume_size, mue, voxel_format
        pose = np.eye(4, dtype=np.float32).T.reshape(-1) if initial_pose is None else \
            np.asarray(initial_pose, np.float32).reshape(-1)
        for i in range(16):
            c.initial_pose[i] = float(pose[i])
        c.icp = icp if icp is not None else IcpConfig(width=width, height=height)
        if vis_volume:  # (n, size, mue) of the visualisation volume
            c.vis_volume_n, c.vis_volume_size, c.vis_mue = vis_volume
        self.cfg = c
        h = C.c_void_p()
        check(lib().kf_pipeline_create(C.byref(c), None, C.byref(h)), "kf_pipeline_create")
        self.h = h
        self.width, self.height = width, height

    def frame(self, raw_device_ptr):
        check(lib().kf_pipeline_frame(self.h, raw_device_ptr), "kf_pipeline_frame")

    def sync(self):
        st = C.c_int(0)
        check(lib().kf_pipeline_sync(self.h, C.byref(st)), "kf_pipeline_sync")
        return st.value

    def pose(self):
        p = np.zeros(16, np.float32)
        check(lib().kf_pipeline_pose(self.h, _p(p)), "kf_pipeline_pose")
        return p

    def map(self, which, level=0):
        ptr, w, h = C.c_void_p(), C.c_int(), C.c_int()
        check(lib().kf_pipeline_map(self.h, which, level, C.byref(ptr), C.byref(w), C.byref(h)), "kf_pipeline_map")
        ch = 1 if which in (0, 1) else 4
        shape = (h.value, w.value) if ch == 1 else (h.value, w.value, ch)
        out = np.empty(shape, np.float32)
        check(lib().kf_memcpy_d2h(_p(out), ptr, out.nbytes, None), "map download")
        return out

    def render(self, shading=True, pose=None):
        """RGBA8 image (h, w) of the model from pose (default: the current pose)."""
        img = DeviceArray((self.height, self.width), np.uint32)
        pp = None if pose is None else _p(np.ascontiguousarray(pose, np.float32))
        check(lib().kf_pipeline_render(self.h, pp, int(shading), img.ptr), "kf_pipeline_render")
        out = img.download()
        img.free()
        return out

    def volume_download(self):
        v = C.c_void_p()
        check(lib().kf_pipeline_volume(self.h, C.byref(v)), "kf_pipeline_volume")
        n = self.cfg.volume_n
        fw = np.empty(n ** 3 * 2, np.float32)
        check(lib().kf_volume_download(v, _p(fw)), "kf_volume_download")
        return fw[0::2].copy(), fw[1::2].copy()

    def touched_total(self):
        v = C.c_void_p()
        check(lib().kf_pipeline_volume(self.h, C.byref(v)), "kf_pipeline_volume")
        c = C.c_int64()
        check(lib().kf_volume_touched_total(v, C.byref(c)), "touched")
        return c.value

    def set_timing(self, on=True):
        check(lib().kf_pipeline_set_timing(self.h, int(on)), "set_timing")

    def stage_times(self):
        t = np.zeros(4, np.float32)
        n = C.c_int64()
        check(lib().kf_pipeline_stage_times(self.h, _p(t), C.byref(n)), "stage_times")
        return dict(zip(STAGE_NAMES, t.tolist())), n.value

    def fuse_kernel_ms(self):
        a = C.c_float()
        n = C.c_int64()
        check(lib().kf_pipeline_fuse_kernel_ms(self.h, C.byref(a), C.byref(n)), "fuse_kernel_ms")
        return a.value, n.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().kf_pipeline_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():  # the HIP runtime may already be torn down; the process frees everything
            return
        try:
            self.close()
        except Exception:
            pass


def SessionConfig(width=640, height=480, fx=525.0, fy=525.0, cx=None, cy=None, depth_scale=1e-3, volume_n=512,
                  volume_size=4.0, mue=0.03, voxel_format=KF_VOXEL_SNORM16_U8, initial_pose=None):
    """kf_session_config; cx, cy default to (w - 1) / 2, (h - 1) / 2 (sensor.h:14 for 640x480)."""
    c = kf_session_config()
    c.frame_width, c.frame_height = width, height
    c.depth_scale = depth_scale
    c.fx, c.fy = fx, fy
    c.cx = (width - 1) / 2.0 if cx is None else cx
    c.cy = (height - 1) / 2.0 if cy is None else cy
    c.volume_n, c.volume_size, c.mue, c.voxel_format = volume_n, volume_size, mue, voxel_format
    pose = np.eye(4, dtype=np.float32).T.reshape(-1) if initial_pose is None else \
        np.asarray(initial_pose, np.float32).reshape(-1)
    for i in range(16):
        c.initial_pose[i] = float(pose[i])
    return c


def session_bytes(cfg):
    n = C.c_size_t()
    check(lib().kf_session_bytes(C.byref(cfg), C.byref(n)), "kf_session_bytes")
    return n.value


class Batch:
    """Many sessions advanced together, one launch per stage per tick (kf_batch_*)."""

    def __init__(self, max_sessions, icp=None, use_graph=True, autotune=False, shape=None, stream=None):
        init()
        c = kf_batch_config()
        c.max_sessions = max_sessions
        c.icp = icp if icp is not None else IcpConfig()
        c.use_graph = int(use_graph)
        c.autotune = int(autotune)
        if shape:
            for k, v in shape.items():
                setattr(c.shape, k, int(v))
        self.cfg = c
        h = C.c_void_p()
        check(lib().kf_batch_create(C.byref(c), stream, C.byref(h)), "kf_batch_create")
        self.h = h
        self.sessions = {}

    def add(self, cfg):
        sid = C.c_int()
        check(lib().kf_batch_add_session(self.h, C.byref(cfg), C.byref(sid)), "kf_batch_add_session")
        self.sessions[sid.value] = cfg
        return sid.value

    def remove(self, sid):
        check(lib().kf_batch_remove_session(self.h, sid), "kf_batch_remove_session")
        self.sessions.pop(sid, None)

    def tick(self, ids, raw_ptrs):
        n = len(ids)
        ia = (C.c_int * max(n, 1))(*ids)
        pa = (C.c_void_p * max(n, 1))(*[int(p) for p in raw_ptrs])
        check(lib().kf_batch_tick(self.h, n, ia, pa), "kf_batch_tick")

    def sync(self):
        check(lib().kf_batch_sync(self.h), "kf_batch_sync")

    def pose(self, sid):
        p = np.zeros(16, np.float32)
        check(lib().kf_batch_pose(self.h, sid, _p(p)), "kf_batch_pose")
        return p

    def status(self, sid):
        st = C.c_int()
        check(lib().kf_batch_session_status(self.h, sid, C.byref(st)), "kf_batch_session_status")
        return st.value

    def map(self, sid, which, level=0):
        ptr, w, h = C.c_void_p(), C.c_int(), C.c_int()
        check(lib().kf_batch_map(self.h, sid, which, level, C.byref(ptr), C.byref(w), C.byref(h)), "kf_batch_map")
        ch = 1 if which in (0, 1) else 4
        shape = (h.value, w.value) if ch == 1 else (h.value, w.value, ch)
        out = np.empty(shape, np.float32)
        self.sync()
        check(lib().kf_memcpy_d2h(_p(out), ptr, out.nbytes, None), "map download")
        return out

    def volume_download(self, sid):
        v = C.c_void_p()
        self.sync()
        check(lib().kf_batch_volume(self.h, sid, C.byref(v)), "kf_batch_volume")
        n = self.sessions[sid].volume_n
        fw = np.empty(n ** 3 * 2, np.float32)
        check(lib().kf_volume_download(v, _p(fw)), "kf_volume_download")
        return fw[0::2].copy(), fw[1::2].copy()

    def save_session(self, sid, path):
        check(lib().kf_batch_session_save(self.h, sid, str(path).encode()), "kf_batch_session_save")

    def restore_session(self, sid, path):
        check(lib().kf_batch_session_restore(self.h, sid, str(path).encode()), "kf_batch_session_restore")

    def shape(self):
        sh = kf_batch_shape()
        tuned = C.c_int()
        check(lib().kf_batch_shape_get(self.h, C.byref(sh), C.byref(tuned)), "kf_batch_shape_get")
        return {k: getattr(sh, k) for k, _ in sh._fields_}, bool(tuned.value)

    def tick_times(self):
        tot, n, mx = C.c_double(), C.c_int64(), C.c_float()
        check(lib().kf_batch_tick_times(self.h, C.byref(tot), C.byref(n), C.byref(mx)), "kf_batch_tick_times")
        return tot.value, n.value, mx.value

    def touched_total(self):
        v = C.c_ulonglong()
        check(lib().kf_batch_touched_total(self.h, C.byref(v)), "kf_batch_touched_total")
        return v.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().kf_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():  # the HIP runtime may already be torn down; the process frees everything
            return
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop (kf_server_*): HBM-budget admission, per-session frame queues,
    continuous batching over a Batch, per-frame latency."""

    def __init__(self, max_sessions, hbm_budget_bytes, icp=None, use_graph=True, autotune=False, shape=None,
                 queue_depth=8, launch_fill=0.5, max_wait_ms=2.0):
        init()
        c = kf_server_config()
        c.batch.max_sessions = max_sessions
        c.batch.icp = icp if icp is not None else IcpConfig()
        c.batch.use_graph = int(use_graph)
        c.batch.autotune = int(autotune)
        if shape:
            for k, v in shape.items():
                setattr(c.batch.shape, k, int(v))
        c.hbm_budget_bytes = int(hbm_budget_bytes)
        c.queue_depth = queue_depth
        c.launch_fill = launch_fill
        c.max_wait_ms = max_wait_ms
        self.cfg = c
        h = C.c_void_p()
        check(lib().kf_server_create(C.byref(c), C.byref(h)), "kf_server_create")
        self.h = h
        self.configs = {}

    @staticmethod
    def now():
        return lib().kf_server_now()

    def open(self, cfg):
        t = C.c_int()
        check(lib().kf_server_open(self.h, C.byref(cfg), C.byref(t)), "kf_server_open")
        self.configs[t.value] = cfg
        return t.value

    def close_session(self, ticket):
        check(lib().kf_server_close(self.h, ticket), "kf_server_close")

    def submit(self, ticket, frame, arrival_s):
        f = np.ascontiguousarray(frame, np.uint16)
        return lib().kf_server_submit(self.h, ticket, _p(f), float(arrival_s))

    def poll(self):
        l = C.c_int()
        check(lib().kf_server_poll(self.h, C.byref(l)), "kf_server_poll")
        return bool(l.value)

    def drain(self):
        check(lib().kf_server_drain(self.h), "kf_server_drain")

    def state(self, ticket):
        s = C.c_int()
        check(lib().kf_server_state(self.h, ticket, C.byref(s)), "kf_server_state")
        return s.value

    def pose(self, ticket):
        p = np.zeros(16, np.float32)
        check(lib().kf_server_pose(self.h, ticket, _p(p)), "kf_server_pose")
        return p

    def volume_download(self, ticket):
        v = C.c_void_p()
        check(lib().kf_server_volume(self.h, ticket, C.byref(v)), "kf_server_volume")
        n = self.configs[ticket].volume_n
        fw = np.empty(n ** 3 * 2, np.float32)
        check(lib().kf_volume_download(v, _p(fw)), "kf_volume_download")
        return fw[0::2].copy(), fw[1::2].copy()

    def checkpoint(self, ticket, path):
        check(lib().kf_server_checkpoint(self.h, ticket, str(path).encode()), "kf_server_checkpoint")

    def restore(self, ticket, path):
        check(lib().kf_server_restore(self.h, ticket, str(path).encode()), "kf_server_restore")

    def stats(self):
        st = kf_server_stats()
        check(lib().kf_server_stats_get(self.h, C.byref(st)), "kf_server_stats_get")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def latencies(self):
        n = C.c_size_t()
        check(lib().kf_server_latencies(self.h, None, 0, C.byref(n)), "kf_server_latencies")
        out = np.zeros(n.value, np.float64)
        check(lib().kf_server_latencies(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)),
              "kf_server_latencies")
        return out

    def run_load(self, configs, banks, frames_per_session, opens_s, rate_hz=30.0, jitter=0.1, seed=7,
                 keep_open=False):
        """Drives the loop with a seeded Poisson load (kf_server_run_load); banks[k] is a
        (frames, h, w) uint16 array replayed ping-pong.  Returns (tickets, stats)."""
        S = len(configs)
        cfg_arr = (kf_session_config * S)(*configs)
        banks = [np.ascontiguousarray(b, np.uint16) for b in banks]
        bank_ptrs = (C.c_void_p * S)(*[b.ctypes.data for b in banks])
        blen = (C.c_int * S)(*[b.shape[0] for b in banks])
        fps = (C.c_int * S)(*[int(x) for x in frames_per_session])
        opens = (C.c_double * S)(*[float(x) for x in opens_s])
        L = kf_poisson_load()
        L.sessions = S
        L.configs = C.cast(cfg_arr, C.c_void_p)
        L.banks = C.cast(bank_ptrs, C.c_void_p)
        L.bank_len = C.cast(blen, C.c_void_p)
        L.frames_per_session = C.cast(fps, C.c_void_p)
        L.opens_s = C.cast(opens, C.c_void_p)
        L.rate_hz, L.jitter, L.seed, L.keep_open = rate_hz, jitter, seed, int(keep_open)
        tickets = (C.c_int * S)()
        st = kf_server_stats()
        check(lib().kf_server_run_load(self.h, C.byref(L), tickets, C.byref(st)), "kf_server_run_load")
        for k in range(S):
            self.configs[tickets[k]] = configs[k]
        return list(tickets), {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().kf_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():  # the HIP runtime may already be torn down; the process frees everything
            return
        try:
            self.close()
        except Exception:
            pass


class TumSequence:
    """A TUM RGB-D sequence (kf_tum_*, rgbd_dataset.cpp:6-97): frame(i) -> (uint16 depth, pose).
    gt_offset=1 reproduces the reference's pairing of frame i with the pose of frame i + 1."""

    def __init__(self, path, gt_offset=0):
        h = C.c_void_p()
        check(lib().kf_tum_open(str(path).encode(), C.byref(h)), "kf_tum_open")
        self.h = h
        self.gt_offset = gt_offset
        n, w, hh = C.c_int(), C.c_int(), C.c_int()
        check(lib().kf_tum_size(h, C.byref(n), C.byref(w), C.byref(hh)), "kf_tum_size")
        self.n, self.width, self.height = n.value, w.value, hh.value

    def __len__(self):
        return self.n

    def frame(self, i):
        depth = np.zeros((self.height, self.width), np.uint16)
        pose = np.zeros(16, np.float32)
        check(lib().kf_tum_frame(self.h, i, self.gt_offset, _p(depth), _p(pose)), "kf_tum_frame")
        return depth, pose

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().kf_tum_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():  # the HIP runtime may already be torn down; the process frees everything
            return
        try:
            self.close()
        except Exception:
            pass


def pose_error(pose1, pose2):
    """poseError (general_helper.cpp:11-21) of two column-major camera -> world poses."""
    a, d = C.c_float(), C.c_float()
    p1 = np.ascontiguousarray(pose1, np.float32).reshape(-1)
    p2 = np.ascontiguousarray(pose2, np.float32).reshape(-1)
    check(lib().kf_pose_error(_p(p1), _p(p2), C.byref(a), C.byref(d)), "kf_pose_error")
    return a.value, d.value

