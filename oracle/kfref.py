"""ctypes binding of the CPU oracle (oracle/kfref.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of ``bench.py``, as the checker / CPU baseline.  The
product library never imports it.

Arrays are numpy; float4 maps are (h, w, 4) float32 (float64 for the FP64
build), matrices are glm column-major flat arrays (9 or 16 values).
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libkfref.so")
REF_LLT_PATH = os.path.join(_HERE, "_ref", "libref_llt.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def ref_llt():
    """Eigen 3.3.7 LLT compiled from the reference tree, or None if not built."""
    if not os.path.exists(REF_LLT_PATH):
        return None
    l = C.CDLL(REF_LLT_PATH)
    l.ref_llt_solve6_f32.argtypes = [C.c_void_p] * 3
    l.ref_llt_solve6_f32.restype = C.c_int
    l.ref_llt_solve6_f64.argtypes = [C.c_void_p] * 3
    l.ref_llt_solve6_f64.restype = C.c_int
    return l


P = C.c_void_p
I = C.c_int
I64 = C.c_int64
F = C.c_float
D = C.c_double


def _declare(l):
    for pre, R in (("kfo_", F), ("kfo64_", D)):
        def fn(name, res, *args):
            f = getattr(l, pre + name)
            f.restype = res
            f.argtypes = list(args)
        fn("u16_to_half_float", F, C.c_uint16)
        fn("cos_threshold", F, F)
        fn("intrinsics", None, R, R, R, R, I, P, P)
        fn("convert_to_depth_meters", None, P, P, I, I, R, I)
        fn("apply_bilateral_filter", None, P, P, I, I)
        fn("down_sample", None, P, I, I, P)
        fn("create_vertex_map", None, P, I, I, P, P)
        fn("create_normal_map", None, P, I, I, P)
        fn("compute_normal", None, P, I, I, I, I, P)
        fn("initialize_grid", None, P, P, I, R, R)
        fn("fuse", I64, P, P, I, R, R, P, I, I, P, P)
        fn("raycast", None, P, I, R, R, P, P, I, I, P, P)
        fn("sample_tsdf", R, P, I, R, R, R)
        fn("construct_icp_residuals", None, P, P, P, I, I, P, P, P, P, P, R, R, P, P)
        fn("normal_equations", I64, P, P, I64, P, P)
        fn("update_pose", None, P, P, P)
        fn("pose_error", None, P, P, P, P, P)
        fn("rigid_inverse", None, P, P, P)
        fn("compute_pose", I, P, P, P, I, I, I, P, P, R, R, R, R, P, P, P, P, P, P)
        fn("interleave", None, P, P, I64, P)
        fn("pipeline_create", P, I, I, I, R, R, R, R, R, R, R, P, R, R, R, R, P)
        fn("pipeline_destroy", None, P)
        fn("pipeline_frame", I, P, P)
        fn("pipeline_pose", None, P, P)
        fn("pipeline_tsdf", P, P)
        fn("pipeline_weight", P, P)
        fn("pipeline_map", P, P, I, I)
        fn("pipeline_depth", P, P, I, I)
        fn("pipeline_stage_ms", D, P, I)
        fn("pipeline_touched", I64, P)
        fn("pipeline_use_count_weights", I, P)
        fn("fuse_counts", I64, P, P, I, R, R, P, I, I, P, P)
    l.kfo_fuse_q.restype = I64
    l.kfo_fuse_q.argtypes = [P, P, I, F, F, P, I, I, P, P]
    l.kfo_initialize_grid_q.restype = None
    l.kfo_initialize_grid_q.argtypes = [P, P, I, F, F]
    l.kfo_raycast_q.restype = None
    l.kfo_raycast_q.argtypes = [P, I, F, F, P, P, I, I, P, P]
    l.kfo_raycast_footprint.restype = I64
    l.kfo_raycast_footprint.argtypes = [P, I, I, F, F, P, P, I, I, P]
    l.kfo_count_marked.restype = I64
    l.kfo_count_marked.argtypes = [P, I64]
    l.kfo_pipeline_use_snorm16.restype = I
    l.kfo_pipeline_use_snorm16.argtypes = [P]
    l.kfo_pipeline_tsdf_q.restype = P
    l.kfo_pipeline_tsdf_q.argtypes = [P]
    l.kfo_pipeline_weight_q.restype = P
    l.kfo_pipeline_weight_q.argtypes = [P]
    l.kfo_threads.restype = I
    l.kfo_threads.argtypes = []
    l.kfo_set_threads.restype = None
    l.kfo_set_threads.argtypes = [I]
    l.kfo_expf.restype = F
    l.kfo_expf.argtypes = [F]
    l.kfo_cholesky_solve6.restype = I
    l.kfo_cholesky_solve6.argtypes = [P, P, P]
    l.kfo_llt_solve6_f32.restype = I
    l.kfo_llt_solve6_f32.argtypes = [P, P, P]
    l.kfo_normals_too_different.restype = I
    l.kfo_normals_too_different.argtypes = [P, P, P, F]
    l.kfo_normals_too_different_acos.restype = I
    l.kfo_normals_too_different_acos.argtypes = [P, P, P, F]
    l.kfo_correspondence.restype = None
    l.kfo_correspondence.argtypes = [P, P, P, P, P, P]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One precision of the oracle: Oracle(np.float32) or Oracle(np.float64)."""

    def __init__(self, dtype=np.float32):
        self.dt = np.dtype(dtype)
        self.pre = "kfo_" if self.dt == np.float32 else "kfo64_"
        self.l = lib()

    def f(self, name):
        return getattr(self.l, self.pre + name)

    def arr(self, a):
        return np.ascontiguousarray(a, dtype=self.dt)

    # --- scalars / small helpers
    def intrinsics(self, fx, fy, cx, cy, level):
        K = np.zeros(9, self.dt)
        Ki = np.zeros(9, self.dt)
        self.f("intrinsics")(fx, fy, cx, cy, level, _p(K), _p(Ki))
        return K, Ki

    # --- measurement
    def convert_to_depth_meters(self, raw, scale, unmirror=False):
        raw = np.ascontiguousarray(raw, dtype=np.uint16)
        h, w = raw.shape
        out = np.zeros((h, w), self.dt)
        self.f("convert_to_depth_meters")(_p(raw), _p(out), w, h, scale, int(unmirror))
        return out

    def apply_bilateral_filter(self, d):
        d = self.arr(d)
        h, w = d.shape
        out = np.zeros_like(d)
        self.f("apply_bilateral_filter")(_p(d), _p(out), w, h)
        return out

    def down_sample(self, d):
        d = self.arr(d)
        h, w = d.shape
        out = np.zeros((h // 2, w // 2), self.dt)
        self.f("down_sample")(_p(d), w, h, _p(out))
        return out

    def create_vertex_map(self, d, Kinv):
        d = self.arr(d)
        h, w = d.shape
        out = np.zeros((h, w, 4), self.dt)
        self.f("create_vertex_map")(_p(d), w, h, _p(self.arr(Kinv)), _p(out))
        return out

    def create_normal_map(self, v):
        v = self.arr(v)
        h, w, _ = v.shape
        out = np.zeros_like(v)
        self.f("create_normal_map")(_p(v), w, h, _p(out))
        return out

    # --- tsdf
    def new_volume(self, n, f0=2.0, w0=0.0):
        F = np.empty(n * n * n, self.dt)
        W = np.empty(n * n * n, self.dt)
        self.f("initialize_grid")(_p(F), _p(W), n, f0, w0)
        return F, W

    def new_volume_q(self, n, f0=2.0, w0=0.0):
        """Precision-form volume: snorm16 F, uint8 W (FP32 oracle only)."""
        F = np.empty(n * n * n, np.int16)
        W = np.empty(n * n * n, np.uint8)
        self.l.kfo_initialize_grid_q(_p(F), _p(W), n, f0, w0)
        return F, W

    def fuse_q(self, F, W, n, width, mue, depth, inv_pose, K):
        depth = np.ascontiguousarray(depth, np.float32)
        h, w = depth.shape
        return self.l.kfo_fuse_q(_p(F), _p(W), n, width, mue, _p(depth), w, h,
                                 _p(np.ascontiguousarray(inv_pose, np.float32)), _p(np.ascontiguousarray(K, np.float32)))

    def raycast_q(self, F, n, width, mue, pose, Kinv, w, h, vout=None, nout=None):
        if vout is None:
            vout = np.zeros((h, w, 4), np.float32)
            vout[..., 3] = -2
        if nout is None:
            nout = np.zeros((h, w, 4), np.float32)
            nout[..., 3] = -2
        self.l.kfo_raycast_q(_p(F), n, width, mue, _p(np.ascontiguousarray(pose, np.float32)),
                             _p(np.ascontiguousarray(Kinv, np.float32)), w, h, _p(vout), _p(nout))
        return vout, nout

    def fuse_counts(self, F, W, n, width, mue, depth, inv_pose, K):
        """Count-weight update (the precision form's rule) on unquantised planes."""
        depth = self.arr(depth)
        h, w = depth.shape
        return self.f("fuse_counts")(_p(F), _p(W), n, width, mue, _p(depth), w, h,
                                     _p(self.arr(inv_pose)), _p(self.arr(K)))

    def fuse(self, F, W, n, width, mue, depth, inv_pose, K):
        depth = self.arr(depth)
        h, w = depth.shape
        return self.f("fuse")(_p(F), _p(W), n, width, mue, _p(depth), w, h,
                              _p(self.arr(inv_pose)), _p(self.arr(K)))

    def raycast(self, F, n, width, mue, pose, Kinv, w, h, vout=None, nout=None):
        if vout is None:
            vout = np.zeros((h, w, 4), self.dt)
            vout[..., 3] = -2
        if nout is None:
            nout = np.zeros((h, w, 4), self.dt)
            nout[..., 3] = -2
        self.f("raycast")(_p(F), n, width, mue, _p(self.arr(pose)), _p(self.arr(Kinv)), w, h,
                          _p(vout), _p(nout))
        return vout, nout

    def raycast_footprint(self, F, n, width, mue, pose, Kinv_levels, w, h):
        """(distinct voxels, samples) of the reference march over the levels whose inverse
        intrinsics are given (level l: w >> l by h >> l); F float32 or int16 snorm16 codes."""
        F = np.ascontiguousarray(F)
        fmt = 1 if F.dtype == np.int16 else 0
        if fmt == 0:
            F = np.ascontiguousarray(F, np.float32)
        mark = np.zeros(n ** 3, np.uint8)
        samples = 0
        for lvl, Ki in enumerate(Kinv_levels):
            samples += self.l.kfo_raycast_footprint(_p(F), fmt, n, width, mue,
                                                    _p(np.ascontiguousarray(pose, np.float32)),
                                                    _p(np.ascontiguousarray(Ki, np.float32)), w >> lvl, h >> lvl,
                                                    _p(mark))
        return int(self.l.kfo_count_marked(_p(mark), mark.size)), int(samples)

    def sample_tsdf(self, F, n, u, v, s):
        return self.f("sample_tsdf")(_p(F), n, u, v, s)

    # --- icp
    def construct_icp_residuals(self, V, TV, TN, Rp, tp, Rc, tc, K, dist_th, angle_th):
        V, TV, TN = self.arr(V), self.arr(TV), self.arr(TN)
        h, w, _ = V.shape
        A = np.zeros((h * w, 6), self.dt)
        b = np.zeros(h * w, self.dt)
        cos_th = self.f("cos_threshold")(angle_th)
        self.f("construct_icp_residuals")(_p(V), _p(TV), _p(TN), w, h, _p(self.arr(Rp)), _p(self.arr(tp)),
                                          _p(self.arr(Rc)), _p(self.arr(tc)), _p(self.arr(K)),
                                          dist_th, cos_th, _p(A), _p(b))
        return A, b

    def normal_equations(self, A, b):
        A, b = self.arr(A), self.arr(b)
        ata = np.zeros(36, np.float64)
        atb = np.zeros(6, np.float64)
        cnt = self.f("normal_equations")(_p(A), _p(b), b.shape[0], _p(ata), _p(atb))
        return ata.reshape(6, 6).T.copy(), atb, cnt

    def update_pose(self, Rm, t, x):
        Rm, t = self.arr(Rm).copy(), self.arr(t).copy()
        self.f("update_pose")(_p(Rm), _p(t), _p(self.arr(x)))
        return Rm, t

    def pose_error(self, R1, t1, R2, t2):
        out = np.zeros(2, self.dt)
        self.f("pose_error")(_p(self.arr(R1)), _p(self.arr(t1)), _p(self.arr(R2)), _p(self.arr(t2)), _p(out))
        return out

    def rigid_inverse(self, Rm, t):
        inv = np.zeros(16, self.dt)
        self.f("rigid_inverse")(_p(self.arr(Rm)), _p(self.arr(t)), _p(inv))
        return inv

    def compute_pose(self, vpyr, tvpyr, tnpyr, iters, Ks, dist_th, angle_th, stop_a, stop_d, Rin, tin,
                     log_sums=False):
        levels = len(vpyr)
        keep = [self.arr(x) for x in list(vpyr) + list(tvpyr) + list(tnpyr)]
        arr_t = C.c_void_p * levels
        vp = arr_t(*[_p(x) for x in keep[:levels]])
        tv = arr_t(*[_p(x) for x in keep[levels:2 * levels]])
        tn = arr_t(*[_p(x) for x in keep[2 * levels:]])
        h0, w0, _ = keep[0].shape
        it = np.array(iters, np.int32)
        Kall = np.ascontiguousarray(np.stack([self.arr(k) for k in Ks]))
        Ro = np.zeros(9, self.dt)
        to = np.zeros(3, self.dt)
        log = np.zeros((int(sum(iters)), 27), np.float64)
        nit = np.zeros(1, np.int32)
        st = self.f("compute_pose")(vp, tv, tn, w0, h0, levels, _p(it), _p(Kall), dist_th, angle_th, stop_a,
                                    stop_d, _p(self.arr(Rin)), _p(self.arr(tin)), _p(Ro), _p(to),
                                    _p(log) if log_sums else None, _p(nit))
        out = dict(status=st, R=Ro, t=to, n_iters=int(nit[0]))
        if log_sums:
            out["sums"] = log[: int(nit[0])]
        return out


class OraclePipeline:
    """The reference's per-frame loop (kinect_fusion.cpp:44-183) on the CPU."""

    def __init__(self, w, h, n, width, mue, fx, fy, cx, cy, depth_scale, iters, dist_th, angle_th,
                 stop_angle=0.0, stop_dist=0.0, init_pose=None, dtype=np.float32, snorm16=False,
                 count_weights=False):
        self.o = Oracle(dtype)
        self.w, self.h, self.n = w, h, n
        it = np.array(iters, np.int32)
        pose = self.o.arr(np.eye(4).T.reshape(-1) if init_pose is None else init_pose)
        self.h_ = self.o.f("pipeline_create")(w, h, n, width, mue, fx, fy, cx, cy, depth_scale, _p(it),
                                              dist_th, angle_th, stop_angle, stop_dist, _p(pose))
        if not self.h_:
            raise MemoryError("oracle pipeline allocation failed")
        self.snorm16 = bool(snorm16)
        if snorm16:
            assert self.o.dt == np.float32
            if self.o.l.kfo_pipeline_use_snorm16(self.h_) != 0:
                raise MemoryError("oracle precision-form volume allocation failed")
        elif count_weights:
            if self.o.f("pipeline_use_count_weights")(self.h_) != 0:
                raise RuntimeError("oracle count-weight switch failed")

    def close(self):
        if self.h_:
            self.o.f("pipeline_destroy")(self.h_)
            self.h_ = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def frame(self, raw):
        raw = np.ascontiguousarray(raw, dtype=np.uint16)
        assert raw.shape == (self.h, self.w)
        return self.o.f("pipeline_frame")(self.h_, _p(raw))

    def pose(self):
        p = np.zeros(16, self.o.dt)
        self.o.f("pipeline_pose")(self.h_, _p(p))
        return p

    def _view(self, ptr, count, shape):
        ct = C.c_float if self.o.dt == np.float32 else C.c_double
        buf = (ct * count).from_address(ptr)
        return np.ctypeslib.as_array(buf).reshape(shape).copy()

    def tsdf_q(self):
        """(int16 F codes, uint8 W counts) of the precision form."""
        n3 = self.n ** 3
        f = np.ctypeslib.as_array((C.c_int16 * n3).from_address(self.o.l.kfo_pipeline_tsdf_q(self.h_))).copy()
        w = np.ctypeslib.as_array((C.c_uint8 * n3).from_address(self.o.l.kfo_pipeline_weight_q(self.h_))).copy()
        return f, w

    def tsdf(self):
        n3 = self.n ** 3
        return (self._view(self.o.f("pipeline_tsdf")(self.h_), n3, (n3,)),
                self._view(self.o.f("pipeline_weight")(self.h_), n3, (n3,)))

    def map(self, which, level):
        w, h = self.w >> level, self.h >> level
        return self._view(self.o.f("pipeline_map")(self.h_, which, level), w * h * 4, (h, w, 4))

    def depth(self, which, level=0):
        w, h = self.w >> level, self.h >> level
        return self._view(self.o.f("pipeline_depth")(self.h_, which, level), w * h, (h, w))

    def stage_ms(self, stage):
        return self.o.f("pipeline_stage_ms")(self.h_, stage)

    def touched(self):
        return self.o.f("pipeline_touched")(self.h_)
