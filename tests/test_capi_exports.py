"""The C-ABI library loads without a GPU and exports every entry point that
include/kfusion.h declares (and the Python binding declares nothing else)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kfusion.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(kf_\w+)\s*\(", text, flags=re.M)))


def test_header_lists_the_reference_surface():
    names = declared_functions()
    for must in ("kf_convert_to_depth_meters", "kf_apply_bilateral_filter", "kf_down_sample", "kf_create_vertex_map",
                 "kf_create_normal_map", "kf_initialize_grid", "kf_fuse", "kf_raycast", "kf_construct_icp_residuals",
                 "kf_solve_normal_equations", "kf_icp_compute_pose", "kf_pipeline_frame", "kf_volume_save",
                 "kf_volume_load"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from kfusion import _capi
    assert os.path.exists(_capi.LIB_PATH), "libkfusion.so not built"
    lib = C.CDLL(_capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_capi.SIGNATURES) == declared_functions()


def test_status_strings_without_device():
    from kfusion import _capi
    lib = _capi.load()
    assert lib.kf_status_string(0) == b"ok"
    assert b"finite" in lib.kf_status_string(_capi.KF_ERR_NONFINITE)


def test_sensor_math_without_device():
    """kf_sensor_init / set_pose are host-only: intrinsics per level (sensor.h:14-19) and the rigid inverse."""
    import numpy as np
    from kfusion import _capi
    lib = _capi.load()
    s = _capi.kf_sensor()
    assert lib.kf_sensor_init(C.byref(s), 525.0, 525.0, 319.5, 239.5) == 0
    for lvl, (f, cx, cy) in enumerate([(525, 319.5, 239.5), (262.5, 159.5, 119.5), (131.25, 79.5, 59.5)]):
        K = np.array(s.intr[lvl][:])
        assert K[0] == f and K[4] == f and K[6] == cx and K[7] == cy
    a = 0.3
    P = np.eye(4)
    P[:3, :3] = [[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]]
    P[:3, 3] = [0.1, -0.2, 0.3]
    pose = np.ascontiguousarray(P.T.reshape(-1), np.float32)
    assert lib.kf_sensor_set_pose(C.byref(s), pose.ctypes.data_as(C.c_void_p)) == 0
    inv = np.array(s.inv_pose[:]).reshape(4, 4).T
    np.testing.assert_allclose(inv, np.linalg.inv(P), atol=1e-6)


def test_oracle_and_library_intrinsics_agree():
    """The library's host-side glm::inverse restatement equals the oracle's bit for bit."""
    import numpy as np
    from kfusion import _capi
    from oracle.kfref import Oracle
    lib = _capi.load()
    o = Oracle()
    s = _capi.kf_sensor()
    lib.kf_sensor_init(C.byref(s), 517.3, 516.5, 318.6, 255.3)
    for lvl in range(3):
        K, Ki = o.intrinsics(517.3, 516.5, 318.6, 255.3, lvl)
        np.testing.assert_array_equal(np.array(s.intr[lvl][:], np.float32), K)
        np.testing.assert_array_equal(np.array(s.inv_intr[lvl][:], np.float32), Ki)


def test_icp_map_size_limit_without_device():
    """ICP maps above 2^24 pixels are refused before any device call (the kernels
    take pixel coordinates from an FP32 quotient that is exact only below it)."""
    from kfusion import _capi
    lib = _capi.load()
    cfg = _capi.kf_icp_config()
    cfg.iters_per_layer[0] = 1
    cfg.levels = 1
    cfg.width, cfg.height = 8192, 4097  # one row more than 2^24 pixels
    out = C.c_void_p()
    assert lib.kf_icp_create(C.byref(cfg), C.byref(out)) == _capi.KF_ERR_DIMS
    assert b"2^24" in lib.kf_last_error()
    buf = C.c_void_p(1)  # never dereferenced: the size check comes first
    sums = (C.c_double * 27)()
    n = C.c_int64()
    m9 = (C.c_float * 9)()
    m3 = (C.c_float * 3)()
    rc = lib.kf_icp_reduce(buf, buf, buf, 8192, 4097, m9, m3, m9, m3, m9, C.c_float(0.1), C.c_float(0.3), 0, sums,
                           C.byref(n), None, None)
    assert rc == _capi.KF_ERR_DIMS
