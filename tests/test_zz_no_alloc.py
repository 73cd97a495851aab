"""kf_solve_normal_equations_noalloc (the allocation-free LinearLeastSquares::solve,
linear_least_squares.cpp:23-44) against kf_solve_normal_equations through the C ABI:
the same solution bits and status on two calls in a row, the second smaller (the
fixed working array keeps the first call's contents), a system that is not positive
definite, and the size limit rejected on the host."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


def solve_both(kf, A, b):
    lib = kf.lib()
    n = A.shape[0]
    dA, db = kf.DeviceArray.from_host(A), kf.DeviceArray.from_host(b)
    x_ref = np.zeros(6, np.float32)
    st_ref = lib.kf_solve_normal_equations(dA.ptr, db.ptr, n, x_ref.ctypes.data_as(C.c_void_p), None, None)
    x_dev, st_dev = kf.DeviceArray(6, np.float32), kf.DeviceArray(1, np.int32)
    x_dev.upload(np.full(6, np.nan, np.float32))
    st_dev.upload(np.array([-1], np.int32))
    rc = lib.kf_solve_normal_equations_noalloc(dA.ptr, db.ptr, n, x_dev.ptr, st_dev.ptr, None)
    assert rc == kf.KF_OK
    kf.check(lib.kf_device_synchronize(), "sync")
    x, st = x_dev.download(), int(st_dev.download()[0])
    for d in (dA, db, x_dev, st_dev):
        d.free()
    return x_ref, st_ref, x, st


def test_noalloc_solve_matches_the_allocating_solve(kf):
    rng = np.random.default_rng(23)
    for n in (640 * 480, 1000):  # the second call is smaller than the first
        announce(f"solve without allocation, {n} equations")
        A = rng.normal(size=(n, 6)).astype(np.float32)
        A[rng.uniform(size=n) < 0.2] = 0.0
        x_true = rng.normal(size=6).astype(np.float32)
        b = (A @ x_true + 1e-3 * rng.normal(size=n)).astype(np.float32)
        x_ref, st_ref, x, st = solve_both(kf, A, b)
        assert st_ref == kf.KF_OK and st == kf.KF_OK
        np.testing.assert_array_equal(x.view(np.uint32), x_ref.view(np.uint32))
        np.testing.assert_allclose(x, x_true, atol=1e-3)


def test_noalloc_solve_reports_not_positive_definite(kf):
    announce("solve without allocation, singular system")
    A = np.zeros((300, 6), np.float32)
    A[:, :3] = np.random.default_rng(5).normal(size=(300, 3))
    b = np.ones(300, np.float32)
    x_ref, st_ref, x, st = solve_both(kf, A, b)
    assert st_ref == kf.KF_ERR_NOT_PD and st == kf.KF_ERR_NOT_PD
    np.testing.assert_array_equal(x.view(np.uint32), x_ref.view(np.uint32))


def test_noalloc_solve_rejects_oversize_on_the_host(kf):
    lib = kf.lib()
    dummy = kf.DeviceArray(64, np.float32)
    rc = lib.kf_solve_normal_equations_noalloc(dummy.ptr, dummy.ptr, (1 << 22) + 1, dummy.ptr, None, None)
    assert rc == kf.KF_ERR_ARG
    assert lib.kf_solve_normal_equations_noalloc(dummy.ptr, dummy.ptr, 0, dummy.ptr, None, None) == kf.KF_ERR_ARG
    dummy.free()
