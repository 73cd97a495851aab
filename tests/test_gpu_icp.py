"""GPU parity of the ICP (project/src/icp.cu, icp.cpp, linear_least_squares.cpp):
the reference's KATs through the C ABI, residual rows bit-exact against the
oracle, the fused 27-term reduction against FP64 sums, and the device-resident
coarse-to-fine pose against the oracle's (<= 1e-4 rad / 1e-4 m)."""
import json
import os

import numpy as np
import pytest

from gpu_util import announce, rot_err, rt, small_sequence

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


def test_residuals_kat(kf):
    kat = KATS["residuals_2x2"]
    announce("constructIcpResiduals KAT 2x2 (icp_tests.cpp:159-231)")
    maps = {}
    for key in ("vertices", "target_vertices", "target_normals"):
        m = kf.GridMap(2, 2, 4)
        m.upload(np.asarray(kat[key], np.float32).reshape(2, 2, 4))
        maps[key] = m
    A = kf.DeviceArray((4, 6))
    b = kf.DeviceArray((4,))
    I3 = np.eye(3, dtype=np.float32).reshape(-1)
    z = np.zeros(3, np.float32)
    K = np.asarray(KATS["intrinsics_cols"], np.float32).reshape(-1)
    kf.kernel.construct_icp_residuals(maps["vertices"], maps["target_vertices"], maps["target_normals"], I3, z, I3,
                                      z, K, kat["distance_thresh"], kat["angle_thresh"], A, b)
    np.testing.assert_array_equal(A.download(), np.asarray(kat["expected_A"], np.float32))
    np.testing.assert_array_equal(b.download(), np.asarray(kat["expected_b"], np.float32))


def test_solve_kat(kf):
    kat = KATS["solve"]
    announce("LinearLeastSquares::solve KAT 7x6 (icp_tests.cpp:129-157)")
    A = kf.DeviceArray.from_host(np.asarray(kat["A"], np.float32))
    b = kf.DeviceArray.from_host(np.asarray(kat["b"], np.float32))
    x = kf.LinearLeastSquares().solve(A, b, 7)
    np.testing.assert_allclose(x, kat["expected"], atol=kat["tolerance"])
    with pytest.raises(kf.KfError) as e:  # all-zero system: LLT fails (ignored by the reference)
        kf.LinearLeastSquares().solve(kf.DeviceArray.from_host(np.zeros((7, 6), np.float32)),
                                      kf.DeviceArray.from_host(np.zeros(7, np.float32)), 7)
    assert e.value.status == kf.KF_ERR_NOT_PD


def model_maps(kf, oracle32, seed=11, w=320, h=240, n=128, L=3.0, mue=0.03, frames=3):
    """Fuses `frames` frames on the oracle, raycasts the model at the last pose, and
    builds the next frame's measured pyramids: the inputs of one frame's ICP."""
    seq = small_sequence(seed=seed, n_frames=frames + 1, w=w, h=h, L=L)
    raws = seq.frames()
    Ks = [oracle32.intrinsics(seq.fx, seq.fy, seq.cx, seq.cy, l) for l in range(3)]
    F, W = oracle32.new_volume(n)
    for f in range(frames):
        d = oracle32.convert_to_depth_meters(raws[f], 1e-3)
        R, t = rt(seq.pose(f))
        oracle32.fuse(F, W, n, L, mue, d, oracle32.rigid_inverse(R, t), Ks[0][0])
    P = seq.pose(frames - 1)
    tv, tn, vp = [], [], []
    d = oracle32.apply_bilateral_filter(oracle32.convert_to_depth_meters(raws[frames], 1e-3))
    for l in range(3):
        a, b = oracle32.raycast(F, n, L, mue, P, Ks[l][1], w >> l, h >> l)
        tv.append(a)
        tn.append(b)
        if l:
            d = oracle32.down_sample(d)
        vp.append(oracle32.create_vertex_map(d, Ks[l][1]))
    return seq, Ks, P, vp, tv, tn


def upload_pyr(kf, maps):
    out = []
    for m in maps:
        g = kf.GridMap(m.shape[1], m.shape[0], 4)
        g.upload(m)
        out.append(g)
    return out


def test_residual_rows_bit_exact(kf, oracle32):
    seq, Ks, P, vp, tv, tn = model_maps(kf, oracle32)
    R, t = rt(P)
    # a perturbed current estimate, as in a mid-ICP iteration
    Rc = R.copy()
    tc = t + np.array([0.004, -0.003, 0.002], np.float32)
    for l in range(3):
        announce(f"constructIcpResiduals level {l} seed 11")
        h, w = vp[l].shape[:2]
        V, TV, TN = upload_pyr(kf, [vp[l], tv[l], tn[l]])
        A = kf.DeviceArray((h * w, 6))
        b = kf.DeviceArray((h * w,))
        th = np.deg2rad(20.0)
        kf.kernel.construct_icp_residuals(V, TV, TN, R, t, Rc, tc, Ks[l][0], 0.1, th, A, b)
        Ar, br = oracle32.construct_icp_residuals(vp[l], tv[l], tn[l], R, t, Rc, tc, Ks[l][0], 0.1, th)
        Ag, bg = A.download(), b.download()
        assert np.count_nonzero(br) > 0.3 * h * w
        np.testing.assert_array_equal(Ag.view(np.uint32), Ar.view(np.uint32))
        np.testing.assert_array_equal(bg.view(np.uint32), br.view(np.uint32))
        # the fused reduction (plain and compensated) against the FP64 sums of the same rows
        ata, atb, cnt = oracle32.normal_equations(Ar, br)
        iu = np.triu_indices(6)
        ref = np.concatenate([ata[iu], atb])
        for comp in (False, True):
            sums, inl = kf.kernel.icp_reduce(V, TV, TN, R, t, Rc, tc, Ks[l][0], 0.1, th, compensated=comp)
            assert inl == np.count_nonzero(np.any(Ar != 0, axis=1))
            scale = np.abs(ref).max()
            tol = (2e-6 if comp else 2e-5) * scale
            np.testing.assert_allclose(sums, ref, rtol=0, atol=tol, err_msg=f"compensated={comp}")


def test_residual_rows_threshold_sweep(kf, oracle32):
    """The distance and angle tests are decided from cheap forms except inside a
    narrow band around the threshold, where they are evaluated exactly.  Dense
    sweeps of both thresholds over the range the pixels span put pixels inside
    those bands; every row must still equal the oracle's bit for bit."""
    seq, Ks, P, vp, tv, tn = model_maps(kf, oracle32)
    R, t = rt(P)
    Rc = R.copy()
    tc = t + np.array([0.004, -0.003, 0.002], np.float32)
    h, w = vp[0].shape[:2]
    V, TV, TN = upload_pyr(kf, [vp[0], tv[0], tn[0]])
    A = kf.DeviceArray((h * w, 6))
    b = kf.DeviceArray((h * w,))
    sweeps = [(float(d), np.deg2rad(20.0)) for d in np.geomspace(0.002, 0.1, 24)]
    sweeps += [(0.1, float(np.deg2rad(a))) for a in np.linspace(1.0, 45.0, 24)]
    accepted = []
    for dist, ang in sweeps:
        announce(f"constructIcpResiduals threshold sweep distance {dist:.4f} m angle {np.rad2deg(ang):.2f} deg")
        kf.kernel.construct_icp_residuals(V, TV, TN, R, t, Rc, tc, Ks[0][0], dist, ang, A, b)
        Ar, br = oracle32.construct_icp_residuals(vp[0], tv[0], tn[0], R, t, Rc, tc, Ks[0][0], dist, ang)
        np.testing.assert_array_equal(A.download().view(np.uint32), Ar.view(np.uint32))
        np.testing.assert_array_equal(b.download().view(np.uint32), br.view(np.uint32))
        accepted.append(int(np.count_nonzero(np.any(Ar != 0, axis=1))))
    # the sweeps move the decisions: the accepted counts grow with both thresholds
    assert accepted[0] < accepted[23] and accepted[24] < accepted[-1]


def test_compute_pose_matches_oracle(kf, oracle32):
    seq, Ks, P, vp, tv, tn = model_maps(kf, oracle32, seed=13)
    R, t = rt(P)
    iters = (10, 5, 4)
    th = float(np.deg2rad(20.0))
    ref = oracle32.compute_pose(vp, tv, tn, iters, [k[0] for k in Ks], 0.1, th, 0.0, 0.0, R, t)
    assert ref["status"] == 0 and ref["n_iters"] == 19
    for comp in (False, True):
        announce(f"ICP::computePose 320x240 seed 13 compensated={comp}")
        cfg = kf.IcpConfig(iters, 320, 240, 0.1, th, 0.0, 0.0, compensated=comp)
        icp = kf.ICP(cfg)
        sensor = kf.Sensor(seq.fx, seq.fy, seq.cx, seq.cy)
        sensor.set_pose(P)
        out = icp.compute_pose(upload_pyr(kf, vp), upload_pyr(kf, tv), upload_pyr(kf, tn), sensor)
        assert icp.iterations() == 19
        assert rot_err(out.rot_mat, ref["R"]) <= 1e-4
        assert np.abs(out.transl_vec - ref["t"]).max() <= 1e-4
        times = icp.get_execution_times()
        assert times[0] > 0 and times[1] > 0


def test_compute_pose_early_stop_and_failure(kf, oracle32):
    seq, Ks, P, vp, tv, tn = model_maps(kf, oracle32, seed=13)
    R, t = rt(P)
    announce("ICP early stop (0.1 deg chord / 1 cm) and all-invalid target")
    cfg = kf.IcpConfig((10, 5, 4), 320, 240, 0.1, float(np.deg2rad(20.0)), float(np.deg2rad(0.1)), 0.01)
    icp = kf.ICP(cfg)
    sensor = kf.Sensor(seq.fx, seq.fy, seq.cx, seq.cy)
    sensor.set_pose(P)
    icp.compute_pose(upload_pyr(kf, vp), upload_pyr(kf, tv), upload_pyr(kf, tn), sensor)
    assert icp.iterations() < 19
    # no valid correspondences: the normal equations are zero; the reference throws (icp.cpp:67-73)
    bad = [np.zeros_like(m) for m in tv]
    for m in bad:
        m[..., 3] = -2
    with pytest.raises(kf.KfError) as e:
        icp.compute_pose(upload_pyr(kf, vp), upload_pyr(kf, bad), upload_pyr(kf, bad), sensor)
    assert e.value.status in (kf.KF_ERR_NOT_PD, kf.KF_ERR_NONFINITE)
