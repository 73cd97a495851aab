"""Pins the oracle's ICP pieces (oracle/kfref.c) without a GPU: the reference's
own known-answer tests (test/icp_tests.cpp), Eigen 3.3.7's LLT built from the
reference tree, closed forms of the pose update and brute-force sums."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import kfref

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def cols(m):
    """glm column list -> flat column-major float32[9]."""
    return np.asarray(m, np.float32).reshape(-1)


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_correspondence_kat():
    kat = KATS["correspondence"]
    K = cols(KATS["intrinsics_cols"])
    R = cols(KATS["rotation_90_right_cols"])
    t = np.asarray(kat["translation"], np.float32)
    l = kfref.lib()
    for v, exp in zip(kat["vertices"], kat["expected_uv"]):
        s = np.asarray(v, np.float32)
        uv = np.zeros(2, np.int32)
        fuv = np.zeros(2, np.float32)
        l.kfo_correspondence(p(s), p(R), p(t), p(K), p(uv), p(fuv))
        assert list(uv) == exp


def test_normals_too_different_kat():
    kat = KATS["normals_too_different"]
    R = cols(KATS["rotation_90_right_cols"])
    tn = np.asarray(kat["target_normal"], np.float32)
    l = kfref.lib()
    cos_t = l.kfo_cos_threshold(kat["angle_thresh"])
    for key in ("close", "far"):
        n = np.asarray(kat[key], np.float32)
        assert bool(l.kfo_normals_too_different(p(n), p(tn), p(R), cos_t)) == kat["expected"][key]
        assert bool(l.kfo_normals_too_different_acos(p(n), p(tn), p(R), kat["angle_thresh"])) == kat["expected"][key]


def test_cosine_form_matches_acos_form():
    """acos(d) > th  <=>  (d >= -1 and d < cos th), away from the boundary."""
    rng = np.random.default_rng(5)
    l = kfref.lib()
    R = np.eye(3, dtype=np.float32).reshape(-1)
    agree = 0
    for _ in range(4000):
        n = rng.normal(size=3).astype(np.float32)
        tn = rng.normal(size=3)
        tn = (tn / np.linalg.norm(tn)).astype(np.float32)
        th = np.float32(rng.uniform(0.05, 3.1))
        cos_t = l.kfo_cos_threshold(th)
        d = float(np.dot(n / np.linalg.norm(n), tn))
        if abs(d - np.cos(np.float64(th))) < 1e-5:
            continue
        a = l.kfo_normals_too_different_acos(p(n), p(tn), p(R), th)
        b = l.kfo_normals_too_different(p(n), p(tn), p(R), cos_t)
        assert a == b
        agree += 1
    assert agree > 3900


def kat_residual_maps():
    kat = KATS["residuals_2x2"]
    V = np.asarray(kat["vertices"], np.float32).reshape(2, 2, 4)
    TV = np.asarray(kat["target_vertices"], np.float32).reshape(2, 2, 4)
    TN = np.asarray(kat["target_normals"], np.float32).reshape(2, 2, 4)
    return kat, V, TV, TN


def test_fill_a_b_and_residuals_kat(oracle32):
    kat, V, TV, TN = kat_residual_maps()
    I3 = np.eye(3, dtype=np.float32).reshape(-1)
    z = np.zeros(3, np.float32)
    K = cols(KATS["intrinsics_cols"])
    A, b = oracle32.construct_icp_residuals(V, TV, TN, I3, z, I3, z, K, kat["distance_thresh"], kat["angle_thresh"])
    np.testing.assert_array_equal(A, np.asarray(kat["expected_A"], np.float32))
    np.testing.assert_array_equal(b, np.asarray(kat["expected_b"], np.float32))
    # fill_a / fill_b KATs through a one-pixel configuration: s = vertex, n = normal, d = target
    fa, fb = KATS["fill_a"], KATS["fill_b"]
    s = np.asarray(fa["vertex"], np.float64)
    n = np.asarray(fa["normal"], np.float64)
    row = [n[1] * s[2] - n[2] * s[1], n[2] * s[0] - n[0] * s[2], n[0] * s[1] - n[1] * s[0], *n]
    assert row == fa["expected"]
    d = np.asarray(fb["target_vertex"], np.float64)
    assert float(n @ d - n @ s) == fb["expected"]


def test_solve_kat_and_eigen_cross_check(oracle32):
    kat = KATS["solve"]
    A = np.asarray(kat["A"], np.float32)
    b = np.asarray(kat["b"], np.float32)
    ata, atb, cnt = oracle32.normal_equations(A, b)
    assert cnt == 7
    x = np.zeros(6)
    assert kfref.lib().kfo_cholesky_solve6(p(np.ascontiguousarray(ata.T.reshape(-1))), p(atb), p(x)) == 0
    np.testing.assert_allclose(x, kat["expected"], atol=kat["tolerance"])
    eig = kfref.ref_llt()
    if eig is None:
        pytest.skip("oracle/_ref/libref_llt.so not built (reference tree absent)")
    rng = np.random.default_rng(11)
    for _ in range(200):
        M = rng.normal(size=(40, 6))
        ata = (M.T @ M).astype(np.float32)
        atb = rng.normal(size=6).astype(np.float32)
        a_cm = np.ascontiguousarray(ata.T.reshape(-1))
        xe, xo = np.zeros(6, np.float32), np.zeros(6, np.float32)
        assert eig.ref_llt_solve6_f32(p(a_cm), p(atb), p(xe)) == 0
        assert kfref.lib().kfo_llt_solve6_f32(p(a_cm), p(atb), p(xo)) == 0
        # same unblocked algorithm in FP32: agreement to a few ulps of the solution
        np.testing.assert_allclose(xo, xe, rtol=2e-5, atol=1e-6)
        a64 = np.ascontiguousarray(ata.astype(np.float64).T.reshape(-1))
        b64 = atb.astype(np.float64)
        xd, xe64 = np.zeros(6), np.zeros(6)
        assert kfref.lib().kfo_cholesky_solve6(p(a64), p(b64), p(xd)) == 0
        assert eig.ref_llt_solve6_f64(p(a64), p(b64), p(xe64)) == 0
        np.testing.assert_allclose(xd, xe64, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(xd, np.linalg.solve(ata.astype(np.float64), b64), rtol=1e-8, atol=1e-9)


def test_cholesky_reports_not_positive_definite():
    x = np.zeros(6)
    zero = np.zeros(36)
    assert kfref.lib().kfo_cholesky_solve6(p(zero), p(np.zeros(6)), p(x)) == 1
    m = np.eye(6)
    m[3, 3] = -1
    assert kfref.lib().kfo_cholesky_solve6(p(np.ascontiguousarray(m.T.reshape(-1))), p(np.ones(6)), p(x)) == 4


def rot(axis, a):
    c, s = np.cos(a), np.sin(a)
    if axis == "x":
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def test_update_pose_closed_form(oracle32):
    """R_inc = R_y(x1) R_x(x0) R_z(x2) (glm orientate3(vec3(x0, x2, x1))); R <- R_inc^T R, t <- R_inc^T t + x_t."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        x = rng.uniform(-0.3, 0.3, 6).astype(np.float32)
        R0 = rot("z", rng.uniform(-1, 1)) @ rot("x", rng.uniform(-1, 1))
        t0 = rng.normal(size=3)
        Rn, tn = oracle32.update_pose(R0.T.reshape(-1), t0, x)  # column-major in/out
        xd = x.astype(np.float64)
        Ri = rot("y", xd[1]) @ rot("x", xd[0]) @ rot("z", xd[2])
        np.testing.assert_allclose(Rn.reshape(3, 3).T, Ri.T @ R0, atol=2e-7)
        np.testing.assert_allclose(tn, Ri.T @ t0.astype(np.float32) + xd[3:], atol=1e-6)


def test_update_pose_first_order_sign(oracle32):
    """With A = (n x s, n) the solved rotation is -omega; R_inc^T ~ I + [omega]_x for small angles."""
    w = np.array([1e-4, -2e-4, 3e-4], np.float32)
    x = np.concatenate([w, np.zeros(3, np.float32)])
    Rn, _ = oracle32.update_pose(np.eye(3, dtype=np.float32).reshape(-1), np.zeros(3), x)
    skew = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    np.testing.assert_allclose(Rn.reshape(3, 3).T, np.eye(3) - skew, atol=1e-7)


def test_pose_error_chord(oracle32):
    I = np.eye(3, dtype=np.float32).reshape(-1)
    a = 0.2
    Rz = rot("z", a).astype(np.float32).T.reshape(-1)
    e = oracle32.pose_error(I, np.zeros(3), Rz, np.array([0, 3, 4], np.float32))
    v = np.ones(3) / np.sqrt(3)
    np.testing.assert_allclose(e[0], np.linalg.norm(rot("z", a) @ v - v), rtol=1e-5)
    np.testing.assert_allclose(e[1], 5.0, rtol=1e-6)


def test_normal_equations_brute_force(oracle32):
    rng = np.random.default_rng(7)
    A = rng.normal(size=(5000, 6)).astype(np.float32)
    A[rng.uniform(size=5000) < 0.3] = 0
    b = rng.normal(size=5000).astype(np.float32)
    ata, atb, cnt = oracle32.normal_equations(A, b)
    np.testing.assert_allclose(ata, A.astype(np.float64).T @ A.astype(np.float64), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(atb, A.astype(np.float64).T @ b.astype(np.float64), rtol=1e-12, atol=1e-9)
    assert cnt == int(np.count_nonzero(np.any(A != 0, axis=1) | (b != 0)))


def test_intrinsics_levels(oracle32):
    """sensor.h:14-16 constants for 640x480 and K K^-1 = I."""
    for lvl, (f, cx, cy) in enumerate([(525, 319.5, 239.5), (262.5, 159.5, 119.5), (131.25, 79.5, 59.5)]):
        K, Ki = oracle32.intrinsics(525.0, 525.0, 319.5, 239.5, lvl)
        assert K[0] == f and K[4] == f and K[6] == cx and K[7] == cy and K[8] == 1
        exact = np.linalg.inv(K.reshape(3, 3).T.astype(np.float64))
        np.testing.assert_allclose(Ki.reshape(3, 3).T, exact, rtol=3e-7, atol=1e-9)


def test_glm_layout_kat():
    kat = KATS["glm_layout"]
    assert cols(kat["cols"])[4] == kat["flat_index_4"]


def test_compute_pose_identity_on_identical_maps(oracle32):
    """Aligning a frame with itself (target = its own vertex/normal maps, identity pose):
    every b_i = n.d - n.s is a rounding residue of n.d - n.d, so A^T b is ~0 against A^T A
    and the pose stays at the start pose."""
    from kfusion.synth import SyntheticSequence
    seq = SyntheticSequence(seed=4, n_frames=1, L=3.0, width=80, height=60, fx=65.625, fy=65.625, noise_k=0.0)
    K, Ki = oracle32.intrinsics(65.625, 65.625, 39.5, 29.5, 0)
    d = oracle32.convert_to_depth_meters(seq.frames(0, 1)[0], 1e-3)
    v = oracle32.create_vertex_map(d, Ki)
    nrm = oracle32.create_normal_map(v)
    I3 = np.eye(3, dtype=np.float32).reshape(-1)
    z = np.zeros(3, np.float32)
    r = oracle32.compute_pose([v], [v], [nrm], [3], [K], 0.1, 0.35, 0.0, 0.0, I3, z, log_sums=True)
    assert r["status"] == 0 and r["n_iters"] == 3
    assert np.abs(r["sums"][:, 21:]).max() < 1e-6 * np.abs(r["sums"][:, :21]).max()
    np.testing.assert_allclose(r["R"], I3, atol=1e-6)
    np.testing.assert_allclose(r["t"], z, atol=1e-6)


def _exact_world_maps(seq_args, f, levels=3):
    """World-space vertex and normal pyramids of frame f taken from the synthetic scene itself
    (noiseless FP64 depth through the pixel centres, analytic plane / sphere normals oriented away
    from the camera like the raycast's -grad F), at the pyramid's intrinsics f / 2^L,
    c_L = (c + 1/2) / 2^L - 1/2."""
    from kfusion.synth import SyntheticSequence
    V, N = [], []
    for lvl in range(levels):
        s = 1.0 / (1 << lvl)
        w, h = seq_args["width"] >> lvl, seq_args["height"] >> lvl
        q = SyntheticSequence(seed=seq_args["seed"], n_frames=seq_args["n_frames"], L=seq_args["L"], width=w,
                              height=h, fx=525.0 * s, fy=525.0 * s, cx=320.0 * s - 0.5, cy=240.0 * s - 0.5,
                              noise_k=0.0)
        _, td = q.frames(f, 1, true_depth=True)
        z = td[0]
        P = q.poses[f].reshape(4, 4).T
        ii, jj = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
        ray = np.stack([(ii - q.cx) / q.fx, (jj - q.cy) / q.fy, np.ones_like(ii)], -1)
        pw = (ray * z[..., None]) @ P[:3, :3].T + P[:3, 3]
        a = 0.45 * q.L
        resid = [np.abs(pw[..., 1] - a), np.abs(pw[..., 0] - a), np.abs(pw[..., 0] + a), np.abs(pw[..., 2] - a)]
        resid += [np.abs(np.linalg.norm(pw - c[:3], axis=-1) - c[3]) for c in q.spheres]
        which = np.argmin(np.stack(resid, -1), -1)
        nw = np.zeros_like(pw)
        for k, axis in enumerate(([0, 1.0, 0], [1.0, 0, 0], [-1.0, 0, 0], [0, 0, 1.0])):
            nw[which == k] = axis
        for k, c in enumerate(q.spheres):
            m = which == 4 + k
            nw[m] = (pw[m] - c[:3]) / c[3]
        nw[np.sum(nw * (pw - P[:3, 3]), -1) < 0] *= -1
        flag = np.where(z > 0, -1.0, -2.0)
        V.append(np.concatenate([pw, flag[..., None]], -1).astype(np.float32))
        N.append(np.concatenate([nw, flag[..., None]], -1).astype(np.float32))
    return V, N


def _measured_vertex_pyramid(o, raw):
    d = o.convert_to_depth_meters(raw, 1e-3)
    dp = [o.apply_bilateral_filter(d)]
    for _ in range(2):
        dp.append(o.down_sample(dp[-1]))
    V, Ks = [], []
    for lvl in range(3):
        K, Ki = o.intrinsics(525.0, 525.0, 319.5, 239.5, lvl)
        V.append(o.create_vertex_map(dp[lvl], Ki))
        Ks.append(K)
    return V, Ks


@pytest.mark.parametrize("seed,L,n_frames,f", [(1, 3.0, 30, 0), (1, 3.0, 30, 7), (2, 4.0, 300, 0), (2, 4.0, 300, 5)])
def test_compute_pose_recovers_known_motion(oracle32, seed, L, n_frames, f):
    """icp.cpp:25-87 on a known rigid motion: the measured (noisy, filtered) pyramid of frame f + 1
    against exact world-space targets of frame f, starting from the true pose of frame f, must land
    on the true pose of frame f + 1 within 0.05 deg and 0.5 mm.  The start pose is 0.3-0.5 deg and
    6-9 mm away, so an ICP that returns its start pose, solves for the wrong sign, or transposes
    the update fails."""
    from kfusion.synth import SyntheticSequence
    seq = SyntheticSequence(seed=seed, n_frames=n_frames, L=L)
    TV, TN = _exact_world_maps(dict(seed=seed, n_frames=n_frames, L=L, width=640, height=480), f)
    V, Ks = _measured_vertex_pyramid(oracle32, seq.frames(f + 1, 1)[0])
    P0, P1 = seq.poses[f].reshape(4, 4).T, seq.poses[f + 1].reshape(4, 4).T
    Rin = np.ascontiguousarray(P0[:3, :3].T.reshape(-1), np.float32)
    r = oracle32.compute_pose(V, TV, TN, [10, 5, 4], Ks, 0.1, float(np.deg2rad(20.0)), 0.0, 0.0, Rin,
                              P0[:3, 3].astype(np.float32))
    assert r["status"] == 0 and r["n_iters"] == 19

    def angle_deg(Ra, Rb):
        return float(np.degrees(np.arccos(np.clip((np.trace(Ra.T @ Rb) - 1) / 2, -1, 1))))

    Re = r["R"].reshape(3, 3).T.astype(np.float64)
    start_rot, start_trans = angle_deg(P1[:3, :3], P0[:3, :3]), np.linalg.norm(P1[:3, 3] - P0[:3, 3])
    assert start_rot > 0.3 and start_trans > 5e-3  # the motion to recover
    assert angle_deg(P1[:3, :3], Re) <= 0.05
    assert np.linalg.norm(r["t"].astype(np.float64) - P1[:3, 3]) <= 5e-4


def test_rigid_inverse_closed_form(oracle32, oracle64):
    """sensor.h:30-34 (Sensor::setPose stores glm::inverse(pose)): for P = [R | t] the inverse is
    [R^T | -R^T t]; checked against a general 4x4 inverse and P * P^-1 = I, in both precisions."""
    rng = np.random.default_rng(17)
    for _ in range(100):
        a = rng.uniform(-np.pi, np.pi, 3)
        Rm = rot("z", a[2]) @ rot("y", a[1]) @ rot("x", a[0])
        t = rng.uniform(-3.0, 3.0, 3)
        P = np.eye(4)
        P[:3, :3], P[:3, 3] = Rm, t
        for o, tol in ((oracle32, 5e-6), (oracle64, 1e-13)):
            inv = o.rigid_inverse(Rm.T.reshape(-1), t).astype(np.float64).reshape(4, 4).T
            Pq = np.eye(4)
            Pq[:3, :3] = Rm.astype(o.dt)
            Pq[:3, 3] = t.astype(o.dt)
            np.testing.assert_allclose(inv, np.linalg.inv(Pq), atol=tol)
            np.testing.assert_allclose(Pq @ inv, np.eye(4), atol=tol)
