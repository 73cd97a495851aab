"""Helpers shared by the GPU parity tests (the checker side: oracle + synthetic input)."""
import sys

import numpy as np


def announce(msg):
    """Names the case before a GPU step, so a run that dies still says where."""
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def small_sequence(seed=1, n_frames=4, w=160, h=120, L=3.0, noise_k=0.0012):
    """A BASELINE-style seeded scene at reduced frame size (intrinsics scaled with the frame)."""
    from kfusion.synth import SyntheticSequence
    s = w / 640.0
    return SyntheticSequence(seed=seed, n_frames=n_frames, L=L, width=w, height=h, fx=525.0 * s, fy=525.0 * s,
                             cx=w / 2 - 0.5, cy=h / 2 - 0.5, noise_k=noise_k)


def rt(pose16):
    p = np.asarray(pose16, np.float32)
    return p[[0, 1, 2, 4, 5, 6, 8, 9, 10]].copy(), p[12:15].copy()


def rot_err(Ra, Rb):
    """Angle (rad) between two column-major rotations, atan2 of the axis and cosine
    parts of A^T B (arccos of the trace alone is ill-conditioned near identity)."""
    A = np.asarray(Ra, np.float64).reshape(3, 3).T
    B = np.asarray(Rb, np.float64).reshape(3, 3).T
    M = A.T @ B
    w = 0.5 * np.array([M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]])
    c = (np.trace(M) - 1.0) / 2.0
    return float(np.arctan2(np.linalg.norm(w), c))


TH = float(np.deg2rad(20.0))
ITERS = (10, 5, 4)


def pose_diff(pa, pb):
    Ra, ta = rt(pa)
    Rb, tb = rt(pb)
    return rot_err(Ra, Rb), float(np.abs(ta.astype(np.float64) - tb.astype(np.float64)).max())


def decode_q(codes):
    """snorm16 codes as the library decodes them (q / 32767; the spare code -32768 is 2)."""
    return np.where(codes == -32768, np.float32(2.0), codes.astype(np.float32) * np.float32(1.0 / 32767.0))


def replay(oracle32, seq, n, mue, raws, poses, snorm16):
    """Fuses frames 0..len-1 on the oracle at the given poses.  Returns the volume (F, W as the
    library's download decodes it), the touched count of every frame, and the predicted maps of
    the last frame (raycast at pose len-2 of the volume after frame len-2)."""
    K0, _ = oracle32.intrinsics(seq.fx, seq.fy, seq.cx, seq.cy, 0)
    F, W = oracle32.new_volume_q(n) if snorm16 else oracle32.new_volume(n)
    touched, maps = [], None
    for f in range(len(raws)):
        if f == len(raws) - 1:
            maps = []
            for lvl in range(3):
                _, Ki = oracle32.intrinsics(seq.fx, seq.fy, seq.cx, seq.cy, lvl)
                lw, lh = seq.width >> lvl, seq.height >> lvl
                cast = oracle32.raycast_q if snorm16 else oracle32.raycast
                maps.append(cast(F, n, seq.L, mue, poses[f - 1], Ki, lw, lh))
        d = oracle32.convert_to_depth_meters(raws[f], 1e-3)
        R, t = rt(poses[f])
        fuse = oracle32.fuse_q if snorm16 else oracle32.fuse
        touched.append(int(fuse(F, W, n, seq.L, mue, d, oracle32.rigid_inverse(R, t), K0)))
    if snorm16:
        F, W = decode_q(F), W.astype(np.float32)
    return F, W, touched, maps


def assert_maps_equal(got, want, what):
    for lvl, ((gv, gn), (ov, on)) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(gv.view(np.uint32), ov.view(np.uint32), err_msg=f"{what} vertex L{lvl}")
        np.testing.assert_array_equal(gn.view(np.uint32), on.view(np.uint32), err_msg=f"{what} normal L{lvl}")


def independent(seq, raws, n, mue, **kw):
    from oracle.kfref import OraclePipeline
    op = OraclePipeline(seq.width, seq.height, n, seq.L, mue, seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, ITERS, 0.1, TH,
                        init_pose=seq.pose(0), **kw)
    poses = []
    for f in range(len(raws)):
        assert op.frame(raws[f]) == 0
        poses.append(op.pose())
    return op, poses
