"""End-to-end parity of the per-frame loop (project/src/kinect_fusion.cpp:44-183):
the GPU pipeline against the oracle pipeline on the same seeded frames.
poses <= 1e-4 rad / 1e-4 m, TSDF |dF| <= 1e-3 (units of
the truncation), raycast depth within 1 mm on >= 99.9 % of the pixels valid in
both."""
import numpy as np
import pytest

from gpu_util import announce, rot_err, rt, small_sequence

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


def run_pair(kf, seq, n, mue, frames, iters=(10, 5, 4), compensated=False):
    """Runs the GPU pipeline and an independent oracle pipeline on the same frames;
    returns both, the GPU's per-frame poses and the worst pose differences."""
    from oracle.kfref import OraclePipeline
    raws = seq.frames(0, frames)
    w, h = seq.width, seq.height
    th = float(np.deg2rad(20.0))
    gp = kf.Pipeline(w, h, seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, n, seq.L, mue, initial_pose=seq.pose(0),
                     icp=kf.IcpConfig(iters, w, h, 0.1, th, compensated=compensated))
    op = OraclePipeline(w, h, n, seq.L, mue, seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, iters, 0.1, th,
                        init_pose=seq.pose(0))
    dev = kf.DeviceArray.from_host(raws)
    worst = dict(rot=0.0, trans=0.0, per_frame=[])
    gposes = []
    for f in range(frames):
        announce(f"pipeline {w}x{h} n {n} seed {seq.seed} frame {f} compensated={compensated}")
        gp.frame(dev.ptr.value + f * w * h * 2)
        assert gp.sync() == 0
        assert op.frame(raws[f]) == 0
        pg, po = gp.pose(), op.pose()
        gposes.append(pg)
        Rg, tg = rt(pg)
        Ro, to = rt(po)
        er, et = rot_err(Rg, Ro), float(np.abs(tg - to).max())
        worst["rot"] = max(worst["rot"], er)
        worst["trans"] = max(worst["trans"], et)
        worst["per_frame"].append((round(er, 8), round(et, 8)))
    return gp, op, gposes, raws, worst


def replay_with_gpu_poses(kf, oracle32, seq, n, mue, raws, gposes, gp):
    """Fuses every frame on the oracle at the pose the GPU committed for it and
    raycasts the model at the last frame's previous pose: the GPU's TSDF and
    predicted maps must equal it bit for bit (fusion and raycast are exact
    restatements; only the poses could differ, and they are taken from the GPU)."""
    w, h = seq.width, seq.height
    F, W = oracle32.new_volume(n)
    K0, _ = oracle32.intrinsics(seq.fx, seq.fy, seq.cx, seq.cy, 0)
    for f in range(len(raws)):
        d = oracle32.convert_to_depth_meters(raws[f], 1e-3)
        R, t = rt(gposes[f])
        oracle32.fuse(F, W, n, seq.L, mue, d, oracle32.rigid_inverse(R, t), K0)
    Fg, Wg = gp.volume_download()
    np.testing.assert_array_equal(Wg.view(np.uint32), W.view(np.uint32))
    np.testing.assert_array_equal(Fg.view(np.uint32), F.view(np.uint32))
    # predicted maps of the last frame came from the TSDF after frame len-2 at pose len-2
    F2, W2 = oracle32.new_volume(n)
    for f in range(len(raws) - 1):
        d = oracle32.convert_to_depth_meters(raws[f], 1e-3)
        R, t = rt(gposes[f])
        oracle32.fuse(F2, W2, n, seq.L, mue, d, oracle32.rigid_inverse(R, t), K0)
    for lvl in range(3):
        _, Ki = oracle32.intrinsics(seq.fx, seq.fy, seq.cx, seq.cy, lvl)
        vo, no = oracle32.raycast(F2, n, seq.L, mue, gposes[-2], Ki, w >> lvl, h >> lvl)
        vg, ng = gp.map(4, lvl), gp.map(5, lvl)
        np.testing.assert_array_equal(vg.view(np.uint32), vo.view(np.uint32), err_msg=f"vertex L{lvl}")
        np.testing.assert_array_equal(ng.view(np.uint32), no.view(np.uint32), err_msg=f"normal L{lvl}")


def independent_stats(gp, op):
    """TSDF and raycast agreement of the two independent runs (reported)."""
    Fg, Wg = gp.volume_download()
    Fo, Wo = op.tsdf()
    seen = (Wg > 0) & (Wo > 0)
    dF = np.abs(Fg - Fo)[seen]
    out = dict(tsdf_within_1e3=float((dF <= 1e-3).mean()), tsdf_p999=float(np.percentile(dF, 99.9)),
               observed_mask_agree=float(((Wg > 0) == (Wo > 0)).mean()))
    for lvl in range(3):
        vg, vo = gp.map(4, lvl), op.map(1, lvl)
        both = (vg[..., 3] == -1) & (vo[..., 3] == -1)
        dz = np.linalg.norm(vg[..., :3] - vo[..., :3], axis=-1)[both]
        out[f"raycast_L{lvl}_within_1mm"] = float((dz <= 1e-3).mean())
    return out


def test_pipeline_small_compensated(kf, oracle32):
    """160x120 frames, 128^3: with compensated sums the GPU's ICP matches the FP64-summing oracle."""
    seq = small_sequence(seed=21, n_frames=8, w=160, h=120, L=3.0)
    gp, op, gposes, raws, worst = run_pair(kf, seq, 128, 0.03, 8, compensated=True)
    st = independent_stats(gp, op)
    print("small pipeline parity:", worst, st, flush=True)
    assert worst["rot"] <= 1e-4 and worst["trans"] <= 1e-4, worst
    for lvl in range(3):
        assert st[f"raycast_L{lvl}_within_1mm"] >= 0.999, st
    replay_with_gpu_poses(kf, oracle32, seq, 128, 0.03, raws, gposes, gp)


def test_pipeline_640x480_to_256(kf, oracle32):
    """Plain method, small: 30 frames of 640x480, 256^3 FP32 over a 3 m cube, mue 3 cm, seed 1,
    FP32 tree sums: poses within 1e-4 of the independent oracle run; TSDF and predicted
    maps bit-exact against the oracle replayed at the GPU's poses."""
    from kfusion.synth import SyntheticSequence
    seq = SyntheticSequence(seed=1, n_frames=30, L=3.0)
    gp, op, gposes, raws, worst = run_pair(kf, seq, 256, 0.03, 30)
    st = independent_stats(gp, op)
    print("640x480 -> 256^3 parity:", worst, st, flush=True)
    assert worst["rot"] <= 1e-4 and worst["trans"] <= 1e-4, worst
    for lvl in range(3):
        assert st[f"raycast_L{lvl}_within_1mm"] >= 0.999, st
    # independent runs: F differs only where a ~1e-6 pose difference moves a voxel's
    # nearest depth pixel; given identical poses the TSDF is bit-exact (replay below)
    assert st["tsdf_within_1e3"] >= 0.99, st
    replay_with_gpu_poses(kf, oracle32, seq, 256, 0.03, raws, gposes, gp)
