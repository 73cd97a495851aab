"""Oracle parity at the sizes of the benchmark workloads (the per-frame
loop, project/src/kinect_fusion.cpp:44-183, with tsdf.cu:20-75, raycast.cu:44-145
and icp.cpp:25-87 inside it).

Each case runs the GPU path at the workload's own volume and frame sizes,
an independent oracle run on the same frames (poses within 1e-4 rad / 1e-4 m),
and an oracle replay at the poses the GPU committed: fusion (F and W, or the
snorm16 codes and uint8 counts), the touched-voxel counts and the predicted
vertex/normal maps must then be bit-exact, because fusion and raycasting are
exact restatements and only the poses could differ.

 - plain FP32, 640x480 -> 512^3 over 4 m, seed 2;
 - precision form (snorm16 + uint8 counts, Kahan FP32 ICP sums, FP64
   solve), 640x480 -> 256^3 over 3 m, 30 frames, seed 1, with its errors
   against the FP64 oracle of the same algorithm stated per result;
 - 16 batched sessions of uneven volume (128..512)^3 and frame size
   (320x240, 640x480), one hipGraph per tick, autotuned launch shapes;
 - served sessions with volumes 256^3..768^3 under a Poisson load.
"""
import numpy as np
import pytest

from gpu_util import ITERS, TH, announce, assert_maps_equal, independent, pose_diff, replay, rt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


def sequence(seed, n_frames, L, w=640, h=480):
    from kfusion.synth import SyntheticSequence
    s = w / 640.0
    return SyntheticSequence(seed=seed, n_frames=n_frames, L=L, width=w, height=h, fx=525.0 * s, fy=525.0 * s,
                             cx=w / 2 - 0.5, cy=h / 2 - 0.5)


def run_pipeline(kf, seq, raws, n, mue, fmt, compensated):
    """The per-frame GPU pipeline frame by frame: poses, touched counts, final volume, last maps."""
    w, h = seq.width, seq.height
    gp = kf.Pipeline(w, h, seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, n, seq.L, mue, voxel_format=fmt,
                     initial_pose=seq.pose(0), icp=kf.IcpConfig(ITERS, w, h, 0.1, TH, compensated=compensated))
    dev = kf.DeviceArray.from_host(raws)
    poses, touched = [], []
    last = gp.touched_total()
    for f in range(len(raws)):
        announce(f"pipeline {w}x{h} n {n} fmt {fmt} seed {seq.seed} frame {f}")
        gp.frame(dev.ptr.value + f * w * h * 2)
        assert gp.sync() == 0, f"ICP failed on frame {f}"
        poses.append(gp.pose())
        now = gp.touched_total()
        touched.append(now - last)
        last = now
    maps = [(gp.map(4, l), gp.map(5, l)) for l in range(3)]
    F, W = gp.volume_download()
    gp.close()
    dev.free()
    return poses, touched, F, W, maps


def test_plain_640x480_to_512_fp32(kf, oracle32):
    """The plain workload at its own size: 640x480 -> 512^3 FP32 f + FP32 w over a 4 m cube, mue 3 cm, seed 2
    (the 300-frame sequence), the first 5 frames.  At 512^3 the brick cull's margins (bounding
    radius, occlusion slack, pixel box) are tightest relative to a voxel."""
    n, L, mue, frames = 512, 4.0, 0.03, 5
    seq = sequence(2, 300, L)
    raws = seq.frames(0, frames)
    poses, touched, F, W, maps = run_pipeline(kf, seq, raws, n, mue, kf.KF_VOXEL_F32, compensated=False)
    announce("plain 512^3: independent oracle run")
    op, oposes = independent(seq, raws, n, mue)
    op.close()
    worst = [pose_diff(a, b) for a, b in zip(poses, oposes)]
    print("plain 512^3 per-frame pose differences (rad, m):", worst, flush=True)
    assert max(r for r, _ in worst) <= 1e-4 and max(t for _, t in worst) <= 1e-4
    announce("plain 512^3: oracle replay at the GPU's poses")
    Fo, Wo, otouched, omaps = replay(oracle32, seq, n, mue, raws, poses, snorm16=False)
    print("plain 512^3 touched voxels per frame:", touched, flush=True)
    assert touched == otouched and min(touched) > 1_000_000
    np.testing.assert_array_equal(W.view(np.uint32), Wo.view(np.uint32))
    np.testing.assert_array_equal(F.view(np.uint32), Fo.view(np.uint32))
    assert_maps_equal(maps, omaps, "plain 512^3")


def test_precision_640x480_to_256_vs_fp64(kf, oracle32):
    """The precision workload: the seed-1 frames (640x480, 256^3 over 3 m, 30 frames) in the precision
    form.  Bit-exact: snorm16 codes, uint8 counts, touched counts and predicted maps against the
    oracle's precision mode replayed at the GPU's poses.  Against the FP64 oracle of the same
    algorithm (count weights, unquantised F): every frame's pose within 1e-4 rad / 1e-4 m, raycast
    depth within 1 mm on >= 99.9 % of the pixels valid in both, and, at the GPU's poses, the TSDF
    within 1e-3 (units of mue) on >= 99.5 % of the observed voxels, as close as FP32 arithmetic
    without the snorm16 rounding gets."""
    n, L, mue, frames = 256, 3.0, 0.03, 30
    seq = sequence(1, 30, L)
    raws = seq.frames(0, frames)
    poses, touched, F, W, maps = run_pipeline(kf, seq, raws, n, mue, kf.KF_VOXEL_SNORM16_U8, compensated=True)
    announce("precision 256^3: oracle replay (precision mode) at the GPU's poses")
    Fo, Wo, otouched, omaps = replay(oracle32, seq, n, mue, raws, poses, snorm16=True)
    assert touched == otouched
    np.testing.assert_array_equal(W, Wo)
    np.testing.assert_array_equal(F.view(np.uint32), Fo.view(np.uint32))
    assert_maps_equal(maps, omaps, "precision 256^3")

    announce("precision 256^3: FP64 oracle of the precision algorithm")
    from oracle.kfref import Oracle
    op64, p64 = independent(seq, raws, n, mue, dtype=np.float64, count_weights=True)
    errs = [pose_diff(a, b) for a, b in zip(poses, p64)]
    rot, trans = [e[0] for e in errs], [e[1] for e in errs]
    res = {"max_rot_rad": max(rot), "max_trans_m": max(trans), "last_rot_rad": rot[-1], "last_trans_m": trans[-1]}
    F64, W64 = op64.tsdf()
    both = (W > 0) & (W64 > 0)
    dF = np.abs(F.astype(np.float64) - F64)[both]
    res["tsdf_rms_independent"] = float(np.sqrt(np.mean(dF ** 2)))
    res["tsdf_within_1e3_independent"] = float((dF <= 1e-3).mean())
    for lvl in range(3):
        vg, v64 = maps[lvl][0], op64.map(1, lvl)
        ok = (vg[..., 3] == -1) & (v64[..., 3] == -1)
        dz = np.linalg.norm(vg[..., :3].astype(np.float64) - v64[..., :3], axis=-1)[ok]
        res[f"raycast_L{lvl}_within_1mm"] = float((dz <= 1e-3).mean())
        res[f"raycast_L{lvl}_mean_mm"] = 1e3 * float(dz.mean())
    op64.close()
    # the TSDF's own error: the count-weight update replayed at the GPU's poses in FP64 and, to
    # separate the snorm16 rounding from the FP32 arithmetic, in FP32 on unquantised planes
    o64 = Oracle(np.float64)
    Fr, Wr = o64.new_volume(n)
    F32, W32 = oracle32.new_volume(n)
    K0, _ = o64.intrinsics(seq.fx, seq.fy, seq.cx, seq.cy, 0)
    K32, _ = oracle32.intrinsics(seq.fx, seq.fy, seq.cx, seq.cy, 0)
    for f in range(frames):
        R, t = rt(poses[f])
        o64.fuse_counts(Fr, Wr, n, L, mue, o64.convert_to_depth_meters(raws[f], 1e-3),
                        o64.rigid_inverse(R.astype(np.float64), t.astype(np.float64)), K0)
        oracle32.fuse_counts(F32, W32, n, L, mue, oracle32.convert_to_depth_meters(raws[f], 1e-3),
                             oracle32.rigid_inverse(R, t), K32)
    for name, (Fa, Wa) in (("same_poses", (F, W)), ("fp32_counts_same_poses", (F32, W32))):
        seen = (Wa > 0) | (Wr > 0)
        dFr = np.abs(Fa.astype(np.float64) - Fr)[seen]
        res[f"tsdf_rms_{name}"] = float(np.sqrt(np.mean(dFr ** 2)))
        res[f"tsdf_within_1e3_{name}"] = float((dFr <= 1e-3).mean())
        res[f"weights_equal_{name}"] = float((Wa.astype(np.float64) == Wr)[seen].mean())
    print("precision 256^3 vs FP64 oracle:", res, flush=True)
    assert res["max_rot_rad"] <= 1e-4 and res["max_trans_m"] <= 1e-4, res
    for lvl in range(3):
        assert res[f"raycast_L{lvl}_within_1mm"] >= 0.999, res
    # At identical poses about 0.2 % of the observed voxels differ from FP64 by more than 1e-3:
    # voxel z comes from repeated FP32 additions (tsdf.cu:27-31), which moves a voxel's projection
    # by up to ~5e-3 pixel, so voxels that close to a pixel edge read the neighbouring (noisy)
    # depth; FP32 arithmetic on unquantised planes shows the same.  The snorm16 rounding must add
    # nothing measurable to it.
    assert res["tsdf_within_1e3_same_poses"] >= 0.995, res
    assert res["tsdf_within_1e3_same_poses"] >= res["tsdf_within_1e3_fp32_counts_same_poses"] - 1e-4, res


def digest(a):
    import hashlib
    return hashlib.blake2b(np.ascontiguousarray(a).data, digest_size=20).hexdigest()


def uneven_sessions():
    """16 sessions: volumes from {128, 256, 384, 512}^3 and frames from {320x240, 640x480},
    drawn with seed 99; session s has scene seed 100 + s; 4 m cubes, mue 3 cm, snorm16."""
    rng = np.random.RandomState(99)
    sizes = rng.choice([128, 256, 384, 512], 16)
    res = rng.choice(2, 16)
    return [dict(n=int(sizes[s]), w=320 << int(res[s]), h=240 << int(res[s]), seed=100 + s) for s in range(16)]


def test_batch16_uneven_128_to_512(kf, oracle32):
    """16 uneven sessions in one batch (one persistent launch per stage, each tick one
    hipGraph launch, autotuned launch shapes), 5 ticks.  A 512^3 / 640x480 session and a small one
    are replayed on the oracle (poses against an independent oracle run, volume codes, counts and
    predicted maps bit-exact at the batch's poses); the large one also runs alone in a batch and
    must equal its result in the mix bit for bit."""
    specs = uneven_sessions()
    ticks, L, mue = 5, 4.0, 0.03
    seqs = [sequence(sp["seed"], ticks, L, sp["w"], sp["h"]) for sp in specs]
    raws = [q.frames(0, ticks) for q in seqs]
    devs = [kf.DeviceArray.from_host(r) for r in raws]

    def run(idx, per_tick_poses, keep):
        b = kf.Batch(len(idx), icp=kf.IcpConfig(ITERS, 640, 480, 0.1, TH, compensated=True), use_graph=True,
                     autotune=True)
        ids = [b.add(kf.SessionConfig(specs[k]["w"], specs[k]["h"], seqs[k].fx, seqs[k].fy, seqs[k].cx, seqs[k].cy,
                                      1e-3, specs[k]["n"], L, mue, kf.KF_VOXEL_SNORM16_U8, seqs[k].pose(0)))
               for k in idx]
        poses = {k: [] for k in idx}
        for t in range(ticks):
            announce(f"batch16 tick {t} sessions {idx}")
            b.tick(ids, [devs[k].ptr.value + t * specs[k]["w"] * specs[k]["h"] * 2 for k in idx])
            if per_tick_poses:
                b.sync()
                for q, k in enumerate(idx):
                    poses[k].append(b.pose(ids[q]))
        b.sync()
        out = {}
        for q, k in enumerate(idx):
            assert b.status(ids[q]) == 0
            if k in keep:
                out[k] = dict(pose=b.pose(ids[q]), vol=b.volume_download(ids[q]),
                              maps=[(b.map(ids[q], 4, l), b.map(ids[q], 5, l)) for l in range(3)], poses=poses[k])
        shape = b.shape()
        b.close()
        return out, shape

    big = next(k for k, sp in enumerate(specs) if sp["n"] == 512 and sp["w"] == 640)
    small = next(k for k, sp in enumerate(specs) if sp["w"] == 320)
    mix, (shape, tuned) = run(list(range(16)), per_tick_poses=True, keep={big, small})
    print("batch16 sessions:", [(sp["n"], sp["w"]) for sp in specs], "shape:", shape, "tuned:", tuned, flush=True)
    for k in (big, small):
        sp, seq = specs[k], seqs[k]
        announce(f"batch16: oracle checks of session {k} ({sp['n']}^3, {sp['w']}x{sp['h']})")
        op, oposes = independent(seq, raws[k], sp["n"], mue, snorm16=True)
        op.close()
        worst = [pose_diff(a, b) for a, b in zip(mix[k]["poses"], oposes)]
        assert max(r for r, _ in worst) <= 1e-4 and max(t for _, t in worst) <= 1e-4, (k, worst)
        Fo, Wo, _, omaps = replay(oracle32, seq, sp["n"], mue, raws[k], mix[k]["poses"], snorm16=True)
        np.testing.assert_array_equal(mix[k]["vol"][1], Wo)
        np.testing.assert_array_equal(mix[k]["vol"][0].view(np.uint32), Fo.view(np.uint32))
        assert_maps_equal(mix[k]["maps"], omaps, f"batch16 session {k}")
    alone, _ = run([big], per_tick_poses=False, keep={big})
    np.testing.assert_array_equal(alone[big]["pose"], mix[big]["pose"])
    np.testing.assert_array_equal(alone[big]["vol"][0], mix[big]["vol"][0])
    np.testing.assert_array_equal(alone[big]["vol"][1], mix[big]["vol"][1])
    assert_maps_equal(alone[big]["maps"], mix[big]["maps"], "batch16 alone vs mix")
    for d in devs:
        d.free()


def test_served_256_to_768(kf, oracle32):
    """Served sessions of 640x480 depth with volumes 256^3..768^3 (snorm16 + uint8, 4 m
    cubes), Poisson arrivals (30 Hz +-10 %, seed 7) under a 200 GB budget, every volume resident.
    The 256^3 session equals the per-frame pipeline on the frames it received (final pose and volume
    bit for bit), whose poses the oracle replays to the same volume; every other session's result is
    the same served alone as in the mix."""
    sizes = [256, 768, 512, 384, 640]
    L, mue, bank, frames = 4.0, 0.03, 3, 5
    seqs = [sequence(1000 + s, bank, L) for s in range(len(sizes))]
    banks = [q.frames(0, bank) for q in seqs]
    cfgs = [kf.SessionConfig(640, 480, 525.0, 525.0, 319.5, 239.5, 1e-3, sizes[s], L, mue, kf.KF_VOXEL_SNORM16_U8,
                             seqs[s].pose(0)) for s in range(len(sizes))]
    order = [0, 1, 2, 1, 0]  # bank ping-pong of 5 frames

    def serve(idx):
        srv = kf.Server(len(idx), 200 * 10 ** 9, icp=kf.IcpConfig(ITERS, 640, 480, 0.1, TH, compensated=True),
                        use_graph=True)
        announce(f"served 256..768 sessions {idx}")
        tickets, st = srv.run_load([cfgs[k] for k in idx], [banks[k] for k in idx], [frames] * len(idx),
                                   [0.005 * q for q in range(len(idx))], rate_hz=30.0, jitter=0.1, seed=7,
                                   keep_open=True)
        assert st["frames"] == frames * len(idx) and st["failed_frames"] == 0
        out = {}
        for q, k in enumerate(idx):
            F, W = srv.volume_download(tickets[q])
            out[k] = (srv.pose(tickets[q]), (F, W) if k == 0 else (digest(F), digest(W)))
        srv.close()
        return out, st

    mix, st = serve(list(range(len(sizes))))
    print("served 256..768 mix:", {k: st[k] for k in ("frames", "ticks", "mean_frames_per_tick", "latency_p50_ms",
                                               "latency_p99_ms", "peak_resident_bytes")}, flush=True)
    # the 256^3 session against the oracle, through the per-frame pipeline on the same frames
    seq, raws = seqs[0], banks[0][order]
    poses, _, F, W, _ = run_pipeline(kf, seq, raws, 256, mue, kf.KF_VOXEL_SNORM16_U8, compensated=True)
    np.testing.assert_array_equal(mix[0][0], poses[-1])
    np.testing.assert_array_equal(mix[0][1][0], F)
    np.testing.assert_array_equal(mix[0][1][1], W)
    announce("served 256..768: oracle checks of the 256^3 session")
    op, oposes = independent(seq, raws, 256, mue, snorm16=True)
    op.close()
    worst = [pose_diff(a, b) for a, b in zip(poses, oposes)]
    assert max(r for r, _ in worst) <= 1e-4 and max(t for _, t in worst) <= 1e-4, worst
    Fo, Wo, _, _ = replay(oracle32, seq, 256, mue, raws, poses, snorm16=True)
    np.testing.assert_array_equal(W, Wo)
    np.testing.assert_array_equal(F.view(np.uint32), Fo.view(np.uint32))
    for k in range(1, len(sizes)):
        alone, _ = serve([k])
        np.testing.assert_array_equal(alone[k][0], mix[k][0], err_msg=f"session {k} pose")
        assert alone[k][1] == mix[k][1], f"session {k}: volume differs alone and in the mix"
