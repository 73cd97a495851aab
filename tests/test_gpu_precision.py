"""Precision form (snorm16 TSDF + uint8 observation counts, compensated ICP
sums, FP64 solve): GPU against the oracle's precision mode bit for bit (fusion
codes, weights, raycast), and errors against the FP64 oracle of the same
algorithm stated per result (the precision workload at reduced size)."""
import numpy as np
import pytest

from gpu_util import announce, rot_err, rt, small_sequence

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


def test_snorm16_fuse_and_raycast_bit_exact(kf, oracle32):
    n, L, mue, w, h = 128, 3.0, 0.03, 320, 240
    seq = small_sequence(seed=31, n_frames=4, w=w, h=h, L=L)
    raws = seq.frames()
    grid = kf.VoxelGrid(L, n, mue=mue, voxel_format=kf.KF_VOXEL_SNORM16_U8)
    Fq, Wq = oracle32.new_volume_q(n)
    sensor = kf.Sensor(seq.fx, seq.fy, seq.cx, seq.cy)
    raw = kf.DepthMap(w, h)
    dm = kf.GridMap(w, h)
    for f in range(4):
        announce(f"snorm16 fuse seed 31 n {n} frame {f}")
        raw.update(raws[f])
        kf.kernel.convert_to_depth_meters(raw, dm, 1e-3)
        sensor.set_pose(seq.pose(f))
        kf.kernel.fuse(dm, grid, sensor)
        d = oracle32.convert_to_depth_meters(raws[f], 1e-3)
        oracle32.fuse_q(Fq, Wq, n, L, mue, d, sensor.get_inverse_pose(), sensor.get_intr(0))
    Fg, Wg = grid.download()  # decoded floats: q / 32767 (2 for the unobserved code) and counts
    np.testing.assert_array_equal(Wg, Wq.astype(np.float32))
    np.testing.assert_array_equal(Fg, np.where(Fq == -32768, np.float32(2.0),
                                               Fq.astype(np.float32) * np.float32(1.0 / 32767.0)))
    assert Wq.max() == 4 and grid.nbytes() == 3 * n ** 3
    for lvl in range(3):
        announce(f"snorm16 raycast level {lvl}")
        lw, lh = w >> lvl, h >> lvl
        vm, nm = kf.GridMap(lw, lh, 4), kf.GridMap(lw, lh, 4)
        init = np.zeros((lh, lw, 4), np.float32)
        init[..., 3] = -2
        vm.upload(init)
        nm.upload(init)
        kf.kernel.raycast(grid, sensor, vm, nm, lvl)
        rv, rn = oracle32.raycast_q(Fq, n, L, mue, sensor.get_pose(), sensor.get_inverse_intr(lvl), lw, lh)
        np.testing.assert_array_equal(vm.download().view(np.uint32), rv.view(np.uint32))
        np.testing.assert_array_equal(nm.download().view(np.uint32), rn.view(np.uint32))


def test_weight_cap_128(kf, oracle32):
    """uint8 counts saturate at 128 (the precision form's weight cap)."""
    n, L, mue = 32, 2.0, 0.1
    grid = kf.VoxelGrid(L, n, mue=mue, voxel_format=kf.KF_VOXEL_SNORM16_U8)
    sensor = kf.Sensor(60.0, 60.0, 31.5, 23.5)
    d = np.full((48, 64), 0.8, np.float32)
    dm = kf.GridMap(64, 48)
    dm.upload(d)
    announce("snorm16 weight cap: 130 identical frames")
    for _ in range(130):
        kf.kernel.fuse(dm, grid, sensor)
    _, Wg = grid.download()
    assert Wg.max() == 128.0 and (Wg[Wg > 0] == 128.0).all()


def test_precision_pipeline_vs_fp64_oracle(kf):
    """The precision workload at reduced size (320x240, 128^3, 10 frames): the precision-form GPU
    pipeline against the FP64 oracle of the same algorithm (count weights on unquantised F), so the
    difference is the snorm16 rounding and the FP32 arithmetic only: poses within 1e-4 rad / 1e-4 m
    and raycast depth within 1 mm on >= 99.9 % of the pixels.  The FP64 oracle of the plain
    algorithm (1/d weights) is reported beside it: the two weight rules alone differ by ~1.8e-4 rad
    here.  (The full-size case is tests/test_gpu_workloads.py::test_precision_640x480_to_256_vs_fp64.)"""
    from oracle.kfref import OraclePipeline
    seq = small_sequence(seed=1, n_frames=10, w=320, h=240, L=3.0)
    raws = seq.frames()
    th = float(np.deg2rad(20.0))
    gp = kf.Pipeline(320, 240, seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, 128, 3.0, 0.03,
                     voxel_format=kf.KF_VOXEL_SNORM16_U8, initial_pose=seq.pose(0),
                     icp=kf.IcpConfig((10, 5, 4), 320, 240, 0.1, th, compensated=True))
    ref = {name: OraclePipeline(320, 240, 128, 3.0, 0.03, seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, (10, 5, 4), 0.1, th,
                                init_pose=seq.pose(0), dtype=np.float64, count_weights=(name == "counts"))
           for name in ("counts", "plain")}
    dev = kf.DeviceArray.from_host(raws)
    err = {name: ([], []) for name in ref}
    for f in range(10):
        announce(f"precision pipeline frame {f}")
        gp.frame(dev.ptr.value + f * 320 * 240 * 2)
        assert gp.sync() == 0
        Rg, tg = rt(gp.pose())
        for name, o in ref.items():
            assert o.frame(raws[f]) == 0
            p64 = o.pose()
            err[name][0].append(rot_err(Rg, p64[[0, 1, 2, 4, 5, 6, 8, 9, 10]]))
            err[name][1].append(float(np.abs(tg - p64[12:15]).max()))
    Fg, Wg = gp.volume_download()
    res = {}
    for name, o in ref.items():
        F64, W64 = o.tsdf()
        both = (Wg > 0) & (W64 > 0)
        vg, v64 = gp.map(4, 0), o.map(1, 0)
        ok = (vg[..., 3] == -1) & (v64[..., 3] == -1)
        dz = np.linalg.norm(vg[..., :3] - v64[..., :3], axis=-1)[ok]
        res[name] = dict(max_rot=max(err[name][0]), max_trans=max(err[name][1]),
                         tsdf_rms=float(np.sqrt(np.mean((Fg[both] - F64[both]) ** 2))),
                         raycast_mean_mm=1e3 * float(dz.mean()), raycast_within_1mm=float((dz <= 1e-3).mean()))
        o.close()
    print("precision form vs FP64 oracles (counts: same algorithm; plain: 1/d weights):", res, flush=True)
    c = res["counts"]
    assert c["max_rot"] <= 1e-4 and c["max_trans"] <= 1e-4, res
    assert c["raycast_within_1mm"] >= 0.999, res
