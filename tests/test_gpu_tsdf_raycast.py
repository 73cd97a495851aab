"""GPU parity of TSDF fusion and raycasting (project/src/tsdf.cu, raycast.cu)
against the oracle.  Fusion: F, W and the touched-voxel count bit-exact.
Raycast: vertex/normal maps bit-exact (same software trilinear, same order)."""
import os
import tempfile

import numpy as np
import pytest

from gpu_util import announce, rt, small_sequence

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


def fuse_case(kf, oracle32, seed, n, L, mue, w, h, frames):
    seq = small_sequence(seed=seed, n_frames=frames, w=w, h=h, L=L)
    raws = seq.frames()
    grid = kf.VoxelGrid(L, n, mue=mue)
    F, W = oracle32.new_volume(n)
    sensor = kf.Sensor(seq.fx, seq.fy, seq.cx, seq.cy)
    raw = kf.DepthMap(w, h)
    dm = kf.GridMap(w, h)
    K0, _ = oracle32.intrinsics(seq.fx, seq.fy, seq.cx, seq.cy, 0)
    t0 = grid.touched_total()
    touched = 0
    for f in range(frames):
        announce(f"fuse seed {seed} n {n} L {L} mue {mue} {w}x{h} frame {f}")
        raw.update(raws[f])
        kf.kernel.convert_to_depth_meters(raw, dm, 1e-3)
        sensor.set_pose(seq.pose(f))
        kf.kernel.fuse(dm, grid, sensor)
        d = oracle32.convert_to_depth_meters(raws[f], 1e-3)
        touched += oracle32.fuse(F, W, n, L, mue, d, sensor.get_inverse_pose(), sensor.get_intr(0))
    Fg, Wg = grid.download()
    return seq, grid, sensor, F, W, Fg, Wg, touched, grid.touched_total() - t0


@pytest.mark.parametrize("n,L,w,h", [(64, 3.0, 160, 120), (128, 4.0, 320, 240)])
def test_fuse_bit_exact(kf, oracle32, n, L, w, h):
    _, _, _, F, W, Fg, Wg, touched, gtouched = fuse_case(kf, oracle32, 7, n, L, 0.03, w, h, 3)
    assert touched > 0 and gtouched == touched
    np.testing.assert_array_equal(Wg.view(np.uint32), W.view(np.uint32))
    np.testing.assert_array_equal(Fg.view(np.uint32), F.view(np.uint32))


@pytest.mark.parametrize("L", [3.0, 4.0])  # 4 m: power-of-two width (texture coordinates by product)
def test_raycast_bit_exact(kf, oracle32, L):
    n, mue, w, h = 128, 0.03, 320, 240
    seq, grid, sensor, F, W, Fg, Wg, _, _ = fuse_case(kf, oracle32, 9, n, L, mue, w, h, 3)
    for lvl in range(3):
        lw, lh = w >> lvl, h >> lvl
        vm, nm = kf.GridMap(lw, lh, 4), kf.GridMap(lw, lh, 4)
        init = np.zeros((lh, lw, 4), np.float32)
        init[..., 3] = -2
        rv, rn = oracle32.raycast(F, n, L, mue, sensor.get_pose(), sensor.get_inverse_intr(lvl), lw, lh)
        assert (rv[..., 3] == -1).mean() > 0.8
        # level 0 is cast three times in a row: the repeats dispatch their tiles in
        # the order the previous launch recorded, which must not change the maps
        for rep in range(3 if lvl == 0 else 1):
            announce(f"raycast n {n} L {L} level {lvl} launch {rep}")
            vm.upload(init)
            nm.upload(init)
            kf.kernel.raycast(grid, sensor, vm, nm, lvl)
            gv, gn = vm.download(), nm.download()
            np.testing.assert_array_equal(gv.view(np.uint32), rv.view(np.uint32))
            np.testing.assert_array_equal(gn.view(np.uint32), rn.view(np.uint32))


def test_raycast_stale_pixels(kf, oracle32):
    """raycast.cu:61-73: rays that miss the cube or start beyond 6 m leave the pixel untouched."""
    n, L, mue = 64, 2.0, 0.05
    grid = kf.VoxelGrid(L, n, mue=mue)
    sensor = kf.Sensor(60.0, 60.0, 31.5, 23.5)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = [0.0, 0.0, -3.0]       # camera 2 m in front of the cube, looking at it
    sensor.set_pose(pose.T.reshape(-1))
    vm, nm = kf.GridMap(64, 48, 4), kf.GridMap(64, 48, 4)
    marker = np.full((48, 64, 4), 7.0, np.float32)
    vm.upload(marker)
    nm.upload(marker)
    announce("raycast stale-pixel rule, camera outside the cube")
    kf.kernel.raycast(grid, sensor, vm, nm, 0)
    F, _ = oracle32.new_volume(n)
    rv, rn = oracle32.raycast(F, n, L, mue, sensor.get_pose(), sensor.get_inverse_intr(0), 64, 48,
                              marker.copy(), marker.copy())
    gv = vm.download()
    np.testing.assert_array_equal(gv, rv)
    stale = (gv == 7.0).all(-1)
    assert stale.any() and (~stale).any()  # corners miss the cube, the centre hits it


def test_volume_roundtrip_and_save_load(kf, oracle32):
    n, L = 64, 3.0
    announce("volume upload/download/save/load n 64")
    grid = kf.VoxelGrid(L, n, mue=0.05)
    rng = np.random.default_rng(1)
    F = rng.uniform(-1, 1, n ** 3).astype(np.float32)
    W = rng.uniform(0, 50, n ** 3).astype(np.float32)
    grid.upload(F, W)
    F2, W2 = grid.download()
    np.testing.assert_array_equal(F2, F)
    np.testing.assert_array_equal(W2, W)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tsdf.bin")
        grid.save_voxel_grid(path)
        raw = np.fromfile(path, np.float32)
        assert raw.size == 2 * n ** 3  # voxel_grid.cpp:43-60: n^3 (f, w) pairs, x fastest, no header
        np.testing.assert_array_equal(raw[0::2], F)
        np.testing.assert_array_equal(raw[1::2], W)
        g2 = kf.VoxelGrid(L, n, mue=0.05)
        g2.load_voxel_grid(path)
        F3, W3 = g2.download()
        np.testing.assert_array_equal(F3, F)
        np.testing.assert_array_equal(W3, W)
        with open(path, "r+b") as fp:
            fp.truncate(100)
        with pytest.raises(kf.KfError) as e:
            g2.load_voxel_grid(path)
        assert e.value.status == kf.KF_ERR_IO
    kf.kernel.initialize_grid(grid)  # Voxel{} = (2, 0)
    F4, W4 = grid.download()
    assert (F4 == 2.0).all() and (W4 == 0.0).all()
