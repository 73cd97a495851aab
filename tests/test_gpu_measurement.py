"""GPU parity of the measurement stages (project/src/measurement.cu) against the
oracle: bit-exact, since both evaluate the same IEEE operations in the same order."""
import numpy as np
import pytest

from gpu_util import announce, small_sequence

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


def test_convert_exhaustive_codes(kf, oracle32):
    announce("convert_to_depth_meters: all 65536 codes, 256x256, scale 1/5000 and 1/1000, mirror on/off")
    codes = np.arange(65536, dtype=np.uint16).reshape(256, 256)
    raw = kf.GridMap(256, 256, 1, np.uint16)
    raw.upload(codes)
    out = kf.GridMap(256, 256)
    for scale in (1.0 / 5000.0, 1.0 / 1000.0):
        for unmirror in (False, True):
            kf.kernel.convert_to_depth_meters(raw, out, scale, unmirror)
            ref = oracle32.convert_to_depth_meters(codes, scale, unmirror)
            got = out.download()
            np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_measurement_stages_bit_exact(kf, oracle32):
    seq = small_sequence(seed=3, n_frames=2, w=320, h=240)
    frames = seq.frames()
    K = [oracle32.intrinsics(seq.fx, seq.fy, seq.cx, seq.cy, l) for l in range(3)]
    for f in range(2):
        announce(f"bilateral/downsample/vertex/normal: seed 3 frame {f}, 320x240")
        raw_h = frames[f].copy()
        raw_h[::17, ::13] = 0          # holes take part in the filter (measurement.cu:49-63)
        raw_h[5, :40] = 65535          # +inf codes
        raw = kf.DepthMap(320, 240)
        raw.update(raw_h)
        dm = kf.GridMap(320, 240)
        kf.kernel.convert_to_depth_meters(raw, dm, 1e-3)
        ref_dm = oracle32.convert_to_depth_meters(raw_h, 1e-3)
        np.testing.assert_array_equal(dm.download().view(np.uint32), ref_dm.view(np.uint32))
        dpyr = kf.GridMapPyramid(320, 240, 3)
        vpyr = kf.GridMapPyramid(320, 240, 3, channels=4)
        npyr = kf.GridMapPyramid(320, 240, 3, channels=4)
        kf.kernel.apply_bilateral_filter(dm, dpyr[0])
        ref_d = [oracle32.apply_bilateral_filter(ref_dm)]
        for l in (1, 2):
            kf.kernel.down_sample(dpyr[l - 1], dpyr[l])
            ref_d.append(oracle32.down_sample(ref_d[-1]))
        for l in range(3):
            got = dpyr[l].download()
            np.testing.assert_array_equal(got.view(np.uint32), ref_d[l].view(np.uint32), err_msg=f"depth L{l}")
            kf.kernel.create_vertex_map(dpyr[l], vpyr[l], K[l][1])
            rv = oracle32.create_vertex_map(ref_d[l], K[l][1])
            np.testing.assert_array_equal(vpyr[l].download().view(np.uint32), rv.view(np.uint32), err_msg=f"vertex L{l}")
            kf.kernel.create_normal_map(vpyr[l], npyr[l])
            rn = oracle32.create_normal_map(rv)
            np.testing.assert_array_equal(npyr[l].download().view(np.uint32), rn.view(np.uint32), err_msg=f"normal L{l}")


def test_normal_map_kat(kf):
    """transformation_tests.cpp:41-59 with the validity flag corrected: (0,0,0),(2,0,0),(0,4,0) -> (0,0,1)."""
    import json, os
    kat = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))["normal_map_2x2"]
    announce("createNormalMap KAT 2x2")
    v = kf.GridMap(2, 2, 4)
    v.upload(np.asarray(kat["vertices"], np.float32).reshape(2, 2, 4))
    n = kf.GridMap(2, 2, 4)
    kf.kernel.create_normal_map(v, n)
    got = n.download()
    np.testing.assert_array_equal(got[0, 0, :3], kat["expected_normal_00"])
    assert got[0, 0, 3] == -1.0
    # (1,0), (0,1), (1,1) lack a right or lower neighbour -> invalid
    assert (got.reshape(4, 4)[1:, 3] == -2.0).all()


def test_preprocess_fused_matches_stages(kf, oracle32):
    import ctypes as C
    seq = small_sequence(seed=5, n_frames=1, w=640, h=480)
    raw_h = seq.frames()[0]
    announce("kf_preprocess 640x480 seed 5 vs oracle stages")
    raw = kf.DepthMap(640, 480)
    raw.update(raw_h)
    dm = kf.GridMap(640, 480)
    dp = [kf.GridMap(640 >> l, 480 >> l) for l in range(3)]
    vp = [kf.GridMap(640 >> l, 480 >> l, 4) for l in range(3)]
    npm = [kf.GridMap(640 >> l, 480 >> l, 4) for l in range(3)]
    Ks = [oracle32.intrinsics(525.0, 525.0, 319.5, 239.5, l) for l in range(3)]
    kinv = np.ascontiguousarray(np.stack([k[1] for k in Ks]))
    arr = C.c_void_p * 3
    st = kf.lib().kf_preprocess(raw.ptr, 640, 480, 1e-3, 0, 3, kinv.ctypes.data_as(C.c_void_p), dm.ptr,
                                arr(*[d.ptr.value for d in dp]), arr(*[v.ptr.value for v in vp]),
                                arr(*[n.ptr.value for n in npm]), None, None)
    kf.check(st, "kf_preprocess")
    kf.check(kf.lib().kf_device_synchronize())
    r_dm = oracle32.convert_to_depth_meters(raw_h, 1e-3)
    np.testing.assert_array_equal(dm.download().view(np.uint32), r_dm.view(np.uint32))
    d = oracle32.apply_bilateral_filter(r_dm)
    for l in range(3):
        if l:
            d = oracle32.down_sample(d)
        v = oracle32.create_vertex_map(d, Ks[l][1])
        n = oracle32.create_normal_map(v)
        np.testing.assert_array_equal(dp[l].download().view(np.uint32), d.view(np.uint32))
        np.testing.assert_array_equal(vp[l].download().view(np.uint32), v.view(np.uint32))
        np.testing.assert_array_equal(npm[l].download().view(np.uint32), n.view(np.uint32))


def _subnormal_weight_taps(d):
    """(pixel, tap) pairs whose weight exp(-sd/16 - 16 id) is subnormal (argument in [-104, -87.34])."""
    h, w = d.shape
    n = 0
    for x in range(-3, 4):
        for y in range(-3, 4):
            nb = d[np.clip(np.arange(h) + y, 0, h - 1)][:, np.clip(np.arange(w) + x, 0, w - 1)]
            a = -(x * x + y * y) / 16.0 - 16.0 * (d.astype(np.float64) - nb) ** 2
            n += int(((a >= -104.0) & (a < -87.34)).sum())
    return n


def test_bilateral_subnormal_weights_bit_exact(kf, oracle32):
    """Depth steps of 2.34-2.55 m put weights in the subnormal range (the exp's
    arguments in [-104, -87.34]), where the final scaling by 2^k rounds."""
    rng = np.random.default_rng(11)
    d = (0.6 + 0.01 * rng.random((96, 128))).astype(np.float32)
    step = rng.random((96, 128)) < 0.3
    d[step] += rng.uniform(2.34, 2.55, size=int(step.sum())).astype(np.float32)
    assert _subnormal_weight_taps(d) > 1000
    announce("bilateral with subnormal weights: 128x96 random depth steps of 2.34-2.55 m")
    a, b = kf.GridMap(128, 96), kf.GridMap(128, 96)
    a.upload(d)
    kf.kernel.apply_bilateral_filter(a, b)
    np.testing.assert_array_equal(b.download().view(np.uint32), oracle32.apply_bilateral_filter(d).view(np.uint32))
    # the fused preprocessing launch: codes 600 and 3000 mm (a 2.4 m step)
    import ctypes as C
    raw_h = np.where(rng.random((480, 640)) < 0.3, 3000, 600).astype(np.uint16)
    r_dm = oracle32.convert_to_depth_meters(raw_h, 1e-3)
    assert _subnormal_weight_taps(r_dm[:64, :64]) > 1000
    announce("kf_preprocess with subnormal weights: 640x480 codes 600/3000")
    raw = kf.DepthMap(640, 480)
    raw.update(raw_h)
    dm = kf.GridMap(640, 480)
    dp = [kf.GridMap(640 >> l, 480 >> l) for l in range(3)]
    vp = [kf.GridMap(640 >> l, 480 >> l, 4) for l in range(3)]
    npm = [kf.GridMap(640 >> l, 480 >> l, 4) for l in range(3)]
    Ks = [oracle32.intrinsics(525.0, 525.0, 319.5, 239.5, l) for l in range(3)]
    kinv = np.ascontiguousarray(np.stack([k[1] for k in Ks]))
    arr = C.c_void_p * 3
    st = kf.lib().kf_preprocess(raw.ptr, 640, 480, 1e-3, 0, 3, kinv.ctypes.data_as(C.c_void_p), dm.ptr,
                                arr(*[x.ptr.value for x in dp]), arr(*[x.ptr.value for x in vp]),
                                arr(*[x.ptr.value for x in npm]), None, None)
    kf.check(st, "kf_preprocess")
    kf.check(kf.lib().kf_device_synchronize())
    np.testing.assert_array_equal(dp[0].download().view(np.uint32),
                                  oracle32.apply_bilateral_filter(r_dm).view(np.uint32))


def test_dimension_errors(kf):
    announce("dimension checks (no kernel launch)")
    a, b = kf.GridMap(8, 8), kf.GridMap(8, 4)
    with pytest.raises(kf.KfError) as e:
        kf.kernel.apply_bilateral_filter(a, b)
    assert e.value.status == kf.KF_ERR_DIMS
    with pytest.raises(kf.KfError) as e:
        kf.kernel.down_sample(a, kf.GridMap(8, 8))
    assert e.value.status == kf.KF_ERR_DIMS
    with pytest.raises(kf.KfError):
        kf.GridMapPyramid(10, 8, 3)
