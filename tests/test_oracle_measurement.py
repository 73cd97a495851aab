"""Pins the oracle's measurement stages (oracle/kfref.c, restating
project/src/measurement.cu) against definitions and brute force."""
import numpy as np

from oracle import kfref


def test_half_conversion_exhaustive():
    """measurement.cu:19-21: uint16 code -> binary16 (RNE) -> float, all 65536 codes,
    against numpy's IEEE float16 conversion."""
    l = kfref.lib()
    codes = np.arange(65536, dtype=np.uint16)
    with np.errstate(over="ignore"):
        ref = codes.astype(np.float16).astype(np.float32)  # numeric conversion, round to nearest even
    got = np.array([l.kfo_u16_to_half_float(int(c)) for c in codes], np.float32)
    np.testing.assert_array_equal(got, ref)
    assert np.isinf(got[65520]) and got[65519] == 65504.0
    assert (got[:2049] == np.arange(2049)).all()


def test_convert_scale_and_mirror(oracle32):
    raw = np.arange(12, dtype=np.uint16).reshape(3, 4) * 1000
    d = oracle32.convert_to_depth_meters(raw, 1e-3)
    np.testing.assert_array_equal(d, raw.astype(np.float16).astype(np.float32) * np.float32(1e-3))
    m = oracle32.convert_to_depth_meters(raw, 1e-3, unmirror=True)
    np.testing.assert_array_equal(m, d[:, ::-1])


def test_expf_accuracy():
    """The project-written expf (Cody-Waite + Cephes polynomial) against FP64 exp."""
    l = kfref.lib()
    xs = np.concatenate([np.linspace(-103.0, 88.0, 20001), np.linspace(-2.0, 0.0, 4001)]).astype(np.float32)
    got = np.array([l.kfo_expf(float(x)) for x in xs], np.float64)
    ref = np.exp(xs.astype(np.float64))
    normal = ref > 1.2e-38
    rel = np.abs(got[normal] - ref[normal]) / ref[normal]
    assert rel.max() < 4e-7  # a few ulp of float
    sub = ~normal
    assert np.abs(got[sub] - ref[sub]).max() < 2e-45  # denormals: within one denormal step
    assert l.kfo_expf(-200.0) == 0.0 and np.isinf(l.kfo_expf(100.0)) and l.kfo_expf(0.0) == 1.0


def bilateral_fp64(d):
    """measurement.cu:33-66 in FP64 with numpy (x outer / y inner, clamp to edge)."""
    h, w = d.shape
    out = np.zeros_like(d, np.float64)
    for j in range(h):
        for i in range(w):
            c = d[j, i]
            num = den = 0.0
            for x in range(-3, 4):
                for y in range(-3, 4):
                    q = d[min(max(j + y, 0), h - 1), min(max(i + x, 0), w - 1)]
                    f = np.exp(-(x * x + y * y) / 16.0 - (c - q) ** 2 / 0.0625)
                    num += f * q
                    den += f
            out[j, i] = num / den
    return out


def test_bilateral_constant_and_brute_force(oracle32):
    const = np.full((12, 10), 1.25, np.float32)
    np.testing.assert_allclose(oracle32.apply_bilateral_filter(const), const, rtol=2e-7)
    rng = np.random.default_rng(2)
    d = rng.uniform(0.5, 3.0, (14, 11)).astype(np.float32)
    d[3, 4] = 0.0  # an invalid pixel takes part in the average
    d[7:, :] += 1.0
    got = oracle32.apply_bilateral_filter(d)
    ref = bilateral_fp64(d.astype(np.float64))
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)
    # the FP64 oracle agrees with the brute force more tightly
    np.testing.assert_allclose(kfref.Oracle(np.float64).apply_bilateral_filter(d), ref, rtol=1e-12)


def test_down_sample_branches(oracle32):
    """measurement.cu:68-102: average of the taps within 0.75 m of the top-left tap."""
    d = np.array([[1.0, 1.5, 2.0, 9.0],
                  [1.2, 1.9, 2.0, 2.0]], np.float32)
    out = oracle32.down_sample(d)
    assert out.shape == (1, 2)
    np.testing.assert_allclose(out[0, 0], (1.0 + 1.5 + 1.2) / 3, rtol=1e-7)  # 1.9 is 0.9 away
    np.testing.assert_allclose(out[0, 1], (2.0 + 2.0 + 2.0) / 3, rtol=1e-7)   # 9.0 rejected
    const = np.full((8, 6), 0.7, np.float32)
    np.testing.assert_array_equal(oracle32.down_sample(const), np.full((4, 3), 0.7, np.float32))


def test_vertex_map_closed_form(oracle32):
    K, Ki = oracle32.intrinsics(525.0, 525.0, 319.5, 239.5, 0)
    rng = np.random.default_rng(0)
    d = rng.uniform(0.3, 4.0, (48, 64)).astype(np.float32)
    d[0, 0] = 0.0
    d[1, 1] = 0.005   # below 0.01 m: invalid (device_helper.cuh:11-17)
    d[2, 2] = np.inf
    v = oracle32.create_vertex_map(d, Ki)
    jj, ii = np.mgrid[0:48, 0:64]
    x = (ii + 0.5 - 319.5) / 525.0 * d
    y = (jj + 0.5 - 239.5) / 525.0 * d
    ok = v[..., 3] == -1
    assert not ok[0, 0] and not ok[1, 1] and not ok[2, 2] and ok.sum() == 48 * 64 - 3
    np.testing.assert_allclose(v[..., 0][ok], x[ok], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(v[..., 1][ok], y[ok], rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(v[..., 2][ok], d[ok])
    assert (v[~ok][:, :3] == 0).all() and (v[~ok][:, 3] == -2).all()


def test_normal_map_plane(oracle32):
    """A fronto-parallel plane gives +z normals (no orientation flip); the last
    column and row are invalid (their right / lower neighbour is out of range)."""
    _, Ki = oracle32.intrinsics(60.0, 60.0, 31.5, 23.5, 0)
    v = oracle32.create_vertex_map(np.full((24, 32), 2.0, np.float32), Ki)
    n = oracle32.create_normal_map(v)
    assert (n[:-1, :-1, 3] == -1).all() and (n[-1, :, 3] == -2).all() and (n[:, -1, 3] == -2).all()
    np.testing.assert_allclose(n[:-1, :-1, :3], np.broadcast_to([0, 0, 1], (23, 31, 3)), atol=1e-6)
