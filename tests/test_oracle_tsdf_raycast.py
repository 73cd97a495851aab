"""Pins the oracle's TSDF fusion and raycast (oracle/kfref.c, restating
project/src/tsdf.cu and raycast.cu) against closed forms and brute force."""
import numpy as np
import pytest

from oracle import kfref

I4 = np.eye(4, dtype=np.float32).T.reshape(-1)


def wall_case(o, n=64, W=2.0, mue=0.1, D=0.8, frames=1, fmt="f32"):
    K, Ki = o.intrinsics(60.0, 60.0, 31.5, 23.5, 0)
    d = np.full((48, 64), D, np.float32)
    if fmt == "f32":
        F, Wt = o.new_volume(n)
        touched = [o.fuse(F, Wt, n, W, mue, d, I4, K) for _ in range(frames)]
    else:
        F, Wt = o.new_volume_q(n)
        touched = [o.fuse_q(F, Wt, n, W, mue, d, I4, K) for _ in range(frames)]
    return F, Wt, touched, K, Ki


def numpy_fuse_reference(n, W, mue, D, K):
    """tsdf.cu:20-75 evaluated independently with numpy (FP64 geometry)."""
    res = W / (n - 1)
    c = np.arange(n) * res - W / 2
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    px = K[0] * x / z + K[6]
    py = K[4] * y / z + K[7]
    with np.errstate(divide="ignore", invalid="ignore"):
        inside = (z >= 0.4) & (z < 6.0) & (px >= 0) & (px < 64) & (py >= 0) & (py < 48)
    sdf = D - z
    touched = inside & (sdf >= -mue)
    f = np.minimum(1.0, sdf / mue)
    return touched.reshape(-1), f.reshape(-1)


def test_fuse_wall_closed_form(oracle32):
    n, W, mue, D = 64, 2.0, 0.1, 0.8
    F, Wt, touched, K, _ = wall_case(oracle32, n, W, mue, D)
    mask, f = numpy_fuse_reference(n, W, mue, D, K)
    # repeated-addition z differs from i*res by ~1e-7 m: only voxels exactly at a boundary could flip
    assert abs(int(mask.sum()) - touched[0]) <= 2
    both = mask & (Wt > 0)
    np.testing.assert_allclose(F[both], f[both], atol=2e-5)
    np.testing.assert_allclose(Wt[both], 1.0 / np.float32(D), rtol=1e-7)
    assert (F[Wt == 0] == 2.0).all()


def test_fuse_repeats(oracle32):
    """F is invariant under repeated identical frames; W = k/D after k frames."""
    n, D = 32, 0.8
    F1, W1, t1, _, _ = wall_case(oracle32, n=n, D=D, frames=1)
    F3, W3, t3, _, _ = wall_case(oracle32, n=n, D=D, frames=3)
    assert t1[0] == t3[0] == t3[2] > 0
    seen = W1 > 0
    np.testing.assert_allclose(F3[seen], F1[seen], atol=1e-6)
    np.testing.assert_allclose(W3[seen], 3.0 / np.float32(D), rtol=1e-6)


def test_fuse_precision_form(oracle32):
    """snorm16 F codes and uint8 counts: W = number of frames, F = round(f * 32767)."""
    n, D, mue = 32, 0.8, 0.1
    F, Wt, touched, K, _ = wall_case(oracle32, n=n, D=D, mue=mue, frames=3, fmt="q")
    Fr, Wr, _, _, _ = wall_case(oracle32, n=n, D=D, mue=mue, frames=1)
    seen = Wr > 0
    assert (Wt[seen] == 3).all() and (Wt[~seen] == 0).all()
    np.testing.assert_array_less(np.abs(F[seen] / 32767.0 - Fr[seen]), 2.0 / 32767.0)
    assert (F[~seen] == -32768).all()  # the unobserved code: initial f = 2


def test_precision_form_unobserved_reads_as_two(oracle32):
    """The unobserved code decodes to 2, so a raycast of the precision-form volume equals
    the raycast of the FP32 volume whose F is the decoded codes (observation rims included)."""
    n, W, mue, D = 48, 2.0, 0.1, 0.8
    Fq, Wq, _, K, Ki = wall_case(oracle32, n=n, W=W, mue=mue, D=D, frames=2, fmt="q")
    Fd = np.where(Fq == -32768, np.float32(2.0), Fq.astype(np.float32) * np.float32(1.0 / 32767.0))
    vq, nq = oracle32.raycast_q(Fq, n, W, mue, I4, Ki, 64, 48)
    vf, nf = oracle32.raycast(Fd.astype(np.float32), n, W, mue, I4, Ki, 64, 48)
    np.testing.assert_array_equal(vq, vf)
    np.testing.assert_array_equal(nq, nf)


def test_raycast_wall_quirk_one_small_step(oracle32):
    """A fronto-parallel wall at depth D: the hit lies one small step (0.1 mue) beyond
    the linear zero crossing — raycast.cu:114 interpolates from the negative
    sample's distance — and the normal is +z (pointing away from the camera)."""
    n, W, mue, D = 128, 2.0, 0.06, 0.8
    F, Wt, _, K, Ki = wall_case(oracle32, n=n, W=W, mue=mue, D=D)
    # texel-centre sampling reads voxel i at (i + 1/2) W / n - W/2 instead of i W / (n-1) - W/2
    res = W / (n - 1)
    z_sample = ((D + W / 2) / res + 0.5) / n * W - W / 2  # where the fused zero crossing is read
    v, nm = oracle32.raycast(F, n, W, mue, I4, Ki, 64, 48)
    ok = v[..., 3] == -1
    assert ok.mean() > 0.9
    expected = z_sample + 0.1 * mue
    np.testing.assert_allclose(np.median(v[..., 2][ok]), expected, atol=2e-4)
    # border pixels' gradient stencils reach unobserved voxels (F = 2): use the median
    np.testing.assert_allclose(np.median(nm[..., :3][ok], axis=0), [0, 0, 1], atol=1e-3)
    centre = nm[16:32, 20:44]
    np.testing.assert_allclose(centre[..., :3].reshape(-1, 3), np.broadcast_to([0, 0, 1], (16 * 24, 3)), atol=1e-4)


def test_raycast_sphere_normals(oracle32):
    """Analytic sphere TSDF: hits lie on the sphere (up to the one-step quirk) and the
    normals point into the sphere along the radius."""
    n, W, mue = 96, 2.0, 0.05
    res = W / (n - 1)
    c = np.arange(n) * res - W / 2
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    C, r = np.array([0.0, 0.0, 0.75]), 0.25  # front at 0.5 m, beyond the 0.4 m near limit
    # camera at the origin looking along +z: projective distance along z, truncated
    sdf = np.linalg.norm(np.stack([x - C[0], y - C[1], z - C[2]]), axis=0) - r
    F = np.clip(sdf / mue, -1, 1).astype(np.float32).reshape(-1)
    # account for the texel-centre shift by sampling through the oracle itself
    _, Ki = oracle32.intrinsics(60.0, 60.0, 31.5, 23.5, 0)
    v, nm = oracle32.raycast(F, n, W, mue, I4, Ki, 64, 48)
    ok = v[..., 3] == -1
    assert ok.sum() > 200
    p = v[..., :3][ok]
    dist = np.linalg.norm(p - C, axis=1)
    assert np.abs(dist - r).max() < 0.03
    radial = (p - C) / dist[:, None]
    cosang = np.sum(-radial * nm[..., :3][ok], axis=1)
    assert np.median(cosang) > 0.99


def test_raycast_stale_pixels(oracle32):
    """raycast.cu:61-73: misses and ranges beyond 6 m leave the output untouched."""
    n, W, mue = 32, 2.0, 0.1
    F, _ = oracle32.new_volume(n)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = [0, 0, -3.0]
    _, Ki = oracle32.intrinsics(60.0, 60.0, 31.5, 23.5, 0)
    marker = np.full((48, 64, 4), 7.0, np.float32)
    v, _ = oracle32.raycast(F, n, W, mue, pose.T.reshape(-1), Ki, 64, 48, marker.copy(), marker.copy())
    stale = (v == 7.0).all(-1)
    assert stale[0, 0] and not stale[24, 32]
    assert (v[~stale][:, 3] == -2).all()  # an empty volume has no crossing
    far = np.eye(4, dtype=np.float32)
    far[:3, 3] = [0, 0, -9.0]  # the cube starts 8 m away: near >= far = 6
    v2, _ = oracle32.raycast(F, n, W, mue, far.T.reshape(-1), Ki, 64, 48, marker.copy(), marker.copy())
    assert (v2 == 7.0).all()


def test_fp64_oracle_close_to_fp32(oracle32, oracle64):
    F32, W32, _, K, Ki = wall_case(oracle32, n=48, D=0.9)
    F64, W64, _, _, _ = wall_case(oracle64, n=48, D=0.9)
    seen = (W32 > 0) & (W64 > 0)
    assert seen.sum() > 1000
    np.testing.assert_allclose(F32[seen], F64[seen], atol=1e-5)


def test_raycast_footprint_counts(oracle32):
    """The raycast footprint (bench.py's algorithmic bytes of the raycast): over free space every
    ray takes only large steps, so |S| grows with the image; a fused wall adds the small steps and
    gradients near the surface; recording never changes the maps."""
    from gpu_util import small_sequence
    n, L, mue, w, h = 32, 2.0, 0.1, 16, 12
    K, Ki = oracle32.intrinsics(16.0, 16.0, 7.5, 5.5, 0)
    F, W = oracle32.new_volume(n)
    F[:] = 1.0
    pose = np.eye(4, dtype=np.float32).reshape(-1)
    d_free, s_free = oracle32.raycast_footprint(F, n, L, mue, pose, [Ki], w, h)
    assert s_free > 0 and 0 < d_free <= 8 * s_free
    d_small, s_small = oracle32.raycast_footprint(F, n, L, mue, pose, [Ki], 4, 3)
    assert d_small < d_free and s_small < s_free
    depth = np.full((h, w), 0.7, np.float32)
    F2, W2 = oracle32.new_volume(n)
    oracle32.fuse(F2, W2, n, L, mue, depth, pose, K)
    v0, n0 = oracle32.raycast(F2, n, L, mue, pose, Ki, w, h)
    d_wall, s_wall = oracle32.raycast_footprint(F2, n, L, mue, pose, [Ki], w, h)
    v1, n1 = oracle32.raycast(F2, n, L, mue, pose, Ki, w, h)
    assert (v0[..., 3] == -1).mean() > 0.5
    np.testing.assert_array_equal(v0, v1)
    np.testing.assert_array_equal(n0, n1)
    assert s_wall > 0 and d_wall <= 8 * s_wall
    codes = np.where(F2 > 1.0, -32768, np.rint(np.clip(F2, -1, 1) * 32767)).astype(np.int16)
    assert oracle32.raycast_footprint(codes, n, L, mue, pose, [Ki], w, h) == (d_wall, s_wall)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fuse_counts_closed_form(dtype):
    """The count-weight update (the precision form's rule on unquantised planes, the FP64 error
    reference): after walls at depths D1 then D2, an observed voxel holds the mean of the two
    truncated values and W = 2; after 130 frames W saturates at 128.  A wrong weight, a missing
    +1 or an uncapped count fails it."""
    o = kfref.Oracle(dtype)
    n, Wd, mue = 32, 2.0, 0.1
    K, _ = o.intrinsics(60.0, 60.0, 31.5, 23.5, 0)
    d1 = np.full((48, 64), 0.8, np.float32)
    d2 = np.full((48, 64), 0.85, np.float32)
    F, W = o.new_volume(n)
    o.fuse_counts(F, W, n, Wd, mue, d1, I4, K)
    seen1 = W > 0
    f1 = F.copy()
    assert (W[seen1] == 1).all() and (F[~seen1] == 2.0).all()
    _, fa = numpy_fuse_reference(n, Wd, mue, 0.8, K)
    np.testing.assert_allclose(f1[seen1], fa[seen1], atol=2e-5)
    o.fuse_counts(F, W, n, Wd, mue, d2, I4, K)
    _, fb = numpy_fuse_reference(n, Wd, mue, 0.85, K)
    both = seen1 & (W == 2)
    assert both.sum() > 500
    np.testing.assert_allclose(F[both], 0.5 * (fa[both] + fb[both]), atol=4e-5)
    for _ in range(128):
        o.fuse_counts(F, W, n, Wd, mue, d2, I4, K)
    assert W.max() == 128 and (W[both] == 128).all()
