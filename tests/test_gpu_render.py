"""Headless visualisation (display.cu:64-105): the shading and normal images of
raycast maps, against a numpy restatement, and the pipeline's visualisation
volume rendered from the current pose."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import announce, small_sequence

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


def to_byte(v):
    s = np.asarray(v, np.float32) * np.float32(255.0)
    out = np.where(s > 0, np.floor(np.minimum(s, 255.0)), 0.0)
    return np.nan_to_num(out, nan=0.0).astype(np.uint32)


def unpack(img):
    return img & 0xFF, (img >> 8) & 0xFF, (img >> 16) & 0xFF, img >> 24


def test_shading_and_normal_image_closed_form(kf):
    w, h = 64, 48
    rng = np.random.default_rng(3)
    n = rng.normal(size=(h, w, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    nm = np.concatenate([n, np.full((h, w, 1), -1.0, np.float32)], axis=-1)
    nm[5, 7] = [0, 0, 0, -2]  # invalid normal: black shading
    sensor = kf.Sensor(60.0, 60.0, 31.5, 23.5)
    a = 0.3
    pose = np.eye(4, dtype=np.float32)
    pose[:3, :3] = [[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]]
    pose[:3, 3] = [0.1, -0.2, 0.3]
    sensor.set_pose(pose.T.reshape(-1))
    d_nm = kf.DeviceArray.from_host(nm)
    out = kf.DeviceArray((h, w), np.uint32)
    announce("render shading 64x48")
    kf.check(kf.lib().kf_render_shading(d_nm.ptr, w, h, C.byref(sensor.s), out.ptr, None, None), "shading")
    img = out.download()
    # reference: dir = normalize((R Kinv) (i+1/2, j+1/2, 1)), radiance = min(|n . dir|, 1)
    M = pose[:3, :3].astype(np.float64) @ np.linalg.inv(np.array([[60, 0, 31.5], [0, 60, 23.5], [0, 0, 1.0]]))
    jj, ii = np.mgrid[0:h, 0:w]
    d = np.stack([ii + 0.5, jj + 0.5, np.ones_like(ii, float)], -1) @ M.T
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rad = np.minimum(np.abs(np.sum(nm[..., :3] * d, -1)), 1.0)
    r, g, b, al = unpack(img)
    assert (al == 255).all() and (r == g).all() and (g == b).all()
    assert np.abs(r.astype(int) - to_byte(rad).astype(int)).max() <= 1
    assert r[5, 7] == 0
    announce("render normals 64x48")
    rinv = pose[:3, :3].copy()  # row-major R = column-major R^T (the inverse rotation)
    kf.check(kf.lib().kf_render_normals(d_nm.ptr, w, h, rinv.reshape(-1).ctypes.data_as(C.c_void_p), out.ptr,
                                        None, None), "normals")
    img = out.download()
    nr = nm[..., :3] @ pose[:3, :3]  # R^T n for each pixel (row vectors)
    norm = np.linalg.norm(nr, axis=-1, keepdims=True)
    nr = np.divide(nr, norm, out=np.zeros_like(nr), where=norm > 0)  # pixel (7, 5) holds a zero normal
    r, g, b, al = unpack(img)
    for ch, comp in ((r, 0), (g, 1), (b, 2)):
        ok = np.ones((h, w), bool)
        ok[5, 7] = False  # normalize(0) is NaN -> 0 on the GPU
        assert np.abs(ch[ok].astype(int) - to_byte(nr[..., comp])[ok].astype(int)).max() <= 1
    assert r[5, 7] == 0 and g[5, 7] == 0 and b[5, 7] == 0


def test_pipeline_render_visualisation_volume(kf):
    w, h = 160, 120
    seq = small_sequence(seed=41, n_frames=3, w=w, h=h, L=3.0)
    raws = seq.frames(0, 3)
    gp = kf.Pipeline(w, h, seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, 64, 3.0, 0.03, initial_pose=seq.pose(0),
                     icp=kf.IcpConfig((4, 3, 2), w, h), vis_volume=(96, 3.0, 0.02))
    dev = kf.DeviceArray.from_host(raws)
    for f in range(3):
        announce(f"visualisation pipeline frame {f}")
        gp.frame(dev.ptr.value + f * w * h * 2)
    assert gp.sync() == 0
    shade = gp.render(shading=True)
    normals = gp.render(shading=False)
    r, g, b, al = unpack(shade)
    assert (al == 255).all() and (r == g).all()
    assert (r > 0).mean() > 0.75  # the synthetic room fills most of the view (rims stay unobserved)
    nr, ng, nb, _ = unpack(normals)
    assert (nr + ng + nb > 0).mean() > 0.75
    # rendering from an explicit pose equal to the current one gives the same image
    again = gp.render(shading=True, pose=gp.pose())
    np.testing.assert_array_equal(again, shade)
    gp.close()
