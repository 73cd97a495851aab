"""TUM RGB-D loader (rgbd_dataset.cpp:6-97) and poseError (general_helper.cpp:11-21)
on the host: a synthetic sequence directory (PNG encoder below, every PNG row
filter type) read back through the C ABI.  No GPU is used."""
import ctypes as C
import struct
import zlib

import numpy as np
import pytest


def png16(path, img, filters):
    """Writes a 16-bit grayscale PNG; row y uses filter type filters[y % len(filters)]."""
    h, w = img.shape
    be = img.astype(">u2").view(np.uint8).reshape(h, 2 * w).astype(np.int32)
    raw = bytearray()
    prev = np.zeros(2 * w, np.int32)
    for y in range(h):
        ft = filters[y % len(filters)]
        cur = be[y]
        a = np.concatenate([[0, 0], cur[:-2]])
        c = np.concatenate([[0, 0], prev[:-2]])
        if ft == 0:
            pred = np.zeros_like(cur)
        elif ft == 1:
            pred = a
        elif ft == 2:
            pred = prev
        elif ft == 3:
            pred = (a + prev) // 2
        else:
            p = a + prev - c
            pa, pb, pc = np.abs(p - a), np.abs(p - prev), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, c))
        raw.append(ft)
        raw += ((cur - pred) & 0xFF).astype(np.uint8).tobytes()
        prev = cur

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 16, 0, 0, 0, 0))
    data += chunk(b"IDAT", zlib.compress(bytes(raw), 6)) + chunk(b"IEND", b"")
    open(path, "wb").write(data)


def quat_to_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def reference_association(depth_ts, gt_ts):
    """rgbd_dataset.cpp:46-64 restated: forward search, start moved only at a break."""
    last, out = 0, []
    for t in depth_ts:
        best, bi = np.inf, -1
        for i in range(last, len(gt_ts)):
            d = abs(t - gt_ts[i])
            if d < best:
                best, bi = d, i
            else:
                last = bi
                break
        out.append(bi)
    return out


@pytest.fixture(scope="module")
def lib():
    from kfusion import _capi
    return _capi.load()


@pytest.fixture(scope="module")
def seq_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("tum")
    (d / "depth").mkdir()
    rng = np.random.default_rng(5)
    w, h, n = 40, 30, 6
    imgs = [rng.integers(0, 65536, size=(h, w), dtype=np.uint16) for _ in range(n)]
    imgs[0][:4] = 0  # runs of zeros, as missing depth
    depth_ts = [1305031102.175304 + 0.0333 * k for k in range(n)]
    with open(d / "depth.txt", "w") as f:
        f.write("# depth maps\n# file: 'synthetic'\n# timestamp filename\n")
        for k in range(n):
            name = f"depth/{depth_ts[k]:.6f}.png"
            png16(d / name, imgs[k], filters=[k % 5, (k + 1) % 5, (k + 2) % 5, 4, 3])
            f.write(f"{depth_ts[k]:.6f} {name}\n")
    gt_ts = [depth_ts[0] - 0.05 + 0.011 * k for k in range(40)]
    gt = []
    for k, t in enumerate(gt_ts):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        tr = rng.normal(size=3)
        gt.append((t, tr, q))
    with open(d / "groundtruth.txt", "w") as f:
        f.write("# ground truth trajectory\n# file: 'synthetic'\n# timestamp tx ty tz qx qy qz qw\n")
        for t, tr, q in gt:
            f.write(f"{t:.4f} {tr[0]:.4f} {tr[1]:.4f} {tr[2]:.4f} {q[0]:.4f} {q[1]:.4f} {q[2]:.4f} {q[3]:.4f}\n")
    # what the loader sees after the text round trip
    gt_read = [(float(f"{t:.4f}"), np.array([float(f"{v:.4f}") for v in tr]),
                np.array([float(f"{v:.4f}") for v in q])) for t, tr, q in gt]
    return d, imgs, depth_ts, gt_read


def test_png_and_association(lib, seq_dir):
    d, imgs, depth_ts, gt = seq_dir
    h = C.c_void_p()
    assert lib.kf_tum_open(str(d).encode(), C.byref(h)) == 0, lib.kf_last_error()
    n, w, hh = C.c_int(), C.c_int(), C.c_int()
    lib.kf_tum_size(h, C.byref(n), C.byref(w), C.byref(hh))
    assert (n.value, w.value, hh.value) == (len(imgs), 40, 30)
    assoc = reference_association(depth_ts, [g[0] for g in gt])
    for k in range(n.value):
        depth = np.zeros((30, 40), np.uint16)
        pose = np.zeros(16, np.float32)
        assert lib.kf_tum_frame(h, k, 0, depth.ctypes.data_as(C.c_void_p), pose.ctypes.data_as(C.c_void_p)) == 0
        np.testing.assert_array_equal(depth, imgs[k])  # every PNG filter type decoded exactly
        _, tr, q = gt[assoc[k]]
        P = pose.reshape(4, 4).T  # column-major -> matrix
        np.testing.assert_allclose(P[:3, :3], quat_to_mat(q), atol=2e-6)
        np.testing.assert_allclose(P[:3, 3], tr, atol=1e-6)
        np.testing.assert_array_equal(P[3], [0, 0, 0, 1])
        # the reference's off-by-one: the pose of the next frame
        p1 = np.zeros(16, np.float32)
        lib.kf_tum_frame(h, k, 1, None, p1.ctypes.data_as(C.c_void_p))
        nxt = min(k + 1, n.value - 1)
        np.testing.assert_allclose(p1.reshape(4, 4).T[:3, 3], gt[assoc[nxt]][1], atol=1e-6)
    assert lib.kf_tum_frame(h, n.value, 0, None, None) != 0
    lib.kf_tum_close(h)


def test_missing_and_bad_files(lib, tmp_path):
    h = C.c_void_p()
    assert lib.kf_tum_open(str(tmp_path).encode(), C.byref(h)) != 0
    (tmp_path / "groundtruth.txt").write_text("#\n#\n#\n1.0 0 0 0 0 0 0 1\n")
    (tmp_path / "depth.txt").write_text("#\n#\n#\n1.0 d.png\n")
    (tmp_path / "d.png").write_bytes(b"not a png")
    assert lib.kf_tum_open(str(tmp_path).encode(), C.byref(h)) != 0
    assert b"PNG" in lib.kf_last_error()


def test_pose_error_closed_form(lib):
    """poseError: |R1 v - R2 v| with v = (1,1,1)/sqrt(3) and the translation distance."""
    a = 0.2
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    p1 = np.eye(4)
    p2 = np.eye(4)
    p2[:3, :3] = R
    p2[:3, 3] = [0.3, -0.4, 0.0]
    ang, dist = C.c_float(), C.c_float()
    f1 = p1.T.reshape(-1).astype(np.float32)
    f2 = p2.T.reshape(-1).astype(np.float32)
    lib.kf_pose_error(f1.ctypes.data_as(C.c_void_p), f2.ctypes.data_as(C.c_void_p), C.byref(ang), C.byref(dist))
    v = np.ones(3) / np.sqrt(3)
    assert abs(ang.value - np.linalg.norm(v - R @ v)) < 1e-6
    assert abs(dist.value - 0.5) < 1e-6
