"""Serving loop (kf_server_*): continuous batching of Poisson-arriving frames over
resident sessions admitted against an HBM budget.  Each session's result is the
same served alone and served in a mix (bit for bit: the batch computes every
session independently), and equals the single-session pipeline on the frames
the session received."""
import numpy as np
import pytest

from gpu_util import announce, small_sequence

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


SPECS = [
    dict(w=320, h=240, n=128, fmt=1, seed=201),
    dict(w=160, h=120, n=64, fmt=1, seed=202),
    dict(w=320, h=240, n=96, fmt=0, seed=203),
]
BANK = 4      # frames rendered per session, replayed ping-pong
FRAMES = 7    # frames sent per session: bank indices 0 1 2 3 2 1 0


def ping_pong(n, count):
    period = 2 * (n - 1) if n > 1 else 1
    return [p if p < n else period - p for p in (f % period for f in range(count))]


@pytest.fixture(scope="module")
def sessions(kf):
    out = []
    for sp in SPECS:
        seq = small_sequence(seed=sp["seed"], n_frames=BANK, w=sp["w"], h=sp["h"], L=3.0)
        cfg = kf.SessionConfig(sp["w"], sp["h"], seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, sp["n"], 3.0, 0.03, sp["fmt"],
                               seq.pose(0))
        out.append((seq, seq.frames(0, BANK), cfg))
    return out


def icp(kf):
    return kf.IcpConfig((4, 3, 2), 640, 480, 0.1, float(np.deg2rad(20.0)), compensated=True)


def pipeline_result(kf, sp, seq, bank):
    raws = bank[ping_pong(BANK, FRAMES)]
    dev = kf.DeviceArray.from_host(raws)
    gp = kf.Pipeline(sp["w"], sp["h"], seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, sp["n"], 3.0, 0.03,
                     voxel_format=sp["fmt"], initial_pose=seq.pose(0),
                     icp=kf.IcpConfig((4, 3, 2), sp["w"], sp["h"], 0.1, float(np.deg2rad(20.0)), compensated=True))
    for f in range(FRAMES):
        gp.frame(dev.ptr.value + f * sp["w"] * sp["h"] * 2)
    assert gp.sync() == 0
    out = gp.pose(), gp.volume_download()
    gp.close()
    return out


def serve(kf, idx, sessions, budget=64 << 30, keep_open=True, opens=None):
    srv = kf.Server(len(idx), budget, icp=icp(kf))
    cfgs = [sessions[k][2] for k in idx]
    banks = [sessions[k][1] for k in idx]
    opens = opens if opens is not None else [0.01 * q for q in range(len(idx))]
    announce(f"serving sessions {idx} budget {budget} keep_open={keep_open}")
    tickets, stats = srv.run_load(cfgs, banks, [FRAMES] * len(idx), opens, rate_hz=30.0, jitter=0.1, seed=7,
                                  keep_open=keep_open)
    res = {}
    for q, k in enumerate(idx):
        res[k] = (srv.pose(tickets[q]), srv.volume_download(tickets[q]) if keep_open else None)
    lat = srv.latencies()
    srv.close()
    return res, stats, lat


def test_alone_equals_mix_equals_pipeline(kf, sessions):
    mix, stats, lat = serve(kf, [0, 1, 2], sessions)
    assert stats["frames"] == 3 * FRAMES and len(lat) == 3 * FRAMES
    assert stats["failed_frames"] == 0
    assert 0 < stats["latency_p50_ms"] <= stats["latency_p99_ms"] <= stats["latency_max_ms"]
    alone, _, _ = serve(kf, [0], sessions)
    np.testing.assert_array_equal(alone[0][0], mix[0][0])
    np.testing.assert_array_equal(alone[0][1][0], mix[0][1][0])
    np.testing.assert_array_equal(alone[0][1][1], mix[0][1][1])
    for k, sp in enumerate(SPECS):
        announce(f"pipeline reference for served session {k}")
        pose, (F, W) = pipeline_result(kf, sp, sessions[k][0], sessions[k][1])
        np.testing.assert_array_equal(mix[k][0], pose)
        np.testing.assert_array_equal(mix[k][1][0], F)
        np.testing.assert_array_equal(mix[k][1][1], W)


def test_admission_waits_for_budget(kf, sessions):
    """A budget that holds two of the three sessions: the third waits until one closes,
    then runs to the same pose as the pipeline."""
    # the server charges kf_session_bytes plus the session's two device frame buffers
    b = [kf.session_bytes(sessions[k][2]) + 2 * 2 * SPECS[k]["w"] * SPECS[k]["h"] for k in range(3)]
    budget = b[0] + b[2] + 1  # sessions 0 and 2 fit; 1 must wait for one of them to close
    assert budget < b[0] + b[1] + b[2]
    res, stats, lat = serve(kf, [0, 2, 1], sessions, budget=budget, keep_open=False, opens=[0.0, 0.0, 0.02])
    assert stats["queued_admissions"] >= 1 and stats["admitted"] == 3
    assert stats["peak_resident_bytes"] <= budget
    assert stats["frames"] == 3 * FRAMES
    for k in (0, 1, 2):
        pose, _ = pipeline_result(kf, SPECS[k], sessions[k][0], sessions[k][1])
        np.testing.assert_array_equal(res[k][0], pose)


def test_submit_errors(kf, sessions):
    srv = kf.Server(1, 64 << 30, icp=icp(kf), queue_depth=2)
    t = srv.open(sessions[1][2])
    assert srv.state(t) == 1
    now = srv.now()
    frame = sessions[1][1][0]
    assert srv.submit(t, frame, now) == 0
    assert srv.submit(t, frame, now) == 0
    assert srv.submit(t, frame, now) == kf.KF_ERR_BUDGET  # queue of 2 is full
    srv.drain()
    t2 = srv.open(sessions[1][2])  # no free slot: waits
    assert srv.state(t2) == 0
    srv.close_session(t)
    srv.drain()
    assert srv.state(t) == -1 and srv.state(t2) == 1
    with pytest.raises(kf.KfError):
        kf.Server(1, 1024, icp=icp(kf)).open(sessions[0][2])  # larger than the whole budget
    srv.close()
