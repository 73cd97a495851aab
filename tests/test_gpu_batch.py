"""Batch ticks (kf_batch_*): every session of a batch computes exactly what the
single-session pipeline (kf_pipeline, itself checked against the oracle in
test_gpu_pipeline.py) computes for the same frames — poses, TSDF, vertex and
predicted maps bit for bit — whatever else shares its ticks, whether ticks run
as hipGraphs or direct launches, and whatever launch shapes the autotuner
picks."""
import numpy as np
import pytest

from gpu_util import announce, small_sequence

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kf():
    import kfusion
    kfusion.init()
    return kfusion


# uneven sessions: frame size, volume size and format, seed, first tick
SESSIONS = [
    dict(w=320, h=240, n=128, L=3.0, fmt=0, seed=101, start=0),
    dict(w=160, h=120, n=64, L=3.0, fmt=1, seed=102, start=0),
    dict(w=320, h=240, n=96, L=3.0, fmt=1, seed=103, start=1),
    dict(w=160, h=120, n=128, L=4.0, fmt=0, seed=104, start=2),
]
TICKS = 5
ITERS = (4, 3, 2)


def make_inputs(kf, spec):
    seq = small_sequence(seed=spec["seed"], n_frames=TICKS, w=spec["w"], h=spec["h"], L=spec["L"])
    raws = seq.frames(0, TICKS)
    return seq, raws, kf.DeviceArray.from_host(raws)


def session_cfg(kf, spec, seq):
    return kf.SessionConfig(spec["w"], spec["h"], seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, spec["n"], spec["L"], 0.03,
                            spec["fmt"], seq.pose(0))


def icp_cfg(kf, w=640, h=480, compensated=True):
    return kf.IcpConfig(ITERS, w, h, 0.1, float(np.deg2rad(20.0)), compensated=compensated)


def run_batch(kf, specs, inputs, frames_of, **kw):
    """Runs the sessions through one batch; frames_of(s, tick) -> frame index or None."""
    b = kf.Batch(len(specs), icp=icp_cfg(kf), **kw)
    ids = [b.add(session_cfg(kf, sp, inputs[k][0])) for k, sp in enumerate(specs)]
    for t in range(TICKS):
        act, ptrs = [], []
        for k, sp in enumerate(specs):
            f = frames_of(k, t)
            if f is None:
                continue
            dev = inputs[k][2]
            act.append(ids[k])
            ptrs.append(dev.ptr.value + f * sp["w"] * sp["h"] * 2)
        announce(f"batch tick {t} sessions {act} {kw}")
        b.tick(act, ptrs)
    b.sync()
    out = []
    for k in range(len(specs)):
        assert b.status(ids[k]) == 0
        maps = [b.map(ids[k], which, l) for which in (2, 4, 5) for l in range(3)] + [b.map(ids[k], 0)]
        out.append(dict(pose=b.pose(ids[k]), vol=b.volume_download(ids[k]), maps=maps))
    b.close()
    return out


def run_pipeline(kf, spec, inputs, frames):
    seq, raws, dev = inputs
    gp = kf.Pipeline(spec["w"], spec["h"], seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, spec["n"], spec["L"], 0.03,
                     voxel_format=spec["fmt"], initial_pose=seq.pose(0), icp=icp_cfg(kf, spec["w"], spec["h"]))
    for f in frames:
        gp.frame(dev.ptr.value + f * spec["w"] * spec["h"] * 2)
    assert gp.sync() == 0
    maps = [gp.map(which, l) for which in (2, 4, 5) for l in range(3)] + [gp.map(0)]
    out = dict(pose=gp.pose(), vol=gp.volume_download(), maps=maps)
    gp.close()
    return out


def assert_same(a, b, what):
    np.testing.assert_array_equal(a["pose"], b["pose"], err_msg=f"{what}: pose")
    np.testing.assert_array_equal(a["vol"][0], b["vol"][0], err_msg=f"{what}: F")
    np.testing.assert_array_equal(a["vol"][1], b["vol"][1], err_msg=f"{what}: W")
    for q, (x, y) in enumerate(zip(a["maps"], b["maps"])):
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32), err_msg=f"{what}: map {q}")


def staggered(k, t):
    f = t - SESSIONS[k]["start"]
    return f if f >= 0 else None


@pytest.fixture(scope="module")
def inputs(kf):
    return [make_inputs(kf, sp) for sp in SESSIONS]


@pytest.fixture(scope="module")
def mixed(kf, inputs):
    return run_batch(kf, SESSIONS, inputs, staggered, use_graph=True)


def test_batch_equals_pipeline(kf, inputs, mixed):
    for k, sp in enumerate(SESSIONS):
        announce(f"pipeline reference for session {k}")
        ref = run_pipeline(kf, sp, inputs[k], range(TICKS - sp["start"]))
        assert_same(mixed[k], ref, f"session {k}")


def test_alone_equals_mix(kf, inputs, mixed):
    for k in (0, 2):
        alone = run_batch(kf, [SESSIONS[k]], [inputs[k]], lambda _, t, k=k: staggered(k, t), use_graph=True)
        assert_same(alone[0], mixed[k], f"session {k} alone vs mixed")


def test_direct_and_tuned_launches_equal_graph(kf, inputs, mixed):
    direct = run_batch(kf, SESSIONS, inputs, staggered, use_graph=False,
                       shape=dict(raycast_blocks=1024, icp_blocks=64, fuse_blocks=512, preprocess_blocks=256,
                                  cull_blocks=64))
    for k in range(len(SESSIONS)):
        assert_same(direct[k], mixed[k], f"session {k} direct launches")
    tuned = run_batch(kf, SESSIONS, inputs, staggered, use_graph=True, autotune=True)
    for k in range(len(SESSIONS)):
        assert_same(tuned[k], mixed[k], f"session {k} autotuned")


def test_skipped_ticks(kf, inputs):
    """A session absent from some ticks advances only on the ticks that carry its frames."""
    pattern = {0: [0, 1, None, 2, 3], 1: [0, None, None, 1, 2]}
    specs = SESSIONS[:2]
    got = run_batch(kf, specs, inputs[:2], lambda k, t: pattern[k][t], use_graph=True)
    for k in range(2):
        frames = [f for f in pattern[k] if f is not None]
        ref = run_pipeline(kf, specs[k], inputs[k], frames)
        assert_same(got[k], ref, f"session {k} with skipped ticks")


def test_session_admission_and_errors(kf, inputs):
    b = kf.Batch(2, icp=icp_cfg(kf))
    seq = inputs[0][0]
    a = b.add(session_cfg(kf, SESSIONS[0], seq))
    c = b.add(session_cfg(kf, SESSIONS[1], inputs[1][0]))
    with pytest.raises(kf.KfError):
        b.add(session_cfg(kf, SESSIONS[0], seq))  # no free slot
    with pytest.raises(kf.KfError):
        b.tick([a, a], [inputs[0][2].ptr.value] * 2)  # a session twice in one tick
    b.remove(c)
    c2 = b.add(session_cfg(kf, SESSIONS[1], inputs[1][0]))
    assert c2 == c
    bad = session_cfg(kf, SESSIONS[0], seq)
    bad.frame_width = 322
    with pytest.raises(kf.KfError):
        kf.session_bytes(bad)
    assert kf.session_bytes(session_cfg(kf, SESSIONS[0], seq)) > 128 ** 3 * 8
    b.close()


def test_checkpoint_restore_resumes_bit_exact(kf, inputs, tmp_path):
    """Session checkpoint (raw volume + pose, voxel_grid.cpp:43-80): saving after 3 ticks and
    resuming in a new batch gives the uninterrupted run's poses, volumes and maps."""
    specs, ins = SESSIONS[:2], inputs[:2]
    whole = run_batch(kf, specs, ins, lambda k, t: t, use_graph=True)
    b = kf.Batch(2, icp=icp_cfg(kf))
    ids = [b.add(session_cfg(kf, sp, ins[k][0])) for k, sp in enumerate(specs)]
    for t in range(3):
        announce(f"checkpoint run tick {t}")
        b.tick(ids, [ins[k][2].ptr.value + t * sp["w"] * sp["h"] * 2 for k, sp in enumerate(specs)])
    for k in range(2):
        b.save_session(ids[k], tmp_path / f"s{k}.bin")
    b.close()
    b = kf.Batch(2, icp=icp_cfg(kf))
    ids = [b.add(session_cfg(kf, sp, ins[k][0])) for k, sp in enumerate(specs)]
    for k in range(2):
        b.restore_session(ids[k], tmp_path / f"s{k}.bin")
    for t in range(3, TICKS):
        announce(f"resumed run tick {t}")
        b.tick(ids, [ins[k][2].ptr.value + t * sp["w"] * sp["h"] * 2 for k, sp in enumerate(specs)])
    b.sync()
    for k in range(2):
        maps = [b.map(ids[k], which, l) for which in (2, 4, 5) for l in range(3)] + [b.map(ids[k], 0)]
        got = dict(pose=b.pose(ids[k]), vol=b.volume_download(ids[k]), maps=maps)
        assert_same(got, whole[k], f"session {k} resumed")
    b.close()
