"""The oracle's per-frame loop (kinect_fusion.cpp:44-183) against its own pieces:
re-fusing a run's frames at the poses the run committed, and raycasting before
the last frame at the previous pose, reproduces the run's volume, touched
counts and predicted maps bit for bit (the replay the GPU parity tests rely
on), in the FP32 form and in the precision form."""
import numpy as np
import pytest

from gpu_util import assert_maps_equal, decode_q, independent, replay, small_sequence


@pytest.mark.parametrize("snorm16", [False, True])
def test_replay_reproduces_the_oracle_pipeline(oracle32, snorm16):
    n, mue, frames = 64, 0.03, 4
    seq = small_sequence(seed=41, n_frames=frames, w=160, h=120, L=3.0)
    raws = seq.frames()
    op, poses = independent(seq, raws, n, mue, snorm16=snorm16)
    touched_last = op.touched()
    if snorm16:
        fq, wq = op.tsdf_q()
        F, W = decode_q(fq), wq.astype(np.float32)
    else:
        F, W = op.tsdf()
    maps = [(op.map(1, l), op.map(2, l)) for l in range(3)]
    op.close()
    Fr, Wr, touched, rmaps = replay(oracle32, seq, n, mue, raws, poses, snorm16)
    assert touched[-1] == touched_last and min(touched) > 0
    np.testing.assert_array_equal(W, Wr)
    np.testing.assert_array_equal(F.view(np.uint32), Fr.view(np.uint32))
    assert_maps_equal(maps, rmaps, "oracle pipeline")
