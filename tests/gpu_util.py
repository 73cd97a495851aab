"""Helpers shared by the GPU parity tests (the checker side: oracle + synthetic input)."""
import sys

import numpy as np


def announce(msg):
    """Names the case before a GPU step, so a run that dies still says where."""
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def small_sequence(seed=1, n_frames=4, w=160, h=120, L=3.0, noise_k=0.0012):
    """A BASELINE-style seeded scene at reduced frame size (intrinsics scaled with the frame)."""
    from kfusion.synth import SyntheticSequence
    s = w / 640.0
    return SyntheticSequence(seed=seed, n_frames=n_frames, L=L, width=w, height=h, fx=525.0 * s, fy=525.0 * s,
                             cx=w / 2 - 0.5, cy=h / 2 - 0.5, noise_k=noise_k)


def rt(pose16):
    p = np.asarray(pose16, np.float32)
    return p[[0, 1, 2, 4, 5, 6, 8, 9, 10]].copy(), p[12:15].copy()


def rot_err(Ra, Rb):
    """Angle (rad) between two column-major rotations, atan2 of the axis and cosine
    parts of A^T B (arccos of the trace alone is ill-conditioned near identity)."""
    A = np.asarray(Ra, np.float64).reshape(3, 3).T
    B = np.asarray(Rb, np.float64).reshape(3, 3).T
    M = A.T @ B
    w = 0.5 * np.array([M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]])
    c = (np.trace(M) - 1.0) / 2.0
    return float(np.arctan2(np.linalg.norm(w), c))
