"""Seeded synthetic depth sequences (libkfsynth.so, synth/kfsynth.c).

A sequence = a trajectory of camera-to-world poses (column-major 4x4), 8
spheres in a room of floor + 3 walls inside a cube of side L, and uint16
millimetre depth frames ray-cast in FP64 with depth-dependent noise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._capi import SYNTH_PATH

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"libkfsynth.so is not built ({SYNTH_PATH})")
        l = C.CDLL(SYNTH_PATH)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        l.kfs_trajectory.argtypes = [C.c_uint64, I, P]
        l.kfs_trajectory.restype = I
        l.kfs_spheres.argtypes = [C.c_uint64, D, P, I, I, P]
        l.kfs_spheres.restype = I
        l.kfs_render.argtypes = [D, P, I, P, I, I, I, I, D, D, D, D, C.c_uint64, D, D, P, P]
        l.kfs_render.restype = None
        _lib = l
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class SyntheticSequence:
    """Seeded synthetic scene (a room with spheres): camera trajectory and noisy depth frames."""

    def __init__(self, seed=1, n_frames=30, L=3.0, width=640, height=480, fx=525.0, fy=525.0, cx=None, cy=None,
                 noise_k=0.0012, dropout=0.0, n_spheres=8):
        l = _load()
        self.seed, self.n_frames, self.L = int(seed), int(n_frames), float(L)
        self.width, self.height = int(width), int(height)
        self.fx, self.fy = float(fx), float(fy)
        self.cx = float(width / 2 - 0.5 if cx is None else cx)
        self.cy = float(height / 2 - 0.5 if cy is None else cy)
        self.noise_k, self.dropout = float(noise_k), float(dropout)
        self.poses = np.zeros((self.n_frames, 16), np.float64)
        if l.kfs_trajectory(self.seed, self.n_frames, _p(self.poses)) != 0:
            raise RuntimeError("kfs_trajectory failed")
        self.spheres = np.zeros((n_spheres, 4), np.float64)
        rej = l.kfs_spheres(self.seed, self.L, _p(self.poses), self.n_frames, n_spheres, _p(self.spheres))
        if rej < 0:
            raise RuntimeError("kfs_spheres: could not place spheres")
        self.rejected_spheres = rej

    def frames(self, first=0, count=None, true_depth=False):
        """uint16 frames [first, first+count) as (count, h, w); optionally the noiseless depth."""
        count = self.n_frames - first if count is None else count
        assert 0 <= first and first + count <= self.n_frames
        out = np.zeros((count, self.height, self.width), np.uint16)
        td = np.zeros((count, self.height, self.width), np.float64) if true_depth else None
        _load().kfs_render(self.L, _p(self.spheres), self.spheres.shape[0], _p(self.poses), first, count,
                           self.width, self.height, self.fx, self.fy, self.cx, self.cy, self.seed, self.noise_k,
                           self.dropout, _p(out), _p(td) if td is not None else None)
        return (out, td) if true_depth else out

    def pose(self, f, dtype=np.float32):
        """Camera-to-world pose of frame f, glm column-major (16 values)."""
        return self.poses[f].astype(dtype)

    def describe(self):
        return dict(seed=self.seed, frames=self.n_frames, cube_m=self.L, width=self.width, height=self.height,
                    fx=self.fx, fy=self.fy, cx=self.cx, cy=self.cy, noise_k=self.noise_k, dropout=self.dropout,
                    spheres=self.spheres.round(4).tolist())
