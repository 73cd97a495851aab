"""Runs the reference's C++ tests restated against the C++ class surface
(tests/cpp/test_class_surface.cpp -> include/kfusion/kfusion.hpp -> C ABI)."""
import os
import subprocess

import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "kf_class_tests")


def test_cpp_class_surface():
    assert os.path.exists(BIN), "build with __graft_entry__.build() (tests/cpp/Makefile)"
    announce("C++ class surface tests (tests/cpp/test_class_surface.cpp)")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "0 failed" in r.stdout
