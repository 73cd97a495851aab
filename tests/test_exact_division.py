"""The fuse kernels form diff / mue and the count update's quotient from an exact
reciprocal instead of a division (tsdf.hip, div_by_exact_reciprocal); the C
program tests/cpp/exact_div_check.c checks on the CPU that this gives the
correctly rounded quotient over the operand ranges the kernels meet."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_exact_reciprocal_division_matches_rounded_quotient():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "exact_div_check")
        subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-o", exe,
                               os.path.join(HERE, "cpp", "exact_div_check.c"), "-lm"])
        r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
        print(r.stdout)
        assert r.returncode == 0, r.stdout
        assert r.stdout.startswith("0 of ")
