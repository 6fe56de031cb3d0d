"""The C++ mirror header (include/lbx/polymat.hpp) compiles against the C ABI, and its
test program (tests/cpp/test_polymat.cpp, the reference's test-fft / test-matpoly-mult /
test-polynomial-matrix checks in C++) passes on the GPU."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cpp"))
import build_cpp_test  # noqa: E402


def test_cpp_mirror_compiles():
    assert os.path.exists(build_cpp_test.build())


@pytest.mark.gpu
def test_cpp_mirror_runs():
    exe = build_cpp_test.OUT
    if not os.path.exists(exe):
        build_cpp_test.build()
    print("[case] C++ mirror test program", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0 and "PASS" in r.stdout
