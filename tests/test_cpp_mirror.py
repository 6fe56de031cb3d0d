"""The C++ drop-in boundary: include/lbx/linbox_shim.hpp (specialisations of
PolynomialMatrixFFTMulDomain, PolynomialMatrixNaiveMulDomain, OrderBasis and FFT
for Givaro::Modular<uint32_t,uint64_t>) compiled against
include/lbx/linbox_mirror.hpp, which has LinBox's template shapes, with the
test program tests/cpp/test_linbox_shim.cpp (LinBox-shaped test code: mul,
midproduct, midproductgen through PolynomialMatrixDomain in all three layouts
and through SubPolynomialMatrix views, M_Basis and PM_Basis through OrderBasis,
FFT_direct / FFT_inverse), every result checked against the oracle."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cpp"))
import build_cpp_test  # noqa: E402


def test_linbox_shim_compiles():
    assert os.path.exists(build_cpp_test.build())


def test_shim_requires_linbox_declarations(tmp_path):
    """Included without LinBox's headers (or the mirror) the shim stops with an error."""
    src = tmp_path / "bad.cpp"
    src.write_text('#include "lbx/linbox_shim.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(os.path.dirname(HERE), "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode != 0 and "polynomial-matrix-domain.h" in r.stderr


@pytest.mark.gpu
def test_linbox_shim_runs():
    exe = build_cpp_test.OUT
    if not os.path.exists(exe):
        build_cpp_test.build()
    print("[case] C++ LinBox-shaped test program through the shim", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout[-6000:], r.stderr[-3000:])
    assert r.returncode == 0 and "PASS" in r.stdout
