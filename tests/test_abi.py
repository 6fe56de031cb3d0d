"""CPU-side checks of the C ABI library and the host-side mirror (no GPU compute).

* liblbx.so loads and exports every function include/lbx/lbx.h declares.
* lbx_strerror maps every status to the reference's message.
* The PolynomialMatrix index formulas follow polynomial-matrix.h:131-137 (polfirst),
  :422-428 (matfirst), :660-683 (matrowfirst), and views follow :940-1036.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, LinboxError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lbx", "lbx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:lbx_status|const char\s*\*|int)\s*(lbx_\w+)\s*\(", text, re.M)))


def test_header_declarations_match_binding_list():
    assert declared_functions() == sorted(lbx.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = lbx.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = os.popen("nm -D --defined-only %s" % lbx.LIB_PATH).read()
    for name in declared_functions():
        assert re.search(r"\bT %s\b" % name, out), name


def test_strerror_messages():
    L = lbx.lib()
    assert L.lbx_strerror(1) == b"LinBox ERROR: bad FFT Prime"
    assert L.lbx_strerror(2) == b"FFT: k is larger that log2(p-1)"
    assert L.lbx_strerror(3) == b"FFT: k must be positive"
    assert L.lbx_strerror(4) == b"w is not a primitive k-root of unity"
    for s in range(0, 13):
        assert L.lbx_strerror(s) != b"unknown status"
    assert L.lbx_strerror(99) == b"unknown status"


def test_null_arguments_rejected_without_device():
    L = lbx.lib()
    assert L.lbx_init(None, 0) == 7
    assert L.lbx_destroy(None) == 7
    assert L.lbx_polymat_mul_u32(None, 7, None, None, None, 0, 0, None) == 7
    assert L.lbx_order_basis_pm_u32(None, 7, None, None, 4, None, None, 0, None) == 7
    assert L.lbx_order_basis_m_u32(None, 7, None, None, 4, None, None, None) == 7
    assert L.lbx_polymat_copy_u32(None, None, None, 0, 0, 0, None) == 7


def test_layout_index_formulas():
    rows, cols, size = 3, 4, 5
    for layout in (lbx.POLFIRST, lbx.MATFIRST, lbx.MATROWFIRST):
        M = PolynomialMatrix(rows, cols, size, layout)
        seen = set()
        for i in range(rows):
            for j in range(cols):
                for k in range(size):
                    idx = M.index(i * cols + j, k)
                    if layout == lbx.POLFIRST:
                        assert idx == (i * cols + j) * size + k
                    elif layout == lbx.MATFIRST:
                        assert idx == i * cols + j + k * rows * cols
                    else:
                        assert idx == i * (cols * size) + j + k * cols
                    seen.add(idx)
        assert seen == set(range(rows * cols * size))


def test_layout_roundtrip_and_views():
    rng = np.random.default_rng(1)
    arr = rng.integers(0, 97, (6, 9), dtype=np.uint32)
    for layout in ("polfirst", "matfirst", "matrowfirst"):
        M = PolynomialMatrix.from_polfirst(arr, 2, 3, layout)
        assert np.array_equal(M.to_polfirst(), arr)
        assert M.get(1, 2, 4) == arr[5, 4]
        V = M.at(2, 6)
        assert V.size() == 5 and V.shift == 2
        assert np.array_equal(V.to_polfirst(), arr[:, 2:7])
        W = V.at(1, 3)
        assert W.shift == 3 and np.array_equal(W.to_polfirst(), arr[:, 3:6])
        with pytest.raises(LinboxError):
            V.resize(6)
        V.resize(2)
        assert np.array_equal(V.to_polfirst(), arr[:, 2:4])


def test_resize_whole_matrix_keeps_prefix():
    arr = np.arange(12, dtype=np.uint32).reshape(4, 3)
    for layout in ("polfirst", "matfirst", "matrowfirst"):
        M = PolynomialMatrix.from_polfirst(arr, 2, 2, layout)
        M.resize(5)
        got = M.to_polfirst()
        assert np.array_equal(got[:, :3], arr) and not got[:, 3:].any()
        M.resize(2)
        assert np.array_equal(M.to_polfirst(), arr[:, :2])
