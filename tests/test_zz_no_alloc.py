"""lbx_polymat_mul_u32_noalloc: the product with every device buffer taken
from one fixed device array (no allocation in the call), checked word for word
against lbx_polymat_mul_u32 on the same inputs, through the C ABI, on calls in
a row (each call rebuilds what it needs, so a smaller second call must not see
the first call's contents), and its error paths."""
import ctypes

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import Polmat
from linbox_amd.hip import DeviceBuffer, Stream, synchronize

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = lbx.Context(0)
    yield c
    c.close()


def pm(buf, rows, cols, size, layout=0):
    return Polmat(buf.data_ptr(), layout, rows, cols, size, size, 0)


def run_both(ctx, p, m, k, n, sa, sb, method, max_rowdeg=0, stream=None, seed=0):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, p, m * k * sa, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, p, k * n * sb, dtype=np.uint64).astype(np.uint32)
    deg = max_rowdeg if max_rowdeg else sa + sb - 2
    da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
    outs = []
    for fn in (lbx.lib().lbx_polymat_mul_u32, lbx.lib().lbx_polymat_mul_u32_noalloc):
        dc = DeviceBuffer.from_numpy(np.full(m * n * (deg + 1), 0xABCDEF, dtype=np.uint32))
        A, B, C = pm(da, m, k, sa), pm(db, k, n, sb), pm(dc, m, n, deg + 1)
        st = fn(ctx.handle, p, ctypes.byref(C), ctypes.byref(A), ctypes.byref(B), max_rowdeg, method, stream)
        assert st == 0, lbx.lib().lbx_strerror(st)
        assert C.size == deg + 1
        synchronize()
        outs.append(dc.download())
    return outs


CASES = [  # (p, m, k, n, sa, sb, method, max_rowdeg)
    (2013265921, 16, 16, 16, 512, 512, lbx.METHOD_INT8_MFMA, 0),     # split transforms, 1024 points
    (998244353, 9, 17, 33, 70, 31, lbx.METHOD_PLAIN_U64, 0),
    (469762049, 20, 30, 40, 500, 300, lbx.METHOD_AUTO, 0),
    (2013265921, 33, 65, 17, 40, 90, lbx.METHOD_INT8_MFMA, 0),      # whole transforms, INT8 pointwise
    (167772161, 5, 7, 3, 100, 37, lbx.METHOD_PLAIN_U64, 120),        # max_rowdeg wrap
    (2013265921, 6, 4, 5, 1, 1, lbx.METHOD_AUTO, 3),                 # constants, coefficients 1..3 zero
    (998244353, 2, 2, 2, 40001, 30000, lbx.METHOD_PLAIN_U64, 0),    # 2^17 points (column passes)
]


@pytest.mark.parametrize("case", CASES)
def test_noalloc_matches_product(ctx, case):
    p, m, k, n, sa, sb, method, mr = case
    say("noalloc vs mul", case)
    ref, got = run_both(ctx, p, m, k, n, sa, sb, method, mr, seed=sa * 7 + m)
    assert np.array_equal(ref, got)


def test_noalloc_two_calls_second_smaller(ctx):
    """A large call fills the fixed array with its transforms; a smaller call
    right after (same stream, no synchronisation in between) must give its own
    exact result."""
    s = Stream()
    say("noalloc big then small")
    big = run_both(ctx, 2013265921, 32, 32, 32, 1000, 1000, lbx.METHOD_INT8_MFMA, stream=s.value, seed=1)
    small = run_both(ctx, 2013265921, 3, 4, 5, 60, 50, lbx.METHOD_INT8_MFMA, stream=s.value, seed=2)
    assert np.array_equal(big[0], big[1])
    assert np.array_equal(small[0], small[1])
    # and back to back on the device: queue both no-alloc calls before reading either
    rng = np.random.default_rng(3)
    jobs = [(32, 32, 32, 1000, 1000), (3, 4, 5, 60, 50)]
    bufs, refs = [], []
    p = 2013265921
    for (m, k, n, sa, sb) in jobs:
        a = rng.integers(0, p, m * k * sa, dtype=np.uint64).astype(np.uint32)
        b = rng.integers(0, p, k * n * sb, dtype=np.uint64).astype(np.uint32)
        da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
        dr = DeviceBuffer(m * n * (sa + sb - 1) * 4)
        dn = DeviceBuffer(m * n * (sa + sb - 1) * 4)
        st = lbx.lib().lbx_polymat_mul_u32(ctx.handle, p, ctypes.byref(pm(dr, m, n, sa + sb - 1)),
                                           ctypes.byref(pm(da, m, k, sa)), ctypes.byref(pm(db, k, n, sb)), 0,
                                           lbx.METHOD_INT8_MFMA, s.value)
        assert st == 0
        s.synchronize()
        bufs.append((da, db, dn, m, k, n, sa, sb))
        refs.append(dr.download())
    say("noalloc queued back to back")
    for (da, db, dn, m, k, n, sa, sb) in bufs:
        st = lbx.lib().lbx_polymat_mul_u32_noalloc(ctx.handle, p, ctypes.byref(pm(dn, m, n, sa + sb - 1)),
                                                   ctypes.byref(pm(da, m, k, sa)), ctypes.byref(pm(db, k, n, sb)), 0,
                                                   lbx.METHOD_INT8_MFMA, s.value)
        assert st == 0
    s.synchronize()
    for i, t in enumerate(bufs):
        assert np.array_equal(t[2].download(), refs[i])


def test_noalloc_errors(ctx):
    p = 2013265921
    L = lbx.lib()
    # beyond the fixed array: 64x64x64 at 2^13 points needs 3 x 128 MiB of transforms
    da, db, dc = DeviceBuffer(64 * 64 * 4096 * 4), DeviceBuffer(64 * 64 * 4096 * 4), DeviceBuffer(64 * 64 * 8191 * 4)
    say("noalloc too large")
    st = L.lbx_polymat_mul_u32_noalloc(ctx.handle, p, ctypes.byref(pm(dc, 64, 64, 8191)),
                                       ctypes.byref(pm(da, 64, 64, 4096)), ctypes.byref(pm(db, 64, 64, 4096)), 0,
                                       lbx.METHOD_PLAIN_U64, None)
    assert st == 13  # LBX_ERR_WORKSPACE_TOO_SMALL
    assert b"fixed workspace" in L.lbx_strerror(st)
    # host operands are refused (they would need staging buffers)
    a = np.ones(4, dtype=np.uint32)
    A = Polmat(a.ctypes.data, 0, 2, 2, 1, 1, 0)
    C = Polmat(a.ctypes.data, 0, 2, 2, 1, 1, 0)
    assert L.lbx_polymat_mul_u32_noalloc(ctx.handle, p, ctypes.byref(C), ctypes.byref(A), ctypes.byref(A), 0, 0,
                                         None) == 7
    # a modulus that needs the few-primes route
    d1 = DeviceBuffer(4 * 4 * 100 * 4)
    d2 = DeviceBuffer(4 * 4 * 199 * 4)
    assert L.lbx_polymat_mul_u32_noalloc(ctx.handle, 1000000007, ctypes.byref(pm(d2, 4, 4, 199)),
                                         ctypes.byref(pm(d1, 4, 4, 100)), ctypes.byref(pm(d1, 4, 4, 100)), 0, 0,
                                         None) == 11
