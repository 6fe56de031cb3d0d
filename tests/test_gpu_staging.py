"""GPU tests of the host-operand copies (csrc/staging.hip): matrices in host memory
larger than the direct-copy limit (4 MiB) go through the context's pinned buffers
(16 MiB each, four of them) and its copy threads.  Every result must equal the same
call on device-resident operands word for word: whole chunks, a ragged last chunk
(sizes that are multiples of neither 16 MiB nor 4 KiB), several rounds of the
four-buffer ring, a view whose storage is copied in and out around it, one copy
thread and sixteen, the transform entry points and the layout conversion."""
import os
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, PolynomialMatrixFFTMulDomain, hip

pytestmark = pytest.mark.gpu
P = 998244353


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = lbx.Context(0)
    yield c
    c.close()


def rand(rng, n):
    return rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32)


def device_product(ctx, m, k, n, a, sa, b, sb, method):
    da, db = hip.DeviceBuffer.from_numpy(a), hip.DeviceBuffer.from_numpy(b)
    sc = sa + sb - 1
    dc = hip.DeviceBuffer(m * n * sc * 4)
    C = PolynomialMatrix(m, n, sc, data=dc)
    PolynomialMatrixFFTMulDomain(P, ctx, method).mul(C, PolynomialMatrix(m, k, sa, data=da),
                                                     PolynomialMatrix(k, n, sb, data=db))
    hip.synchronize()
    out = dc.download(count=m * n * sc)
    for x in (da, db, dc):
        x.free()
    return out


def host_product(ctx, m, k, n, a, sa, b, sb, method):
    A = PolynomialMatrix(m, k, sa, data=a.copy())
    B = PolynomialMatrix(k, n, sb, data=b.copy())
    C = PolynomialMatrix(m, n, sa + sb - 1)
    PolynomialMatrixFFTMulDomain(P, ctx, method).mul(C, A, B)
    assert np.array_equal(A.data, a) and np.array_equal(B.data, b)  # operands untouched
    return C.data


@pytest.mark.parametrize("m,k,n,sa,sb", [
    (16, 16, 16, 8192, 8192),      # 8 MiB operands, 16 MiB product: one chunk each
    (3, 5, 7, 300001, 200001),     # ragged: 18,000,060-byte a, 42 MB product
    (32, 32, 32, 16384, 16384),    # 64 MiB operands, 128 MiB product: the ring turns over
])
def test_host_operands_equal_device(ctx, m, k, n, sa, sb):
    rng = np.random.default_rng(m * 1000 + sa)
    a, b = rand(rng, m * k * sa), rand(rng, k * n * sb)
    say("staging product", m, k, n, sa, sb, "bytes", a.nbytes, b.nbytes, m * n * (sa + sb - 1) * 4)
    ref = device_product(ctx, m, k, n, a, sa, b, sb, lbx.METHOD_AUTO)
    got = host_product(ctx, m, k, n, a, sa, b, sb, lbx.METHOD_AUTO)
    assert np.array_equal(got, ref)


def test_host_view_keeps_the_rest_of_its_storage(ctx):
    """c is a view inside a larger host matrix: its storage is copied in, written
    over the view's window only, and copied back."""
    m, k, n, sa, sb = 8, 8, 8, 20000, 20000
    rng = np.random.default_rng(7)
    a, b = rand(rng, m * k * sa), rand(rng, k * n * sb)
    ref = device_product(ctx, m, k, n, a, sa, b, sb, lbx.METHOD_AUTO).reshape(m * n, sa + sb - 1)
    big = 3 + (sa + sb - 1) + 5
    fill = rand(rng, m * n * big)
    S = PolynomialMatrix(m, n, big, data=fill.copy())
    V = S.at(3, 3 + sa + sb - 2)
    say("staging view", S.data.nbytes)
    PolynomialMatrixFFTMulDomain(P, ctx).mul(V, PolynomialMatrix(m, k, sa, data=a.copy()),
                                             PolynomialMatrix(k, n, sb, data=b.copy()))
    got = S.data.reshape(m * n, big)
    want = fill.reshape(m * n, big).copy()
    want[:, 3:3 + sa + sb - 1] = ref
    assert np.array_equal(got, want)


@pytest.mark.parametrize("threads", ["1", "16"])
def test_copy_thread_counts(threads):
    old = os.environ.get("LBX_COPY_THREADS")
    os.environ["LBX_COPY_THREADS"] = threads
    try:
        c = lbx.Context(0)
        try:
            m, k, n, sa, sb = 4, 6, 5, 150001, 120001
            rng = np.random.default_rng(11)
            a, b = rand(rng, m * k * sa), rand(rng, k * n * sb)
            say("staging threads", threads)
            ref = device_product(c, m, k, n, a, sa, b, sb, lbx.METHOD_PLAIN_U64)
            assert np.array_equal(host_product(c, m, k, n, a, sa, b, sb, lbx.METHOD_PLAIN_U64), ref)
            c.trim()  # frees the pinned buffers; the next call makes new ones
            assert np.array_equal(host_product(c, m, k, n, a, sa, b, sb, lbx.METHOD_PLAIN_U64), ref)
        finally:
            c.close()
    finally:
        if old is None:
            del os.environ["LBX_COPY_THREADS"]
        else:
            os.environ["LBX_COPY_THREADS"] = old


def test_ntt_host_data(ctx):
    L, count = 14, 200  # 12.5 MiB
    rng = np.random.default_rng(3)
    x = rand(rng, count << L)
    say("staging ntt", x.nbytes)
    dx = hip.DeviceBuffer.from_numpy(x)
    w = lbx.ntt_forward(P, L, dx, count=count, ctx=ctx)
    hip.synchronize()
    ref = dx.download(count=count << L)
    dx.free()
    y = x.copy()
    assert lbx.ntt_forward(P, L, y, count=count, ctx=ctx) == w
    assert np.array_equal(y, ref)
    wi = lbx.FFT(P, L, w, ctx=ctx).invroot()
    lbx.ntt_inverse(P, L, y, count=count, root=wi, ctx=ctx)
    assert np.array_equal(y, (x.astype(np.uint64) * (1 << L) % P).astype(np.uint32))


def test_layout_copy_host(ctx):
    m, n, s = 64, 64, 3000  # 47 MiB each way
    rng = np.random.default_rng(5)
    A = PolynomialMatrix.from_polfirst(rand(rng, m * n * s).reshape(m * n, s), m, n)
    B = PolynomialMatrix(m, n, s, layout=lbx.MATFIRST)
    say("staging layout copy", A.data.nbytes)
    B.copy(A, ctx=ctx)
    assert np.array_equal(B.to_polfirst(), A.to_polfirst())


def test_two_threads_two_contexts():
    """Two host threads, each with its own context, run host-operand products at
    the same time (ctypes releases the interpreter lock during the calls): each
    context's pinned buffers and copy threads serve only its own calls."""
    import threading
    m, k, n, sa, sb = 8, 8, 8, 40000, 30000  # 10 / 7.7 MiB operands, 17.9 MiB products
    rng = np.random.default_rng(21)
    cases = [(rand(rng, m * k * sa), rand(rng, k * n * sb)) for _ in range(2)]
    refs = []
    c0 = lbx.Context(0)
    try:
        for a, b in cases:
            refs.append(device_product(c0, m, k, n, a, sa, b, sb, lbx.METHOD_AUTO))
    finally:
        c0.close()
    say("staging two threads")
    results, errors = [None, None], []

    def run(i):
        try:
            c = lbx.Context(0)
            try:
                a, b = cases[i]
                outs = [host_product(c, m, k, n, a, sa, b, sb, lbx.METHOD_AUTO) for _ in range(3)]
                results[i] = outs
            finally:
                c.close()
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=200)
    assert not errors, errors
    for i in range(2):
        for out in results[i]:
            assert np.array_equal(out, refs[i])
