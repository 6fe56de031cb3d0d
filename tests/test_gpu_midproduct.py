"""GPU parity of the middle products (lbx_polymat_midproduct_u32, _naive_u32).

* The committed cyclic-correlation fixtures (Python big integers), including
  outputs past the middle product where the correlation wraps, and the
  one-point case that FFT rejects.
* Random shapes against the oracle's restatement of the FFT-prime midproduct,
  for both operand orders, both pointwise forms and transform lengths 2^2 to
  2^14 (the split transforms from 2^8 on).
* Primes without the needed 2-valuation and primes >= 2^31 through the CRT
  route, against the correlation computed with Python integers.
* The naive midproduct (c.size() <= 4 in midproductgen) against the oracle,
  including the coefficients it must leave untouched.
* Views and layouts, the PolynomialMatrixMulDomain dispatch, and error codes.
* A PM-Basis level at full size (midproductgen(serie2, sigma1, serie, true,
  ord1+1, ord1+ord2), order-basis.h:280): the window against the device
  product's middle coefficients, and rows against the oracle.
"""
import json
import os
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import (PolynomialMatrix, PolynomialMatrixMulDomain, PolynomialMatrixFFTMulDomain,
                        PolynomialMatrixNaiveMulDomain, LinboxError)
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PRIMES = [998244353, 2013265921, 469762049, 167772161]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = lbx.Context(0)
    yield c
    c.close()


def rand(rng, p, n):
    return rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32)


def gpu_mid(ctx, p, m, k, n, a, sa, b, sb, sc, small_left, n0=0, n1=0, method=lbx.METHOD_AUTO,
            layouts=("polfirst",) * 3):
    A = PolynomialMatrix.from_polfirst(np.asarray(a).reshape(m * k, sa), m, k, layouts[0])
    B = PolynomialMatrix.from_polfirst(np.asarray(b).reshape(k * n, sb), k, n, layouts[1])
    C = PolynomialMatrix(m, n, sc, layouts[2])
    PolynomialMatrixFFTMulDomain(p, ctx, method).midproduct(C, A, B, small_left, n0, n1)
    assert C.size() == sc
    return C.to_polfirst().reshape(-1)


def correlation(p, m, k, n, a, sa, b, sb, sc, small_left, n0, n1):
    """The cyclic correlation the FFT midproduct computes, with Python integers."""
    hdeg = n0 or sc
    deg = n1 or 2 * hdeg - 1
    pts = 1
    while pts < deg:
        pts <<= 1
    A = np.zeros((m * k, pts), dtype=object)
    B = np.zeros((k * n, pts), dtype=object)
    A[:, :sa] = np.asarray(a, dtype=np.uint64).reshape(m * k, sa).astype(object)
    B[:, :sb] = np.asarray(b, dtype=np.uint64).reshape(k * n, sb).astype(object)
    if small_left:
        A[:, :hdeg] = A[:, :hdeg][:, ::-1].copy()
    else:
        B[:, :hdeg] = B[:, :hdeg][:, ::-1].copy()
    out = np.zeros((m * n, sc), dtype=object)
    idx = np.arange(pts)
    for i in range(m):
        for j in range(n):
            for q in range(k):
                pa, pb = A[i * k + q], B[q * n + j]
                for t in range(sc):
                    if small_left:
                        out[i * n + j, t] += int(np.dot(pa[(idx - t) % pts], pb))
                    else:
                        out[i * n + j, t] += int(np.dot(pa, pb[(idx - t) % pts]))
    return np.array([int(x) % p for x in out.reshape(-1)], dtype=np.uint64).astype(np.uint32)


def test_midproduct_golden(ctx):
    with open(os.path.join(GOLD, "midproduct.json")) as f:
        cases = json.load(f)
    for c in cases:
        say("midproduct golden", c["p"], c["m"], c["k"], c["n"], c["sa"], c["sb"], c["sc"], c["small_left"], c["n0"],
            c["n1"])
        args = (ctx, c["p"], c["m"], c["k"], c["n"], np.array(c["a"], dtype=np.uint32), c["sa"],
                np.array(c["b"], dtype=np.uint32), c["sb"], c["sc"], c["small_left"], c["n0"], c["n1"])
        if c["status"]:
            with pytest.raises(LinboxError) as e:
                gpu_mid(*args)
            assert e.value.status == c["status"]
            continue
        for method in (lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA):
            assert gpu_mid(*args, method=method).tolist() == c["c"]


@pytest.mark.parametrize("small_left", [True, False])
def test_midproduct_vs_oracle(ctx, small_left):
    rng = np.random.default_rng(11 + small_left)
    # (p, m, k, n, hdeg, big, sc, n1): lengths 2^2 .. 2^14 (pts >= deg), wrapped outputs, both pointwise forms
    cases = [(998244353, 2, 3, 2, 3, 5, 3, 0), (2013265921, 3, 4, 5, 9, 17, 9, 0), (469762049, 4, 2, 3, 20, 39, 20, 0),
             (167772161, 16, 16, 16, 64, 127, 64, 0), (998244353, 17, 20, 18, 100, 180, 120, 200),
             (2013265921, 16, 17, 33, 257, 513, 257, 0), (469762049, 20, 16, 16, 700, 1300, 800, 1400),
             (2013265921, 16, 16, 16, 2048, 4095, 2048, 0), (998244353, 16, 16, 17, 4096, 8191, 4100, 0),
             (469762049, 16, 16, 16, 8192, 16383, 8192, 0)]
    for (p, m, k, n, hdeg, big, sc, n1) in cases:
        sa, sb = (hdeg, big) if small_left else (big, hdeg)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.midproduct_fft_prime(p, m, k, n, a, sa, b, sb, sc, small_left, hdeg, n1)
        assert st == 0
        for method in (lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA):
            say("midproduct vs oracle", p, m, k, n, sa, sb, sc, small_left, n1, method)
            got = gpu_mid(ctx, p, m, k, n, a, sa, b, sb, sc, small_left, hdeg, n1, method)
            assert np.array_equal(got, ref)


@pytest.mark.parametrize("p", [1000000007, 4294967291, 3221225473, 97])
def test_midproduct_crt_route(ctx, p):
    """Primes with too small a 2-valuation and primes >= 2^31 go through FFT
    primes and CRT; the result is the same correlation mod p."""
    rng = np.random.default_rng(p % 1000)
    for (m, k, n, hdeg, big, sc, small_left, n1) in [(2, 3, 2, 5, 9, 5, True, 0), (3, 2, 4, 7, 12, 9, False, 13),
                                                     (4, 5, 3, 33, 60, 33, True, 0)]:
        sa, sb = (hdeg, big) if small_left else (big, hdeg)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        say("midproduct crt", p, m, k, n, sa, sb, sc, small_left, n1)
        got = gpu_mid(ctx, p, m, k, n, a, sa, b, sb, sc, small_left, hdeg, n1)
        assert np.array_equal(got, correlation(p, m, k, n, a, sa, b, sb, sc, small_left, hdeg, n1))


def test_midproduct_naive(ctx):
    rng = np.random.default_rng(21)
    for p in PRIMES + [4294967291]:
        for (m, k, n, sa, sb, sc, n0, n1) in [(3, 4, 2, 2, 3, 4, 2, 0), (2, 2, 2, 1, 1, 1, 0, 0),
                                              (4, 3, 5, 3, 9, 4, 3, 6), (2, 5, 3, 6, 2, 4, 4, 0)]:
            a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
            sentinel = np.full(m * n * sc, 7, dtype=np.uint32)
            st, ref = orc.midproduct_naive(p, m, k, n, a, sa, b, sb, sentinel, True, n0, n1)
            assert st == 0
            say("midproduct naive", p, m, k, n, sa, sb, sc, n0, n1)
            A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
            B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n, "matfirst")
            C = PolynomialMatrix.from_polfirst(sentinel.reshape(m * n, sc), m, n, "matrowfirst")
            PolynomialMatrixNaiveMulDomain(p, ctx).midproduct(C, A, B, True, n0, n1)
            assert np.array_equal(C.to_polfirst().reshape(-1), ref)


def test_midproduct_domain_dispatch_and_views(ctx):
    """PolynomialMatrixMulDomain.midproduct / midproductgen pick the FFT or naive
    route like polynomial-matrix-domain.h:77-103, on views and mixed layouts."""
    p = 469762049
    rng = np.random.default_rng(31)
    m, k, n, hdeg = 3, 4, 2, 30
    Abig = PolynomialMatrix.from_polfirst(rand(rng, p, m * k * 45).reshape(m * k, 45), m, k, "matfirst")
    Bbig = PolynomialMatrix.from_polfirst(rand(rng, p, k * n * 70).reshape(k * n, 70), k, n, "polfirst")
    A, B = Abig.at(5, 5 + hdeg - 1), Bbig.at(2, 2 + 2 * hdeg - 2)
    Cbig = PolynomialMatrix.from_polfirst(np.full((m * n, 50), 7, dtype=np.uint32), m, n, "matrowfirst")
    C = Cbig.at(10, 10 + hdeg - 1)
    say("midproduct views")
    PolynomialMatrixMulDomain(p, ctx).midproduct(C, A, B)
    st, ref = orc.midproduct_fft_prime(p, m, k, n, A.to_polfirst().reshape(-1), hdeg, B.to_polfirst().reshape(-1),
                                       2 * hdeg - 1, hdeg, True, 0, 0)
    full = Cbig.to_polfirst()
    assert np.array_equal(full[:, 10:10 + hdeg].reshape(-1), ref)
    assert (full[:, :10] == 7).all() and (full[:, 10 + hdeg:] == 7).all()
    # midproductgen with c.size() <= 4 takes the naive route: coefficients past
    # min(sa+sb-1, deg) stay as they were
    As = PolynomialMatrix.from_polfirst(rand(rng, p, m * k * 2).reshape(m * k, 2), m, k)
    Bs = PolynomialMatrix.from_polfirst(rand(rng, p, k * n * 3).reshape(k * n, 3), k, n)
    Cs = PolynomialMatrix.from_polfirst(np.full((m * n, 4), 7, dtype=np.uint32), m, n)
    PolynomialMatrixMulDomain(p, ctx).midproductgen(Cs, As, Bs, True, 2, 0)
    prod = orc.mul_naive(p, m, k, n, As.to_polfirst().reshape(-1), 2, Bs.to_polfirst().reshape(-1), 3).reshape(m * n, 4)
    got = Cs.to_polfirst()
    assert np.array_equal(got[:, :2], prod[:, 1:3]) and (got[:, 2:] == 7).all()


def test_midproduct_errors(ctx):
    p = 998244353
    dom = PolynomialMatrixFFTMulDomain(p, ctx)
    with pytest.raises(LinboxError) as e:  # small operand longer than hdeg
        dom.midproduct(PolynomialMatrix(2, 2, 4), PolynomialMatrix(2, 3, 5), PolynomialMatrix(3, 2, 7))
    assert e.value.status == 5
    with pytest.raises(LinboxError) as e:  # other operand size >= hdeg + deg
        dom.midproduct(PolynomialMatrix(2, 2, 4), PolynomialMatrix(2, 3, 4), PolynomialMatrix(3, 2, 11))
    assert e.value.status == 5
    with pytest.raises(LinboxError) as e:  # c.size() < deg - hdeg
        dom.midproduct(PolynomialMatrix(2, 2, 2), PolynomialMatrix(2, 3, 4), PolynomialMatrix(3, 2, 6), True, 4, 9)
    assert e.value.status == 5
    with pytest.raises(LinboxError) as e:  # dimension mismatch
        dom.midproduct(PolynomialMatrix(2, 2, 4), PolynomialMatrix(2, 3, 4), PolynomialMatrix(2, 2, 7))
    assert e.value.status == 5
    with pytest.raises(LinboxError) as e:  # a one-point transform
        dom.midproduct(PolynomialMatrix(1, 1, 1), PolynomialMatrix(1, 1, 1), PolynomialMatrix(1, 1, 1))
    assert e.value.status == 3


def test_midproduct_pm_basis_level(ctx):
    """One PM-Basis level: sigma1 (m x m, ord1+1 coefficients) times serie
    (m x n, ord1+ord2), hdeg = ord1+1, deg = ord1+ord2, ord1 = ord2 = 4096."""
    p = 2013265921
    m, n, ord1, ord2 = 32, 16, 4096, 4096
    rng = np.random.default_rng(41)
    sig = rand(rng, p, m * m * (ord1 + 1))
    ser = rand(rng, p, m * n * (ord1 + ord2))
    S1 = PolynomialMatrix.from_polfirst(sig.reshape(m * m, ord1 + 1), m, m)
    SE = PolynomialMatrix.from_polfirst(ser.reshape(m * n, ord1 + ord2), m, n)
    S2 = PolynomialMatrix(m, n, ord2)
    say("pm-basis level midproductgen")
    PolynomialMatrixMulDomain(p, ctx).midproductgen(S2, S1, SE, True, ord1 + 1, ord1 + ord2)
    P = PolynomialMatrix(m, n, 1)
    PolynomialMatrixMulDomain(p, ctx).mul(P, S1, SE)
    got = S2.to_polfirst()
    assert np.array_equal(got, P.to_polfirst()[:, ord1:ord1 + ord2])
    # one output row against the oracle's product
    a0 = sig.reshape(m, m, ord1 + 1)[0].reshape(-1)
    st, row = orc.mul_fft_prime(p, 1, m, n, a0, ord1 + 1, ser, ord1 + ord2)
    assert st == 0
    assert np.array_equal(got[:n], row.reshape(n, -1)[:, ord1:ord1 + ord2])
