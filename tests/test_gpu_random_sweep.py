"""Seeded random sweep over the entry points, each case against the oracle.

Shapes, degrees, primes, layouts and pointwise forms are drawn from fixed seeds
so the sweep is reproducible; every drawn case prints its parameters before it
runs.  Products: any prime below 2^32 (FFT primes, word primes that take the CRT
route, primes of too small 2-valuation), both forms, all layouts.  Middle
products: both operand orders and explicit windows.  Order bases: M_Basis and
PM_Basis with random shifts.  Layout conversions: random layouts, windows,
offsets and views on either side.
"""
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, PolynomialMatrixFFTMulDomain, OrderBasis
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
FFT_PRIMES = [998244353, 2013265921, 469762049, 167772161, 754974721, 1811939329]
WORD_PRIMES = [1000000007, 4294967291, 3221225473, 65537, 97, 2]
LAYOUTS = ["polfirst", "matfirst", "matrowfirst"]


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


def ref_product(p, m, k, n, a, sa, b, sb):
    D = sa + sb - 2
    pts = 1
    while pts < D + 1:
        pts <<= 1
    if sa + sb > 2 and p < (1 << 31) and (p - 1) % pts == 0:
        st, c = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        return c
    if m * k * n * sa * sb <= 1 << 22:
        return orc.mul_naive(p, m, k, n, a, sa, b, sb)
    st, c, _ = orc.mul_crt(np.array([p], dtype=np.uint32), m, k, n, a, sa, b, sb)  # exact product mod p
    assert st == 0
    return c


def test_random_products(ctx):
    rng = np.random.default_rng(20261202)
    for case in range(120):
        p = int(rng.choice(FFT_PRIMES + WORD_PRIMES))
        m, k, n = (int(x) for x in rng.integers(1, 40, 3))
        sa = int(np.exp(rng.uniform(0, np.log(3000))))
        sb = max(1, int(sa * rng.uniform(0.2, 1.5)))
        method = int(rng.choice([lbx.METHOD_AUTO, lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA]))
        lay = [str(rng.choice(LAYOUTS)) for _ in range(3)]
        say("product", case, p, m, k, n, sa, sb, method, lay)
        a = rand(rng, p, m * k * sa)
        b = rand(rng, p, k * n * sb)
        A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k, lay[0])
        B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n, lay[1])
        C = PolynomialMatrix(m, n, 1, lay[2])
        PolynomialMatrixFFTMulDomain(p, ctx, method).mul(C, A, B)
        ref = ref_product(p, m, k, n, a, sa, b, sb)
        assert np.array_equal(C.to_polfirst().reshape(-1), ref), case


def test_random_midproducts(ctx):
    rng = np.random.default_rng(20261203)
    for case in range(60):
        p = int(rng.choice(FFT_PRIMES[:4] + [1000000007, 4294967291]))
        m, k, n = (int(x) for x in rng.integers(1, 20, 3))
        hdeg = int(rng.integers(2, 600))
        deg = hdeg + int(rng.integers(1, 600))
        small_left = bool(rng.integers(0, 2))
        ss = int(rng.integers(1, hdeg + 1))
        sbig = int(rng.integers(1, deg + 1))
        sc = deg - hdeg + 1
        sa, sb = (ss, sbig) if small_left else (sbig, ss)
        say("midproduct", case, p, m, k, n, hdeg, deg, small_left, sa, sb)
        a = rand(rng, p, m * k * sa)
        b = rand(rng, p, k * n * sb)
        A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
        B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
        C = PolynomialMatrix(m, n, sc)
        PolynomialMatrixFFTMulDomain(p, ctx).midproduct(C, A, B, small_left, hdeg, deg)
        # coefficients hdeg-1 .. deg-1 of a.b
        full = ref_product(p, m, k, n, a, sa, b, sb).reshape(m * n, sa + sb - 1)
        want = np.zeros((m * n, sc), dtype=np.uint32)
        top = min(sa + sb - 1, deg)
        want[:, :max(0, top - hdeg + 1)] = full[:, hdeg - 1:top]
        assert np.array_equal(C.to_polfirst(), want), case


def test_random_order_bases(ctx):
    rng = np.random.default_rng(20261204)
    for case in range(32):
        p = int(rng.choice([998244353, 2013265921, 1000000007, 469762049, 4294967291]))
        m = int(rng.integers(1, 24))
        n = int(rng.integers(1, min(m, 12) + 1))
        order = int(rng.integers(1, 160))
        shift = [int(x) for x in rng.integers(0, 6, m)]
        which = "pm" if rng.integers(0, 2) else "m"
        say("order basis", case, which, p, m, n, order, shift)
        serie = rand(rng, p, m * n * order)
        F = PolynomialMatrix.from_polfirst(serie.reshape(m * n, order), m, n)
        S = PolynomialMatrix(m, m, order + 1)
        sh = list(shift)
        ob = OrderBasis(p, ctx)
        d = (ob.PM_Basis if which == "pm" else ob.M_Basis)(S, F, order, sh)
        st, ref, dref, shref = (orc.pm_basis if which == "pm" else orc.m_basis)(p, m, n, serie, order, order, shift)
        assert st == 0 and d == dref and sh == [int(x) for x in shref], case
        assert np.array_equal(S.to_polfirst(), ref.reshape(m * m, order + 1)[:, :d + 1]), case


def test_random_copies(ctx):
    rng = np.random.default_rng(20261205)
    for case in range(100):
        r, c = (int(x) for x in rng.integers(1, 24, 2))
        if rng.integers(0, 3) == 0:  # entry counts that are multiples of four (the 16-byte kernel)
            r = 4 * int(rng.integers(1, 8))
        ss = int(rng.integers(1, 300))
        beg = int(rng.integers(0, ss))
        end = int(rng.integers(beg, ss))
        cnt = end - beg + 1
        ds = cnt + int(rng.integers(0, 40))
        start = int(rng.integers(0, ds - cnt + 1))
        ls, ld = str(rng.choice(LAYOUTS)), str(rng.choice(LAYOUTS))
        sview, dview = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
        say("copy", case, r, c, ss, ds, beg, end, start, ls, ld, sview, dview)
        s0 = int(rng.integers(0, 9)) if sview else 0  # views: extra coefficients before the window
        d0 = int(rng.integers(0, 9)) if dview else 0
        sdata = rng.integers(0, 1 << 32, (r * c, ss + s0 + 3), dtype=np.uint64).astype(np.uint32)
        S = PolynomialMatrix.from_polfirst(sdata, r, c, ls)
        src = S.at(s0, s0 + ss - 1) if sview else S
        if not sview:
            S = PolynomialMatrix.from_polfirst(sdata[:, :ss], r, c, ls)
            src = S
        ddata = rng.integers(0, 1 << 32, (r * c, ds + d0 + 2), dtype=np.uint64).astype(np.uint32)
        D = PolynomialMatrix.from_polfirst(ddata if dview else ddata[:, :ds], r, c, ld)
        dst = D.at(d0, d0 + ds - 1) if dview else D
        dst.copy(src, beg, end, start, ctx=ctx)
        want = (ddata if dview else ddata[:, :ds]).copy()
        want[:, d0 + start:d0 + start + cnt] = sdata[:, s0 + beg:s0 + end + 1]
        assert np.array_equal(D.to_polfirst(), want), case
