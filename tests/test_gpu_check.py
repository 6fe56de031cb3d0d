"""GPU tests of the product check: lbx_polymat_check_mul_u32, the random-point
form of check_mul (matpoly-mult-fft.h:87-163, which evaluates at 1), and
lbx_set_check / LBX_CHECK_MUL, the run-time form of CHECK_MATPOL_MUL
(polynomial-matrix-domain.h:72-74).

* The check accepts the oracle's product and rejects it with one coefficient
  changed, at 1 (the reference's point) and at random points, for FFT
  primes, a non-FFT prime and a prime >= 2^31, on polfirst, matfirst and
  matrowfirst storage and on a SubPolynomialMatrix view, host and device.
* deg < c.size() evaluates only c's first deg coefficients.
* With the switch on, products of every route (single prime in both
  pointwise forms, few primes, constant matrices) pass, a truncated product
  (max_rowdeg) is left unchecked, and the switch can be turned off again.
* The middle-product check (lbx_polymat_check_midproduct_u32, check_midproduct
  of matpoly-mult-fft.h:168-262 by random row and column projections, and the
  CHECK_MATPOL_MIDP switch): it accepts the middle products of the FFT, CRT
  and naive routes, both operand orders and explicit windows, rejects any
  checked coefficient changed (first, last, in between, any entry), and
  ignores the coefficients past deg-hdeg that wrap around.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import (PolynomialMatrix, PolynomialMatrixFFTMulDomain, PolynomialMatrixNaiveMulDomain, check_mul,
                        check_midproduct)
from linbox_amd.hip import DeviceBuffer
from oracle import oracle as orc
from test_gpu_parity import ctx, rand, say  # noqa: F401

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _operands(rng, p, m, k, n, sa, sb):
    a = rand(rng, p, m * k * sa)
    b = rand(rng, p, k * n * sb)
    ref = orc.mul_naive(p, m, k, n, a, sa, b, sb)
    return a, b, ref.reshape(m * n, sa + sb - 1)


@pytest.mark.parametrize("p", [998244353, 2013265921, 1000000007, 4294967291])
def test_check_accepts_and_rejects(ctx, p):
    rng = np.random.default_rng(p % 1000)
    m, k, n, sa, sb = 5, 7, 3, 40, 33
    a, b, ref = _operands(rng, p, m, k, n, sa, sb)
    for layout in ("polfirst", "matfirst", "matrowfirst"):
        say("check", p, layout)
        A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k, layout)
        B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n, layout)
        C = PolynomialMatrix.from_polfirst(ref, m, n, layout)
        for x in (1, 0, 12345, p - 1, int(rng.integers(0, p))):
            assert check_mul(p, C, A, B, x=x, ctx=ctx)
        bad = ref.copy()
        bad[4, 17] = (int(bad[4, 17]) + 1) % p
        Cb = PolynomialMatrix.from_polfirst(bad, m, n, layout)
        # a wrong coefficient changes C(x) at every x but the roots of x^17
        for x in (1, 3, 777):
            assert not check_mul(p, Cb, A, B, x=x, ctx=ctx)


def test_check_deg_and_views(ctx):
    p, m, k, n, sa, sb = 998244353, 4, 6, 5, 20, 25
    rng = np.random.default_rng(3)
    a, b, ref = _operands(rng, p, m, k, n, sa, sb)
    A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
    B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
    # c with extra garbage coefficients past the product: deg selects the product part
    big = np.concatenate([ref, rand(rng, p, m * n * 7).reshape(m * n, 7)], axis=1)
    Cbig = PolynomialMatrix.from_polfirst(big, m, n)
    assert check_mul(p, Cbig, A, B, deg=sa + sb - 1, x=99, ctx=ctx)
    assert not check_mul(p, Cbig, A, B, x=99, ctx=ctx)
    # a SubPolynomialMatrix view of a larger matrix holding the product at shift 5
    store = np.zeros((m * n, sa + sb - 1 + 9), dtype=np.uint32)
    store[:, 5:5 + sa + sb - 1] = ref
    S = PolynomialMatrix.from_polfirst(store, m, n)
    V = S.at(5, 5 + sa + sb - 2)
    assert check_mul(p, V, A, B, x=4242, ctx=ctx)
    # device storage
    dA = PolynomialMatrix(m, k, sa, data=DeviceBuffer.from_numpy(a))
    dB = PolynomialMatrix(k, n, sb, data=DeviceBuffer.from_numpy(b))
    dC = PolynomialMatrix(m, n, sa + sb - 1, data=DeviceBuffer.from_numpy(ref.reshape(-1)))
    assert check_mul(p, dC, dA, dB, x=31337, ctx=ctx)


@pytest.mark.parametrize("p,method", [(2013265921, lbx.METHOD_PLAIN_U64), (2013265921, lbx.METHOD_INT8_MFMA),
                                      (1000000007, lbx.METHOD_AUTO), (3221225473, lbx.METHOD_AUTO)])
def test_switch_runs_every_route(ctx, p, method):
    rng = np.random.default_rng(p % 997 + method)
    ctx.set_check(True)
    try:
        for (m, k, n, sa, sb) in [(12, 9, 10, 201, 150), (16, 16, 16, 64, 64), (3, 4, 2, 1, 1)]:
            say("check switch", p, method, m, k, n, sa, sb)
            a, b, ref = _operands(rng, p, m, k, n, sa, sb)
            A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
            B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
            C = PolynomialMatrix(m, n, 1)
            PolynomialMatrixFFTMulDomain(p, ctx, method).mul(C, A, B)
            assert np.array_equal(C.to_polfirst(), ref)
        # a truncated product is not checked (and stays correct)
        a, b, ref = _operands(rng, p, 4, 4, 4, 100, 100)
        A = PolynomialMatrix.from_polfirst(a.reshape(16, 100), 4, 4)
        B = PolynomialMatrix.from_polfirst(b.reshape(16, 100), 4, 4)
        C = PolynomialMatrix(4, 4, 1)
        PolynomialMatrixFFTMulDomain(p, ctx, method).mul(C, A, B, 150)
        assert C.size() == 151
    finally:
        ctx.set_check(False)


def test_environment_switch():
    """A product on a context made with LBX_CHECK_MUL=1 in the environment (child process)."""
    code = (
        "import sys; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "import numpy as np, linbox_amd as lbx\n"
        "from linbox_amd import PolynomialMatrix, PolynomialMatrixFFTMulDomain\n"
        "ctx = lbx.Context(0)\n"
        "p = 998244353\n"
        "A = PolynomialMatrix.from_polfirst(np.arange(8 * 30, dtype=np.uint32).reshape(8, 30), 2, 4)\n"
        "B = PolynomialMatrix.from_polfirst(np.arange(12 * 30, dtype=np.uint32).reshape(12, 30), 4, 3)\n"
        "C = PolynomialMatrix(2, 3, 1)\n"
        "PolynomialMatrixFFTMulDomain(p, ctx).mul(C, A, B)\n"
        "print('checked ok', C.size())\n" % (ROOT, os.path.join(ROOT, "tests")))
    env = dict(os.environ, LBX_CHECK_MUL="1")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "checked ok 59" in r.stdout


# (m, k, n, small operand size, other operand size, c.size, small_left, n0, n1)
MID_CASES = [
    (3, 4, 2, 40, 79, 40, True, 0, 0),     # hdeg = deg-hdeg+1 = c.size
    (2, 5, 3, 40, 79, 40, False, 0, 0),    # b the small operand
    (4, 3, 4, 17, 64, 32, True, 17, 48),   # PM_Basis window: hdeg = n0, deg = n1
    (2, 2, 2, 12, 16, 16, True, 12, 16),   # c.size > deg-hdeg+1: coefficients past it wrap, unchecked
    (6, 8, 5, 300, 700, 300, True, 0, 0),  # split transforms (2^10 points)
]


def _mid(ctx, p, case, naive=False):
    m, k, n, s_small, s_big, sc, small_left, n0, n1 = case
    rng = np.random.default_rng(sum(case[:6]) + p % 97)
    sa, sb = (s_small, s_big) if small_left else (s_big, s_small)
    A = PolynomialMatrix.from_polfirst(rand(rng, p, m * k * sa).reshape(m * k, sa), m, k)
    B = PolynomialMatrix.from_polfirst(rand(rng, p, k * n * sb).reshape(k * n, sb), k, n)
    C = PolynomialMatrix(m, n, sc)
    if naive:
        PolynomialMatrixNaiveMulDomain(p, ctx).midproduct(C, A, B, small_left, n0, n1)
    else:
        PolynomialMatrixFFTMulDomain(p, ctx).midproduct(C, A, B, small_left, n0, n1)
    hdeg = n0 or sc
    deg = n1 or 2 * hdeg - 1
    return A, B, C, min(sc, deg - hdeg + 1)


@pytest.mark.parametrize("p", [998244353, 2013265921, 1000000007, 4294967291])
def test_midproduct_check_accepts_and_rejects(ctx, p):
    rng = np.random.default_rng(p % 1009)
    for case in MID_CASES:
        say("midproduct check", p, case)
        A, B, C, count = _mid(ctx, p, case)
        n0, n1 = case[7], case[8]
        assert check_midproduct(p, C, A, B, n0, n1, ctx=ctx)
        m, n, sc = case[0], case[2], case[5]
        for t in sorted({0, count - 1, int(rng.integers(0, count))}):
            e = int(rng.integers(0, m * n))
            i, j = divmod(e, n)
            v = C.get(i, j, t)
            C.ref_set(i, j, t, (v + 1 + int(rng.integers(0, p - 1))) % p)
            assert not check_midproduct(p, C, A, B, n0, n1, ctx=ctx), (case, t, e)
            C.ref_set(i, j, t, v)
        if sc > count:  # a wrapped coefficient is outside the middle product
            C.ref_set(0, 0, sc - 1, (C.get(0, 0, sc - 1) + 1) % p)
            assert check_midproduct(p, C, A, B, n0, n1, ctx=ctx)


def test_midproduct_check_naive_route(ctx):
    p = 998244353
    for case in [(3, 4, 2, 3, 6, 3, True, 0, 0), (2, 3, 2, 4, 9, 4, True, 4, 7)]:
        say("midproduct check naive", case)
        A, B, C, count = _mid(ctx, p, case, naive=True)
        assert check_midproduct(p, C, A, B, case[7], case[8], ctx=ctx)
        C.ref_set(1, 1, 0, (C.get(1, 1, 0) + 5) % p)
        assert not check_midproduct(p, C, A, B, case[7], case[8], ctx=ctx)


def test_midproduct_switch(ctx):
    """With the switch on, middle products of every route run their check and pass."""
    ctx.set_check(True)
    try:
        for p in (998244353, 1000000007, 4294967291):
            for case in MID_CASES[:3]:
                say("midproduct switch", p, case)
                _mid(ctx, p, case)
        _mid(ctx, 998244353, (3, 4, 2, 3, 6, 3, True, 0, 0), naive=True)
    finally:
        ctx.set_check(False)
