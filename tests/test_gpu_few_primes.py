"""GPU parity of the few-primes route (PolynomialMatrixThreePrimesFFTMulDomain,
three-primes.inl:57-328): word primes below 2^29 that are not FFT primes for
the product length.

* Against the oracle's restatement (or_polymat_mul_three_primes) and the naive
  product on small shapes, every layout, device pointers.
* A modulus whose prime basis would include a prime below it (536870909 with
  469762049 in the basis) falls back to the multiprecision route: same product.
* At 64 x 64, degree 1023: the few-primes route against the multiprecision route
  on the same modulus (given as two limbs, which keeps the general CRT), both
  bit-exact, and the device's oracle product.
* Middle products and PM_Basis through the route.
* The mirror class rejects primes above 29 bits.
* One-word moduli from 2^29 to 2^32 (the reference's multiprecision domain) take
  the same route here, with a residue pass when a basis prime is below p and
  64-bit accumulation from 2^31: against the naive product and the general CRT.
"""
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import (PolynomialMatrix, PolynomialMatrixThreePrimesFFTMulDomain, PolynomialMatrixFFTMulDomain,
                        OrderBasis, LinboxError, mul_mp)
from linbox_amd.hip import DeviceBuffer, synchronize
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
NON_FFT = [97, 1048573, 268435399, 536870909]
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


def device_mul(ctx, p, m, k, n, a, sa, b, sb, lay=("polfirst",) * 3, method=lbx.METHOD_AUTO):
    A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k, lay[0])
    B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n, lay[1])
    C = PolynomialMatrix(m, n, 1, lay[2])
    PolynomialMatrixThreePrimesFFTMulDomain(p, ctx, method).mul(C, A, B)
    return C.to_polfirst().reshape(-1)


@pytest.mark.parametrize("p", NON_FFT)
def test_small_shapes_against_oracle(ctx, p):
    rng = np.random.default_rng(p % 977)
    for (m, k, n, sa, sb) in [(2, 3, 2, 17, 9), (4, 4, 4, 64, 64), (3, 1, 5, 40, 1), (5, 6, 2, 33, 120),
                              (1, 1, 1, 2, 2), (7, 9, 3, 300, 250)]:
        say("few primes", p, m, k, n, sa, sb)
        a = rand(rng, p, m * k * sa)
        b = rand(rng, p, k * n * sb)
        st, ref, primes = orc.mul_three_primes(p, m, k, n, a, sa, b, sb)
        assert st == 0
        if m * k * n * sa * sb <= 1 << 21:
            assert np.array_equal(ref, orc.mul_naive(p, m, k, n, a, sa, b, sb))
        for method in (lbx.METHOD_AUTO, lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA):
            assert np.array_equal(device_mul(ctx, p, m, k, n, a, sa, b, sb, method=method), ref), method


def test_extreme_values(ctx):
    """All coefficients p-1: the integer product reaches the bound (three primes)."""
    p, m, k, n, sa, sb = 268435399, 3, 8, 3, 100, 100
    a = np.full(m * k * sa, p - 1, np.uint32)
    b = np.full(k * n * sb, p - 1, np.uint32)
    st, ref, primes = orc.mul_three_primes(p, m, k, n, a, sa, b, sb)
    assert st == 0 and len(primes) == 3
    say("extreme values", p)
    assert np.array_equal(device_mul(ctx, p, m, k, n, a, sa, b, sb), ref)


def test_layouts_and_device_pointers(ctx):
    p = 1048573
    rng = np.random.default_rng(11)
    m, k, n, sa, sb = 5, 4, 3, 70, 90
    a = rand(rng, p, m * k * sa)
    b = rand(rng, p, k * n * sb)
    ref = orc.mul_naive(p, m, k, n, a, sa, b, sb)
    for lay in [("matfirst", "matrowfirst", "polfirst"), ("matrowfirst", "polfirst", "matfirst"),
                ("matfirst", "matfirst", "matrowfirst")]:
        say("layouts", lay)
        assert np.array_equal(device_mul(ctx, p, m, k, n, a, sa, b, sb, lay), ref)
    say("device pointers")
    da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
    dc = DeviceBuffer(m * n * (sa + sb - 1) * 4)
    A = PolynomialMatrix(m, k, sa, data=da)
    B = PolynomialMatrix(k, n, sb, data=db)
    C = PolynomialMatrix(m, n, sa + sb - 1, data=dc)
    PolynomialMatrixThreePrimesFFTMulDomain(p, ctx).mul(C, A, B)
    synchronize()
    assert np.array_equal(dc.download(), ref)


def test_large_shape_against_multiprecision_route(ctx):
    """64 x 64 x 64, degree 1023: few primes vs the general CRT (the same modulus as two limbs)."""
    p, m, k, n, s = 268435399, 64, 64, 64, 1024
    rng = np.random.default_rng(64)
    a = rand(rng, p, m * k * s)
    b = rand(rng, p, k * n * s)
    say("large few primes", p, m, k, n, s)
    got = device_mul(ctx, p, m, k, n, a, s, b, s)
    a2 = np.zeros(m * k * s * 2, np.uint32)
    a2[0::2] = a
    b2 = np.zeros(k * n * s * 2, np.uint32)
    b2[0::2] = b
    A2 = PolynomialMatrix(m, k, s, nlimbs=2)
    A2.data[:] = a2
    B2 = PolynomialMatrix(k, n, s, nlimbs=2)
    B2.data[:] = b2
    C2 = PolynomialMatrix(m, n, 2 * s - 1, nlimbs=2)
    say("large multiprecision two limbs")
    mul_mp(p, C2, A2, B2, ctx=ctx)
    mp = np.asarray(C2.data).reshape(-1, 2)
    assert not mp[:, 1].any()
    assert np.array_equal(got, mp[:, 0])
    # one entry against the oracle's restatement on a 1 x 64 x 1 slice
    st, ref, _ = orc.mul_three_primes(p, 1, k, 1, a[:k * s], s, b.reshape(k, n, s)[:, 0].reshape(-1), s)
    assert st == 0 and np.array_equal(got.reshape(m, n, 2 * s - 1)[0, 0], ref)


def test_fallback_when_a_basis_prime_is_smaller(ctx):
    """536870909 > 469762049, the third prime of the basis: the multiprecision route takes over."""
    p, m, k, n, sa, sb = 536870909, 6, 16, 4, 300, 300
    rng = np.random.default_rng(9)
    a = rand(rng, p, m * k * sa)
    b = rand(rng, p, k * n * sb)
    st, ref, primes = orc.mul_three_primes(p, m, k, n, a, sa, b, sb)
    assert st == 0 and any(q < p for q in primes)
    say("fallback", p, primes)
    assert np.array_equal(device_mul(ctx, p, m, k, n, a, sa, b, sb), ref)


def test_midproducts(ctx):
    rng = np.random.default_rng(77)
    for case, p in enumerate([1048573, 268435399, 97]):
        m, k, n = 3, 5, 2
        hdeg, deg = 200, 450
        sa, sb = 150, 400
        say("few primes midproduct", p)
        a = rand(rng, p, m * k * sa)
        b = rand(rng, p, k * n * sb)
        A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
        B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
        C = PolynomialMatrix(m, n, deg - hdeg + 1)
        PolynomialMatrixThreePrimesFFTMulDomain(p, ctx).midproduct(C, A, B, True, hdeg, deg)
        full = orc.mul_naive(p, m, k, n, a, sa, b, sb).reshape(m * n, sa + sb - 1)
        want = full[:, hdeg - 1:deg]
        assert np.array_equal(C.to_polfirst(), want), case


def test_order_basis_through_few_primes(ctx):
    p, m, n, order = 1048573, 6, 3, 150
    rng = np.random.default_rng(150)
    serie = rand(rng, p, m * n * order)
    shift = [int(x) for x in rng.integers(0, 4, m)]
    say("pm basis few primes", p)
    F = PolynomialMatrix.from_polfirst(serie.reshape(m * n, order), m, n)
    S = PolynomialMatrix(m, m, order + 1)
    sh = list(shift)
    d = OrderBasis(p, ctx).PM_Basis(S, F, order, sh)
    st, ref, dref, shref = orc.pm_basis(p, m, n, serie, order, order, shift)
    assert st == 0 and d == dref and sh == [int(x) for x in shref]
    assert np.array_equal(S.to_polfirst(), ref.reshape(m * m, order + 1)[:, :d + 1])


def test_mirror_rejects_large_primes(ctx):
    with pytest.raises(LinboxError) as e:
        PolynomialMatrixThreePrimesFFTMulDomain(1000000007, ctx)
    assert e.value.status == 8
    # an FFT prime below 2^29 takes the single-prime route, as the reference's mul_fft does
    p = 469762049
    rng = np.random.default_rng(5)
    a = rand(rng, p, 2 * 2 * 50)
    b = rand(rng, p, 2 * 2 * 50)
    assert np.array_equal(device_mul(ctx, p, 2, 2, 2, a, 50, b, 50), orc.mul_naive(p, 2, 2, 2, a, 50, b, 50))
    # the generic domain gives the same product for a non-FFT prime
    p = 1048573
    A = PolynomialMatrix.from_polfirst(a.reshape(4, 50) % p, 2, 2)
    B = PolynomialMatrix.from_polfirst(b.reshape(4, 50) % p, 2, 2)
    C = PolynomialMatrix(2, 2, 1)
    PolynomialMatrixFFTMulDomain(p, ctx).mul(C, A, B)
    assert np.array_equal(C.to_polfirst().reshape(-1),
                          orc.mul_naive(p, 2, 2, 2, (a % p).astype(np.uint32), 50, (b % p).astype(np.uint32), 50))


@pytest.mark.parametrize("p", [1000000007, 2147483647, 3221225473, 4294967291])
def test_word_moduli_above_2_29(ctx, p):
    rng = np.random.default_rng(p % 4093)
    for (m, k, n, sa, sb) in [(3, 5, 2, 60, 45), (2, 9, 3, 400, 300)]:
        for extreme in (False, True):
            say("word modulus", p, m, k, n, sa, sb, extreme)
            if extreme:
                a = np.full(m * k * sa, p - 1, np.uint32)
                b = np.full(k * n * sb, p - 1, np.uint32)
            else:
                a = rand(rng, p, m * k * sa)
                b = rand(rng, p, k * n * sb)
            A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
            B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
            C = PolynomialMatrix(m, n, 1)
            PolynomialMatrixFFTMulDomain(p, ctx).mul(C, A, B)
            got = C.to_polfirst().reshape(-1)
            if m * k * n * sa * sb <= 1 << 21:
                assert np.array_equal(got, orc.mul_naive(p, m, k, n, a, sa, b, sb))
            st, ref, _ = orc.mul_crt(np.array([p], np.uint32), m, k, n, a, sa, b, sb)
            assert st == 0 and np.array_equal(got, ref)


@pytest.mark.parametrize("p", [268435399, 1000000007])
def test_longest_transform(ctx, p):
    """2^15 points (degree 32766), the longest LDS-resident transform: operands used as they are
    (268435399) and reduced first (1000000007, below which a basis prime lies)."""
    rng = np.random.default_rng(p % 997)
    m, k, n, s = 2, 2, 2, 16384
    a = rand(rng, p, m * k * s)
    b = rand(rng, p, k * n * s)
    say("few primes 2^15 points", p)
    A = PolynomialMatrix.from_polfirst(a.reshape(m * k, s), m, k)
    B = PolynomialMatrix.from_polfirst(b.reshape(k * n, s), k, n)
    C = PolynomialMatrix(m, n, 1)
    PolynomialMatrixFFTMulDomain(p, ctx).mul(C, A, B)  # the three-primes mirror refuses 30-bit primes
    got = C.to_polfirst().reshape(-1)
    st, ref, primes = orc.mul_crt(np.array([p], np.uint32), m, k, n, a, s, b, s)
    assert st == 0 and np.array_equal(got, ref)
