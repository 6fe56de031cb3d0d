"""Transforms longer than the LDS-resident 2^15 points (2^16 .. 2^20 here; the
kernels go to 2^27): HBM column passes of up to four stages around the
LDS-resident transform of 2^14-point blocks.  The reference's FFT domain takes
any length up to the prime's 2-adic valuation (fft-utils.h:88-93,
matpoly-mult-fft-wordsize.inl:57-65), and so do products, middle products and
PM_Basis here.  Every result is compared word for word with the oracle."""
import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, PolynomialMatrixFFTMulDomain, OrderBasis
from oracle import oracle as orc
from test_gpu_parity import ctx, rand, say, gpu_mul, METHODS  # noqa: F401

pytestmark = pytest.mark.gpu

PRIMES = [998244353, 2013265921, 469762049, 167772161]


@pytest.mark.parametrize("p", PRIMES)
def test_long_ntt_vs_oracle(ctx, p):
    rng = np.random.default_rng(p + 16)
    for L in range(16, 21):
        count = 3 if L <= 18 else 1
        say("long ntt vs oracle", p, L, count)
        x = rand(rng, p, count << L)
        y = x.copy()
        w = lbx.ntt_forward(p, L, y, ctx=ctx)
        assert w == orc.primitive_root(p, L)
        ref = np.concatenate([orc.fft_direct(p, L, x[i << L:(i + 1) << L])[0] for i in range(count)])
        assert np.array_equal(y, ref), (p, L)
        wi = orc.invroot(p, L, w)
        z = y.copy()
        lbx.ntt_inverse(p, L, z, root=wi, ctx=ctx)
        ref = np.concatenate([orc.fft_inverse(p, L, y[i << L:(i + 1) << L], root=wi)[0] for i in range(count)])
        assert np.array_equal(z, ref), (p, L)
        assert np.array_equal(z.astype(np.uint64), x.astype(np.uint64) * (1 << L) % p)


def test_long_ntt_device_pointers(ctx):
    """In-place transforms of device-resident polynomials (2^17 and 2^21 points)."""
    from linbox_amd.hip import DeviceBuffer, synchronize
    p = 2013265921
    rng = np.random.default_rng(5)
    for L in (17, 21):
        x = rand(rng, p, 2 << L)
        d = DeviceBuffer.from_numpy(x)
        say("long ntt device", L)
        w = lbx.ntt_forward(p, L, d, count=2, ctx=ctx)
        synchronize()
        ref = np.concatenate([orc.fft_direct(p, L, x[i << L:(i + 1) << L])[0] for i in range(2)])
        assert np.array_equal(d.download(), ref)
        lbx.ntt_inverse(p, L, d, count=2, root=orc.invroot(p, L, w), ctx=ctx)
        synchronize()
        assert np.array_equal(d.download().astype(np.uint64), x.astype(np.uint64) * (1 << L) % p)


@pytest.mark.parametrize("method", METHODS)
def test_long_products_vs_oracle(ctx, method):
    """Products of degree 40000 and 100000 (2^17 points) and 2^18 points, both pointwise forms."""
    for (p, m, k, n, sa, sb) in [(998244353, 2, 2, 2, 40001, 30000), (2013265921, 1, 2, 1, 100000, 100001),
                                 (469762049, 2, 3, 2, 65536, 65536), (167772161, 3, 1, 2, 130000, 2)]:
        rng = np.random.default_rng(sa + sb)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        st, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        say("long product", p, m, k, n, sa, sb, method)
        got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, method)
        assert np.array_equal(got.reshape(-1), ref)


def test_long_product_layouts_and_int8_shapes(ctx):
    """Long transforms reading matfirst / matrowfirst operands and writing a
    matrowfirst result, and an INT8 shape with full 32-row bands at 2^16 points."""
    p = 998244353
    rng = np.random.default_rng(77)
    m, k, n, sa, sb = 2, 3, 2, 33000, 32000
    a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
    _, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
    for method in METHODS:
        say("long layouts", method)
        got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, method, layouts=("matfirst", "matrowfirst", "matrowfirst"))
        assert np.array_equal(got.reshape(-1), ref)
    m, k, n, sa, sb = 32, 16, 32, 20000, 20000
    a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
    _, ref = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
    say("long int8 32x16x32")
    got = gpu_mul(ctx, p, m, k, n, a, sa, b, sb, lbx.METHOD_INT8_MFMA)
    assert np.array_equal(got.reshape(-1), ref)


def test_long_midproduct_vs_oracle(ctx):
    """Middle products over 2^16 and 2^17 points (both operand orders)."""
    for (p, m, k, n, hdeg, big, small_left) in [(998244353, 2, 2, 3, 30000, 60000, True),
                                                (2013265921, 3, 2, 2, 70000, 100000, False)]:
        rng = np.random.default_rng(hdeg)
        sa, sb = (hdeg, big) if small_left else (big, hdeg)
        a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
        sc = hdeg
        st, ref = orc.midproduct_fft_prime(p, m, k, n, a, sa, b, sb, sc, small_left)
        assert st == 0
        A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
        B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
        C = PolynomialMatrix(m, n, sc)
        say("long midproduct", p, m, k, n, hdeg, big, small_left)
        PolynomialMatrixFFTMulDomain(p, ctx).midproduct(C, A, B, small_left)
        assert np.array_equal(C.to_polfirst().reshape(-1), ref)


def test_pm_basis_order_65536(ctx):
    """PM_Basis at order 2^16 (the top-level middle product and product run at
    2^16 and 2^17 points) against the oracle, word for word."""
    p, m, n, order = 998244353, 2, 1, 1 << 16
    rng = np.random.default_rng(65536)
    serie = rand(rng, p, m * n * order)
    st, sref, dref, shref = orc.pm_basis(p, m, n, serie, order, order, [0] * m)
    assert st == 0
    S = PolynomialMatrix(m, m, order + 1)
    shift = [0] * m
    say("pm_basis order 65536")
    d = OrderBasis(p, ctx).PM_Basis(S, PolynomialMatrix.from_polfirst(serie.reshape(m * n, order), m, n), order,
                                    shift)
    assert d == dref
    assert shift == [int(x) for x in shref]
    assert np.array_equal(S.to_polfirst(), sref.reshape(m * m, order + 1)[:, :d + 1])
