"""GPU tests of lbx_polymat_mul_zz, PolynomialMatrixFFTMulDomain<ZRing<integer>>::mul
(matpoly-mult-fft-multiprecision.inl:43-89, mul_crtla :141-277): polynomial matrix
products over the integers, two's complement coefficients of 1 to 32 words.

* Small shapes against the product computed with Python integers: random signed
  coefficients of several bit sizes, the extreme values -2^(32w-1), 2^(32w-1)-1, -1,
  0, 1, zero matrices, constant matrices, both pointwise forms.
* max_rowdeg: the cyclic product mod x^pts - 1 truncated, as lbx_polymat_mul_mp.
* Larger shapes (many primes, so the per-prime products go through the batch
  kernels) by evaluation at random points modulo a 127-bit prime.
* Errors: an output word count too small for the bound, limb counts out of range.
"""
import random
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, LinboxError
from linbox_amd.polymat import mul_zz, to_limbs, from_limbs

pytestmark = pytest.mark.gpu
Q127 = (1 << 127) - 1  # a Mersenne prime


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = lbx.Context(0)
    yield c
    c.close()


def rand_signed(rnd, count, bits):
    return [rnd.randrange(-(1 << bits) + 1, 1 << bits) for _ in range(count)]


def naive(A, B, m, k, n, sa, sb):
    C = [0] * (m * n * (sa + sb - 1))
    for i in range(m):
        for j in range(n):
            for q in range(k):
                a = A[(i * k + q) * sa:(i * k + q + 1) * sa]
                b = B[(q * n + j) * sb:(q * n + j + 1) * sb]
                base = (i * n + j) * (sa + sb - 1)
                for u, x in enumerate(a):
                    if x:
                        for v, y in enumerate(b):
                            C[base + u + v] += x * y
    return C


def out_words(A, B, k, sa, sb):
    """Words the call needs: bound = 2 maxA maxB k min(sa,sb), maxA = 2^(max bit size)."""
    la = max(abs(x).bit_length() for x in A)
    lb = max(abs(x).bit_length() for x in B)
    bound = 2 * (1 << la) * (1 << lb) * k * min(sa, sb)
    return (bound.bit_length() + 31) // 32


def gpu_zz(ctx, A, B, m, k, n, sa, sb, wi, wo, method=lbx.METHOD_AUTO, max_rowdeg=0):
    a = PolynomialMatrix(m, k, sa, data=to_limbs(A, wi), nlimbs=wi)
    b = PolynomialMatrix(k, n, sb, data=to_limbs(B, wi), nlimbs=wi)
    sc = max_rowdeg + 1 if max_rowdeg else sa + sb - 1
    c = PolynomialMatrix(m, n, sc, nlimbs=wo)
    mul_zz(c, a, b, ctx=ctx, method=method, max_rowdeg=max_rowdeg)
    assert c.size() == sc
    return from_limbs(c.data, wo)


@pytest.mark.parametrize("wi,bits", [(1, 20), (1, 31), (2, 50), (4, 127), (8, 200), (32, 1000)])
@pytest.mark.parametrize("method", [lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA])
def test_zz_small_vs_python(ctx, wi, bits, method):
    rnd = random.Random(wi * 1000 + bits + method)
    m, k, n, sa, sb = 3, 4, 2, 9, 7
    A = rand_signed(rnd, m * k * sa, bits)
    B = rand_signed(rnd, k * n * sb, bits)
    wo = out_words(A, B, k, sa, sb)
    say("zz", wi, bits, method, wo)
    assert gpu_zz(ctx, A, B, m, k, n, sa, sb, wi, wo, method) == naive(A, B, m, k, n, sa, sb)
    # a wider output than needed gives the same values
    assert gpu_zz(ctx, A, B, m, k, n, sa, sb, wi, wo + 2, method) == naive(A, B, m, k, n, sa, sb)


@pytest.mark.parametrize("wi", [1, 3, 32])
def test_zz_extreme_values(ctx, wi):
    lo, hi = -(1 << (32 * wi - 1)), (1 << (32 * wi - 1)) - 1
    rnd = random.Random(wi)
    m, k, n, sa, sb = 2, 5, 3, 6, 4
    pool = [lo, hi, -1, 0, 1, lo + 1, hi - 1]
    A = [rnd.choice(pool) for _ in range(m * k * sa)]
    B = [rnd.choice(pool) for _ in range(k * n * sb)]
    wo = out_words(A, B, k, sa, sb)
    say("zz extremes", wi, wo)
    assert gpu_zz(ctx, A, B, m, k, n, sa, sb, wi, wo) == naive(A, B, m, k, n, sa, sb)
    # all lo: every product coefficient at its most positive
    A = [lo] * (m * k * sa)
    B = [lo] * (k * n * sb)
    assert gpu_zz(ctx, A, B, m, k, n, sa, sb, wi, out_words(A, B, k, sa, sb)) == naive(A, B, m, k, n, sa, sb)


def test_zz_zero_and_constant(ctx):
    m, k, n = 3, 3, 3
    Z = [0] * (m * k * 5)
    assert gpu_zz(ctx, Z, [7] * (k * n * 5), m, k, n, 5, 5, 1, 1) == [0] * (m * n * 9)
    rnd = random.Random(9)
    A = rand_signed(rnd, m * k, 60)
    B = rand_signed(rnd, k * n, 60)
    assert gpu_zz(ctx, A, B, m, k, n, 1, 1, 2, out_words(A, B, k, 1, 1)) == naive(A, B, m, k, n, 1, 1)


def test_zz_max_rowdeg_wraps(ctx):
    rnd = random.Random(5)
    m, k, n, sa, sb = 2, 3, 2, 12, 10
    A = rand_signed(rnd, m * k * sa, 90)
    B = rand_signed(rnd, k * n * sb, 90)
    full = naive(A, B, m, k, n, sa, sb)
    for mr in (13, 15, 20, 25):
        sc = mr + 1
        pts = 1
        while pts < sc:
            pts <<= 1
        ref = []
        for e in range(m * n):
            row = [0] * pts
            for t in range(sa + sb - 1):
                row[t % pts] += full[e * (sa + sb - 1) + t]
            ref += row[:sc] + [0] * max(0, sc - pts)
        say("zz max_rowdeg", mr)
        assert gpu_zz(ctx, A, B, m, k, n, sa, sb, 3, out_words(A, B, k, sa, sb), max_rowdeg=mr) == ref


@pytest.mark.parametrize("m,k,n,sa,sb,wi,bits", [(16, 16, 16, 300, 250, 8, 250), (8, 12, 6, 1000, 700, 2, 63),
                                                 (32, 32, 32, 64, 64, 4, 120)])
def test_zz_larger_at_random_points(ctx, m, k, n, sa, sb, wi, bits):
    rnd = random.Random(m * k + sa)
    A = rand_signed(rnd, m * k * sa, bits)
    B = rand_signed(rnd, k * n * sb, bits)
    wo = out_words(A, B, k, sa, sb)
    say("zz larger", m, k, n, sa, sb, wi, bits, wo)
    C = gpu_zz(ctx, A, B, m, k, n, sa, sb, wi, wo)
    sc = sa + sb - 1

    def ev(vals, e, s, x):
        acc = 0
        for t in range(s - 1, -1, -1):
            acc = (acc * x + vals[e * s + t]) % Q127
        return acc
    for _ in range(2):
        x = rnd.randrange(Q127)
        EA = [ev(A, e, sa, x) for e in range(m * k)]
        EB = [ev(B, e, sb, x) for e in range(k * n)]
        for i in range(m):
            for j in range(n):
                want = sum(EA[i * k + q] * EB[q * n + j] for q in range(k)) % Q127
                assert ev(C, i * n + j, sc, x) == want, (i, j)


def test_zz_errors(ctx):
    rnd = random.Random(1)
    m, k, n, sa, sb = 2, 2, 2, 4, 4
    A = rand_signed(rnd, m * k * sa, 40)
    B = rand_signed(rnd, k * n * sb, 40)
    wo = out_words(A, B, k, sa, sb)
    with pytest.raises(LinboxError) as e:
        gpu_zz(ctx, A, B, m, k, n, sa, sb, 2, wo - 1)
    assert e.value.status == 5  # LBX_ERR_SIZE_MISMATCH
    for wi, wo2 in ((33, 70), (2, 81)):
        a = PolynomialMatrix(m, k, sa, nlimbs=wi)
        b = PolynomialMatrix(k, n, sb, nlimbs=wi)
        c = PolynomialMatrix(m, n, sa + sb - 1, nlimbs=wo2)
        with pytest.raises(LinboxError) as e:
            mul_zz(c, a, b, ctx=ctx)
        assert e.value.status == 11  # LBX_ERR_UNSUPPORTED
