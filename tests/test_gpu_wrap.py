"""max_rowdeg on the routes that do not run one FFT prime: word moduli that
are not FFT primes for the length (the few-primes route,
three-primes.inl:67-84) and multiword moduli (mul_crtla,
multiprecision.inl:141-156).  The result is deg+1 = max_rowdeg+1 coefficients
of the product mod x^pts - 1 with pts = 2^ceil(log2(deg+1)): wrapped when
deg < sa+sb-2, zero past the full degree when deg > sa+sb-2.

* polymat_mul_wrap.json: Python big-int fixtures (tests/golden/make_golden.py)
  for 97, 1000000007, 536870909, 3221225473, 4294967291, a 61-bit modulus
  (two limbs) and a 192-bit one, through lbx_polymat_mul_u32 (one-word moduli)
  and lbx_polymat_mul_mp, both pointwise forms, c pre-filled with nonzero words.
* At a larger size, the wrapped result equals the full product folded mod
  x^pts - 1 (the full product from the oracle's few-primes restatement)."""
import json
import os

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, PolynomialMatrixFFTMulDomain
from oracle import oracle as orc
from test_gpu_parity import ctx, rand, say, METHODS  # noqa: F401

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def cases():
    with open(os.path.join(GOLD, "polymat_mul_wrap.json")) as f:
        return json.load(f)


def test_wrap_golden_mp(ctx, cases):
    for cs in cases:
        P, nl, mr = int(cs["P_hex"], 16), cs["nlimbs"], cs["max_rowdeg"]
        m, k, n, sa, sb = cs["m"], cs["k"], cs["n"], cs["sa"], cs["sb"]
        for method in METHODS:
            say("wrap golden mp", hex(P), m, k, n, sa, sb, mr, method)
            A = PolynomialMatrix(m, k, sa, nlimbs=nl, data=np.array(cs["a"], dtype=np.uint32))
            B = PolynomialMatrix(k, n, sb, nlimbs=nl, data=np.array(cs["b"], dtype=np.uint32))
            C = PolynomialMatrix(m, n, mr + 1, nlimbs=nl, data=np.full(m * n * (mr + 1) * nl, 0x1234567,
                                                                        dtype=np.uint32))
            lbx.mul_mp(P, C, A, B, ctx=ctx, method=method, max_rowdeg=mr)
            assert C.size() == mr + 1
            assert C.data.tolist() == cs["c"]


def test_wrap_golden_word_moduli(ctx, cases):
    for cs in cases:
        if cs["nlimbs"] != 1:
            continue
        p, mr = int(cs["P_hex"], 16), cs["max_rowdeg"]
        m, k, n, sa, sb = cs["m"], cs["k"], cs["n"], cs["sa"], cs["sb"]
        for method in METHODS:
            say("wrap golden word", p, m, k, n, sa, sb, mr, method)
            A = PolynomialMatrix.from_polfirst(np.array(cs["a"], dtype=np.uint32).reshape(m * k, sa), m, k)
            B = PolynomialMatrix.from_polfirst(np.array(cs["b"], dtype=np.uint32).reshape(k * n, sb), k, n,
                                               "matfirst")
            C = PolynomialMatrix.from_polfirst(np.full((m * n, mr + 1), 0x7654321, dtype=np.uint32) % p, m, n,
                                               "matrowfirst")
            PolynomialMatrixFFTMulDomain(p, ctx, method).mul(C, A, B, mr)
            assert C.size() == mr + 1
            assert C.to_polfirst().reshape(-1).tolist() == cs["c"]


def fold(full, m, n, sfull, deg, p):
    pts = 1
    while pts <= deg:
        pts <<= 1
    f = full.reshape(m * n, sfull).astype(object)
    out = np.zeros((m * n, pts), dtype=object)
    for t in range(sfull):
        out[:, t % pts] += f[:, t]
    return (out[:, :deg + 1] % p).astype(np.uint64).astype(np.uint32)


@pytest.mark.parametrize("p", [536870909, 1000000007, 4294967291])
def test_wrap_equals_folded_product(ctx, p):
    """8x12x6 with 700 and 900 coefficients: max_rowdeg 1000 (pts 1024, the
    product of 1599 coefficients wraps) equals the oracle's naive product folded."""
    m, k, n, sa, sb, mr = 8, 12, 6, 700, 900, 1000
    rng = np.random.default_rng(p % 1000)
    a, b = rand(rng, p, m * k * sa), rand(rng, p, k * n * sb)
    full = orc.mul_naive(p, m, k, n, a, sa, b, sb)
    ref = fold(full, m, n, sa + sb - 1, mr, p)
    for method in METHODS:
        say("wrap folded", p, method)
        A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
        B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
        C = PolynomialMatrix(m, n, mr + 1)
        PolynomialMatrixFFTMulDomain(p, ctx, method).mul(C, A, B, mr)
        assert np.array_equal(C.to_polfirst(), ref)
