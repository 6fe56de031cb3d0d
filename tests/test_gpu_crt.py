"""GPU parity of the multiprecision (CRT) route against the oracle and fixtures.

* crt_mul.json: products over a seeded random 192-bit prime P (Python big-int
  convolution fixtures).
* 32x32 . 32x32, degree 1023, coefficients uniform in [0,P) for the 192-bit P,
  14 NTT primes, word-for-word against the oracle's CRT product.
* Word-size moduli that are not FFT primes: 1000000007 and the largest 32-bit
  prime, against the oracle's naive product.
* Ten multiword products of distinct shapes queued on device buffers with no
  synchronisation between them: their per-prime products go through the batch
  kernels, and the tenth shape evicts the cached batch plans while earlier
  launches may still be queued; every result against the oracle.
"""
import json
import os
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, PolynomialMatrixMulDomain
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
METHODS = [lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = lbx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def crt():
    with open(os.path.join(GOLD, "crt_mul.json")) as f:
        return json.load(f)


def gpu_mp(ctx, P, nl, m, k, n, a, sa, b, sb, method):
    A = PolynomialMatrix(m, k, sa, nlimbs=nl, data=np.ascontiguousarray(a, dtype=np.uint32))
    B = PolynomialMatrix(k, n, sb, nlimbs=nl, data=np.ascontiguousarray(b, dtype=np.uint32))
    C = PolynomialMatrix(m, n, sa + sb - 1, nlimbs=nl)
    lbx.mul_mp(P, C, A, B, ctx=ctx, method=method)
    return C.data


def test_crt_golden(ctx, crt):
    P = int(crt["P_hex"], 16)
    for cs in crt["cases"]:
        for method in METHODS:
            say("crt golden", cs["m"], cs["k"], cs["n"], cs["sa"], cs["sb"], method)
            got = gpu_mp(ctx, P, crt["nlimbs"], cs["m"], cs["k"], cs["n"], cs["a"], cs["sa"], cs["b"], cs["sb"], method)
            assert got.tolist() == cs["c"]


def random_mp(rng, P, nl, count):
    """count values uniform in [0,P) as nl little-endian words each: the top word is drawn
    in [0, top(P)], and the rare rows whose top word equals top(P) are redrawn until < P."""
    ptop = P >> (32 * (nl - 1))
    words = rng.integers(0, 1 << 32, (count, nl), dtype=np.uint64)
    words[:, nl - 1] = rng.integers(0, ptop + 1, count, dtype=np.uint64)
    for i in np.nonzero(words[:, nl - 1] == ptop)[0]:
        while sum(int(w) << (32 * j) for j, w in enumerate(words[i])) >= P:
            words[i, :nl - 1] = rng.integers(0, 1 << 32, nl - 1, dtype=np.uint64)
    return words.astype(np.uint32).reshape(-1)


def test_config4_full_size(ctx, crt):
    """32x32.32x32 degree 1023 over the 192-bit P: 14 primes, bit-exact vs the oracle."""
    P = int(crt["P_hex"], 16)
    nl = crt["nlimbs"]
    m = k = n = 32
    sa = sb = 1024
    rng = np.random.default_rng(0x4C494E424F58 + 3)
    a = random_mp(rng, P, nl, m * k * sa)
    b = random_mp(rng, P, nl, k * n * sb)
    say("config4 oracle")
    st, ref, primes = orc.mul_crt(crt["P_limbs"], m, k, n, a, sa, b, sb)
    assert st == 0 and len(primes) == 14
    for method in METHODS:
        say("config4 gpu", method)
        got = gpu_mp(ctx, P, nl, m, k, n, a, sa, b, sb, method)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("p", [1000000007, 4294967291, 3])
def test_non_fft_word_primes(ctx, p):
    rng = np.random.default_rng(p)
    for (m, k, n, sa, sb) in [(3, 4, 5, 37, 22), (16, 16, 16, 64, 64), (1, 1, 1, 1, 1)]:
        a = rng.integers(0, p, m * k * sa, dtype=np.uint64).astype(np.uint32)
        b = rng.integers(0, p, k * n * sb, dtype=np.uint64).astype(np.uint32)
        ref = orc.mul_naive(p, m, k, n, a, sa, b, sb)
        for method in METHODS:
            say("non-FFT prime", p, m, k, n, sa, sb, method)
            A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
            B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
            C = PolynomialMatrix(m, n, 1)
            PolynomialMatrixMulDomain(p, ctx, method).mul(C, A, B)
            assert np.array_equal(C.to_polfirst().reshape(-1), ref)


def test_fft_prime_too_short_for_length(ctx):
    """97 = 3*2^5+1 is an FFT prime only up to 32 points: a product needing 64
    points takes the CRT route instead."""
    p = 97
    rng = np.random.default_rng(1)
    m, k, n, sa, sb = 2, 3, 2, 40, 30
    a = rng.integers(0, p, m * k * sa, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, p, k * n * sb, dtype=np.uint64).astype(np.uint32)
    ref = orc.mul_naive(p, m, k, n, a, sa, b, sb)
    say("valuation too small", p)
    A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
    B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
    C = PolynomialMatrix(m, n, 1)
    PolynomialMatrixMulDomain(p, ctx).mul(C, A, B)
    assert np.array_equal(C.to_polfirst().reshape(-1), ref)


@pytest.mark.parametrize("bits", [320, 480, 512, 1024, 2048])
def test_crt_wide_moduli(ctx, bits):
    """Moduli of 10, 15, 16, 32 and 64 limbs (22 to 133 FFT primes, so every recombination
    kernel shape runs) against the product computed with Python integers."""
    rng = np.random.default_rng(bits)
    P = (1 << (bits - 1)) | int(rng.integers(1, 1 << 62)) << 1 | 1
    nl = (P.bit_length() + 31) // 32
    m, k, n, sa, sb = 2, 3, 2, 6, 5
    a = random_mp(rng, P, nl, m * k * sa)
    b = random_mp(rng, P, nl, k * n * sb)

    def val(words, i):
        return sum(int(words[i * nl + l]) << (32 * l) for l in range(nl))
    say("wide modulus", bits)
    got = gpu_mp(ctx, P, nl, m, k, n, a, sa, b, sb, lbx.METHOD_AUTO)
    for i in range(m):
        for j in range(n):
            for t in range(sa + sb - 1):
                ref = 0
                for q in range(k):
                    for u in range(max(0, t - sb + 1), min(t, sa - 1) + 1):
                        ref += val(a, (i * k + q) * sa + u) * val(b, (q * n + j) * sb + (t - u))
                assert val(got, ((i * n + j) * (sa + sb - 1)) + t) == ref % P, (bits, i, j, t)


def test_batched_primes_with_plan_eviction(ctx, crt):
    from linbox_amd.hip import DeviceBuffer, synchronize
    P, nl = int(crt["P_hex"], 16), crt["nlimbs"]
    top = (P >> (32 * (nl - 1))) & 0xFFFFFFFF
    rng = np.random.default_rng(10)
    shapes = [(3 + i % 3, 4 + i, 2 + i % 4, 40 + 7 * i, 33 + 5 * i) for i in range(10)]
    cases, keep = [], []
    for (m, k, n, sa, sb) in shapes:
        def rand_mp(count):
            w = rng.integers(0, 1 << 32, (count, nl), dtype=np.uint64)
            w[:, nl - 1] = rng.integers(0, top, count, dtype=np.uint64)
            return w.astype(np.uint32).reshape(-1)
        a, b = rand_mp(m * k * sa), rand_mp(k * n * sb)
        da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
        dc = DeviceBuffer(m * n * (sa + sb - 1) * nl * 4)
        A = PolynomialMatrix(m, k, sa, nlimbs=nl, data=da)
        B = PolynomialMatrix(k, n, sb, nlimbs=nl, data=db)
        C = PolynomialMatrix(m, n, sa + sb - 1, nlimbs=nl, data=dc)
        say("queued multiword product", m, k, n, sa, sb)
        lbx.mul_mp(P, C, A, B, ctx=ctx, method=lbx.METHOD_PLAIN_U64)
        cases.append((m, k, n, sa, sb, a, b, dc))
        keep.append((da, db))
    synchronize()
    for (m, k, n, sa, sb, a, b, dc) in cases:
        st, ref, primes = orc.mul_crt(np.array([(P >> (32 * i)) & 0xFFFFFFFF for i in range(nl)], np.uint32),
                                      m, k, n, a, sa, b, sb)
        assert st == 0 and len(primes) >= 6
        assert np.array_equal(dc.download(), ref), (m, k, n, sa, sb)
