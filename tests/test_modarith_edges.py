"""Edge values of the word-size modular arithmetic in the kernels
(linbox_amd/csrc/modarith.h and pw_math.h), built for the host and checked
against Python integers: csub, Shoup products (the reference's
mul_precomp_b_without_reduction, fft-simd.h:123-132), the balanced signed
digits of the INT8 form, the MFMA epilogue that recombines the seven digit
anti-diagonals, and the plain form's 64-bit accumulation with its folds.

Values: 0, 1, p-1, p, 2p-1, 2^31-1, 2^32-1 and the extremes each function's
contract allows, for the four FFT primes of the workloads; the host helpers
(powmod, invmod, shoup_pre) also for 4294967291 (p > 2^31, where the
word-size kernels are not used and the few-primes route keeps 64-bit words)."""
import ctypes
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "modarith_harness.cpp")
OUT = os.path.join(HERE, "cpp", "_build", "libmodarith.so")
ROOT = os.path.dirname(HERE)

PRIMES = [998244353, 2013265921, 469762049, 167772161]
WIDE = 4294967291


@pytest.fixture(scope="module")
def lib():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-shared", "-fPIC", "-o", OUT, SRC], check=True)
    L = ctypes.CDLL(OUT)
    u32, u64 = ctypes.c_uint32, ctypes.c_uint64
    for name, args in [("csub", [u32, u32]), ("shoup_pre", [u32, u32]), ("shoup_lazy", [u32, u32, u32]),
                       ("shoup_mul", [u32, u32, u32]), ("add_mod", [u32, u32, u32]), ("sub_mod", [u32, u32, u32]),
                       ("mont_mul", [ctypes.c_int32, u32, u32]), ("inv_mod32", [u32]),
                       ("to_limbs", [u32, u32]), ("powmod", [u32, u64, u32]), ("invmod", [u32, u32]),
                       ("combine", [ctypes.POINTER(ctypes.c_int32), u32, u32]),
                       ("plain_dot", [ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.c_int, u32, u32]),
                       ("plain_fold_count", [u32]),
                       ("plain_s_dot", [ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.c_int, u32, u32])]:
        f = getattr(L, "lbx_test_" + name)
        f.argtypes = args
        f.restype = u32
    return L


def edges(p):
    return sorted({0, 1, 2, p // 2, (p - 1) // 2, (p + 1) // 2, p - 2, p - 1})


def test_csub_and_add_sub(lib):
    for p in PRIMES:
        for x in [0, 1, p - 1, p, p + 1, 2 * p - 2, 2 * p - 1]:
            assert lib.lbx_test_csub(x, p) == x % p, (p, x)
        for a in edges(p):
            for b in edges(p):
                assert lib.lbx_test_add_mod(a, b, p) == (a + b) % p
                assert lib.lbx_test_sub_mod(a, b, p) == (a - b) % p


def test_add_sub_random_and_largest_word(lib):
    """add_mod and sub_mod on random canonical pairs, and for 2^31 - 1, the
    largest modulus their contract (a, b < p < 2^31) allows, where the wrapped
    difference and the sum come closest to 2^32."""
    rng = random.Random(7)
    for p in PRIMES + [(1 << 31) - 1]:
        pairs = [(a, b) for a in edges(p) for b in edges(p)]
        pairs += [(rng.randrange(p), rng.randrange(p)) for _ in range(2000)]
        for a, b in pairs:
            assert lib.lbx_test_add_mod(a, b, p) == (a + b) % p, (p, a, b)
            assert lib.lbx_test_sub_mod(a, b, p) == (a - b) % p, (p, a, b)


def test_shoup_products(lib):
    rng = random.Random(1)
    for p in PRIMES:
        ws = edges(p) + [rng.randrange(p) for _ in range(6)]
        xs = [0, 1, p - 1, p, 2 * p - 1, (1 << 31) - 1, (1 << 32) - 1] + [rng.getrandbits(32) for _ in range(6)]
        for w in ws:
            assert lib.lbx_test_shoup_pre(w, p) == (w << 32) // p
            for x in xs:
                lazy = lib.lbx_test_shoup_lazy(x, w, p)
                assert 0 <= lazy < 2 * p and lazy % p == x * w % p, (p, w, x)
                assert lib.lbx_test_shoup_mul(x, w, p) == x * w % p, (p, w, x)


def test_signed_montgomery(lib):
    """The forward transforms' twiddle product: x*w mod p in [0,p) for signed
    |x| < p (the difference a - b of two canonical words) and w < p, with the
    twiddle in Montgomery form; p^-1 mod 2^32 by Newton's iteration.  Also for
    2^31 - 1, the largest odd modulus below 2^31, where |x wm| and |m p|
    come closest to 2^62."""
    rng = random.Random(11)
    for p in PRIMES + [(1 << 31) - 1]:
        assert lib.lbx_test_inv_mod32(p) * p % (1 << 32) == 1
        ws = edges(p) + [rng.randrange(p) for _ in range(8)]
        xs = [0, 1, -1, p - 1, -(p - 1), p // 2, -(p // 2)] + [rng.randrange(-p + 1, p) for _ in range(40)]
        for w in ws:
            for x in xs:
                assert lib.lbx_test_mont_mul(x, w, p) == x * w % p, (p, w, x)
        for _ in range(3000):
            a, b, w = rng.randrange(p), rng.randrange(p), rng.randrange(p)
            assert lib.lbx_test_mont_mul(a - b, w, p) == (a - b) * w % p, (p, a, b, w)


def _root(p, L):
    """compute_primitive_root (fft-utils.h:83-104): t^((p-1)/2^L) for the first t = 2, 3, ...
    whose result has order exactly 2^L."""
    t = 2
    while True:
        w = pow(t, (p - 1) >> L, p)
        if pow(w, 1 << (L - 1), p) != 1:
            return w
        t += 1


def test_forward_transform_with_montgomery_tables(lib):
    """The forward transforms read twiddle tables of signed Montgomery pairs in the
    pow_w layout: a DIF through such a table equals the DIF through the Shoup
    table for every length to 2^14 and the four primes, and for lengths to 2^8
    both equal the direct evaluation the reference's FFT test checks
    (test-fft.C:157-176: output i, in bit-reversed order, is P(w^i))."""
    lib.lbx_test_dif.restype = ctypes.c_int
    lib.lbx_test_dif.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_int]
    rng = random.Random(13)
    for p in PRIMES:
        for L in range(1, 15):
            n = 1 << L
            w = _root(p, L)
            x = [rng.randrange(p) for _ in range(n)]
            xm = (ctypes.c_uint32 * n)(*x)
            xs = (ctypes.c_uint32 * n)(*x)
            assert lib.lbx_test_dif(xm, L, p, w, 1) == 0
            assert lib.lbx_test_dif(xs, L, p, w, 0) == 0
            assert list(xm) == list(xs), (p, L)
            if L <= 8:
                rev = [int(format(i, "0%db" % L)[::-1], 2) for i in range(n)]
                for i in range(n):
                    wi = pow(w, i, p)
                    acc = 0
                    for c in reversed(x):
                        acc = (acc * wi + c) % p
                    assert xm[rev[i]] == acc, (p, L, i)


def test_balanced_digits(lib):
    for p in PRIMES:
        for x in edges(p) + [p // 3, 2 * p // 3]:
            v = lib.lbx_test_to_limbs(x, p)
            d = [((v >> (8 * s)) & 0xFF) for s in range(4)]
            d = [t - 256 if t >= 128 else t for t in d]
            assert all(-128 <= t <= 127 for t in d)
            c = x if x <= p // 2 else x - p
            assert sum(t << (8 * s) for s, t in enumerate(d)) == c, (p, x, d)
            assert abs(c) < (1 << 30)


def test_mfma_epilogue_extremes(lib):
    """sum_d D_d 2^(8d) * scale mod p for anti-diagonals at the int32 chunk bound
    |D_d| <= 4 * 4096 * 2^14 = 2^28 (both signs), zeros and random values."""
    rng = random.Random(2)
    B = 1 << 28
    for p in PRIMES:
        for scale in [1, p - 1, pow(1 << 14, p - 2, p), rng.randrange(1, p)]:
            cases = [[0] * 7, [B] * 7, [-B] * 7, [B, -B] * 3 + [B], [-B, B] * 3 + [-B], [1, 0, 0, 0, 0, 0, -1]]
            cases += [[rng.randint(-B, B) for _ in range(7)] for _ in range(20)]
            for D in cases:
                arr = (ctypes.c_int32 * 7)(*D)
                want = sum(D[d] << (8 * d) for d in range(7)) * scale % p
                assert lib.lbx_test_combine(arr, p, scale) == want, (p, scale, D)


def test_plain_accumulation_never_wraps(lib):
    """Inner products of (p-1)*(p-1) terms, the largest the plain form adds,
    over lengths around multiples of the fold period: a 64-bit wrap would
    change the result."""
    for p in PRIMES:
        fold = lib.lbx_test_plain_fold_count(p)
        assert 2 <= fold <= 64
        # the documented bound: 2^62 + 2^33 + fold (p-1)^2 < 2^64, and fold is the largest such n (or 64)
        assert (1 << 62) + (1 << 33) + fold * (p - 1) ** 2 < 1 << 64
        if fold < 64:
            assert (1 << 62) + (1 << 33) + (fold + 1) * (p - 1) ** 2 >= (1 << 64) - 1
        for n in [1, fold - 1, fold, fold + 1, 3 * fold, 4096, 4097]:
            if n < 1:
                continue
            a = (ctypes.c_uint32 * n)(*([p - 1] * n))
            for scale in [1, pow(1 << 13, p - 2, p)]:
                want = n * (p - 1) ** 2 * scale % p
                assert lib.lbx_test_plain_dot(a, a, n, p, scale) == want, (p, n, scale)


def test_plain_signed_accumulation_never_wraps(lib):
    """The point-major plain form: centred operands in (-p/2, p/2], signed
    64-bit sums folded every six products.  Runs of the largest products of
    either sign (both operands at -(p-1)/2, or one at +(p-1)/2) and random
    values, over lengths around the fold period and the k of the workloads."""
    lib.lbx_test_centre.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.lbx_test_centre.restype = ctypes.c_int32
    rng = random.Random(4)
    for p in PRIMES:
        h = (p - 1) // 2
        assert lib.lbx_test_centre(h, p) == h and lib.lbx_test_centre(h + 1, p) == h + 1 - p
        assert lib.lbx_test_centre(0, p) == 0 and lib.lbx_test_centre(p - 1, p) == -1
        lo = h + 1  # centred value -(p-1)/2
        for n in [1, 5, 6, 7, 12, 13, 128, 129, 4097]:
            for (x, y) in [(lo, lo), (h, lo), (h, h)]:
                a = (ctypes.c_uint32 * n)(*([x] * n))
                b = (ctypes.c_uint32 * n)(*([y] * n))
                for scale in [1, pow(1 << 14, p - 2, p)]:
                    assert lib.lbx_test_plain_s_dot(a, b, n, p, scale) == n * x * y * scale % p, (p, n, x, y)
            xs = [rng.randrange(p) for _ in range(n)]
            ys = [rng.randrange(p) for _ in range(n)]
            a = (ctypes.c_uint32 * n)(*xs)
            b = (ctypes.c_uint32 * n)(*ys)
            assert lib.lbx_test_plain_s_dot(a, b, n, p, 1) == sum(u * v for u, v in zip(xs, ys)) % p


def test_host_helpers_wide_modulus(lib):
    """powmod / invmod / shoup_pre for moduli up to 2^32-5 (the few-primes route's range)."""
    rng = random.Random(3)
    for p in PRIMES + [WIDE, 3221225473, 2147483647]:
        for a in [1, 2, p - 1, p - 2] + [rng.randrange(1, p) for _ in range(5)]:
            for e in [0, 1, 2, p - 2, p - 1, (1 << 63) + 5]:
                assert lib.lbx_test_powmod(a, e, p) == pow(a, e, p)
            assert lib.lbx_test_invmod(a, p) * a % p == 1
            assert lib.lbx_test_shoup_pre(a, p) == (a << 32) // p


def test_planted_mistake_is_caught(lib):
    """A digit split without the centring step (digits of x itself) misses
    residues above p/2: the reconstruction check would see it."""
    p = 2013265921
    x = p - 1
    raw = (x + 0x80808080) ^ 0x80808080  # no centring
    d = [((raw >> (8 * s)) & 0xFF) for s in range(4)]
    d = [t - 256 if t >= 128 else t for t in d]
    c = x - p
    assert sum(t << (8 * s) for s, t in enumerate(d)) != c
