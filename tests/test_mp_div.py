"""The two-digit Knuth remainder the CRT kernels reduce with (linbox_amd/csrc/mp_div.h),
built for the host and checked against Python integers.

The CRT recombination forms U = X 2^sh with X < 2^36 P, where the shift sh = 32 off + s
moves P to the top of NL words with its top bit set (Q = P 2^sh).  The remainder of U by Q
must be (X mod P) 2^sh for every modulus size 1..NL words, every leading-word size and
every X up to the bound, the edges included (X = 0, P-1, P, multiples of P, 2^36 P - 1)."""
import ctypes
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "mp_div_harness.cpp")
OUT = os.path.join(HERE, "cpp", "_build", "libmpdiv.so")


@pytest.fixture(scope="module")
def lib():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run(["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-o", OUT, SRC], check=True)
    return ctypes.CDLL(OUT)


def limbs(x, n):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(n)]


def value(words):
    return sum(int(w) << (32 * i) for i, w in enumerate(words))


@pytest.mark.parametrize("nl", [2, 4, 8, 16])
def test_remainder_matches_big_integers(lib, nl):
    fn = {2: lib.lbx_test_rem2, 4: lib.lbx_test_rem4, 8: lib.lbx_test_rem8, 16: lib.lbx_test_rem16}[nl]
    rng = random.Random(nl)
    checked = 0
    for _ in range(6000):
        n = rng.randint(1, nl)
        top = rng.choice([1, 2, 3, 0xFFFFFFFF, 1 << 31, rng.getrandbits(32) | 1, rng.getrandbits(8) | 1])
        P = (top << (32 * (n - 1))) | (rng.getrandbits(32 * (n - 1)) if n > 1 else 0)
        if P < 2:
            continue
        s = 32 - (P >> (32 * (n - 1))).bit_length()
        sh = 32 * (nl - n) + s
        Q = P << sh
        X = rng.choice([0, P - 1, P, rng.randrange(1 << 36) * P, (1 << 36) * P - 1,
                        rng.randrange(1 << 36) * P + rng.randrange(P), rng.randrange((1 << 36) * P)])
        u = (ctypes.c_uint32 * (nl + 2))(*limbs(X << sh, nl + 2))
        v = (ctypes.c_uint32 * nl)(*limbs(Q, nl))
        fn(u, v)
        assert value(list(u)[nl:]) == 0
        assert value(list(u)[:nl]) == (X % P) << sh, (nl, n, hex(P), hex(X))
        checked += 1
    assert checked > 5000


def test_a_wrong_divisor_is_caught(lib):
    """A divisor that is not normalized (top bit clear) breaks the two-digit bound: the check sees it."""
    nl, P = 8, (1 << 200) + 12345
    X = (1 << 36) * P - 1
    s = 32 - (P >> (32 * 6)).bit_length()
    sh = 32 * (nl - 7) + s
    u = (ctypes.c_uint32 * (nl + 2))(*limbs(X << sh, nl + 2))
    v = (ctypes.c_uint32 * nl)(*limbs(P << (sh - 1), nl))  # shifted one bit short
    lib.lbx_test_rem8(u, v)
    assert value(list(u)[:nl]) != (X % P) << sh


def test_two_by_one_division(lib):
    """The quotient digit of the remainder's estimate: (u1 2^32 + u0) / d for a
    normalized d and u1 < d through the reciprocal floor((2^64-1)/d) - 2^32
    (Moller and Granlund 2011, Algorithm 4), against Python integers at the
    extremes of d, u1 and u0 and at random."""
    lib.lbx_test_div21.restype = ctypes.c_uint32
    lib.lbx_test_div21.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_uint32)]
    rng = random.Random(77)
    ds = [1 << 31, (1 << 31) + 1, 0xFFFFFFFF, 0xFFFFFFFE, 0xC0000000] + [rng.getrandbits(31) | (1 << 31)
                                                                        for _ in range(200)]
    r = ctypes.c_uint32()
    for d in ds:
        u1s = [0, 1, d - 1, d // 2] + [rng.randrange(d) for _ in range(20)]
        for u1 in u1s:
            for u0 in [0, 1, 0xFFFFFFFF, d - 1, d, rng.getrandbits(32)]:
                num = (u1 << 32) | u0
                q = lib.lbx_test_div21(u1, u0, d, ctypes.byref(r))
                assert (q, r.value) == divmod(num, d), (hex(u1), hex(u0), hex(d))
