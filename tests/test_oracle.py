"""Pin the CPU oracle (oracle/lbx_oracle.c) before anything is checked against it.

Each restated function is checked against something that does not share its
code: Horner known answers (tests/test-fft.C:157-176 form), closed forms,
Python big-integer convolution fixtures, sympy primality, and brute force.
The mutation tests at the end plant plausible mistakes in the oracle's inputs
or outputs (a wrong root, a transposed operand, a dropped term, a permuted
coefficient) and show the checks used here reject each of them.
"""
import json
import os
import random

import numpy as np
import pytest

from oracle import oracle as orc
from polycheck import check_by_evaluation

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PRIMES = [998244353, 2013265921, 469762049, 167772161]


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def v2(x):
    v = 0
    while x % 2 == 0:
        x //= 2
        v += 1
    return v


@pytest.fixture(scope="module", autouse=True)
def built():
    orc.build()


# ---------------------------------------------------------------- NTT -----
def test_ntt_horner_kat():
    for case in load("ntt_kat.json"):
        p, L = case["p"], case["L"]
        assert orc.primitive_root(p, L) == case["root"]
        out_br, w = orc.fft_direct(p, L, case["in"])
        assert w == case["root"]
        assert out_br.tolist() == case["out_br"], (p, L)
        # DIT with the same root reads bit-reversed input, writes natural order
        out, _ = orc.fft_inverse(p, L, case["in_br"], root=case["root"])
        assert out.tolist() == case["out"], (p, L)


def test_root_has_exact_order():
    for p in PRIMES:
        for L in range(1, v2(p - 1) + 1):
            w = orc.primitive_root(p, L)
            assert pow(w, 1 << L, p) == 1
            assert pow(w, 1 << (L - 1), p) == p - 1
    # k larger than log2(p-1) is refused (fft-utils.h:92-93)
    with pytest.raises(ValueError):
        orc.primitive_root(998244353, 24)


def test_root_brute_force_small_prime():
    # p = 97 = 3*2^5+1: brute force over t in the specification's order
    p = 97
    for L in range(1, 6):
        w = orc.primitive_root(p, L)
        for t in range(2, p):
            cand = pow(t, (p - 1) >> L, p)
            if pow(cand, 1 << (L - 1), p) == p - 1:
                assert cand == w
                break


def test_init_powers_closed_form():
    for p in PRIMES:
        L = 7
        n = 1 << L
        w = orc.primitive_root(p, L)
        pw, pwp = orc.init_powers(p, L, w)
        off, width, stride = 0, n // 2, 1
        while width >= 1:
            for j in range(width):
                assert pw[off + j] == pow(w, j * stride, p)
            off += width
            width //= 2
            stride *= 2
        assert off == n - 1
        assert all(int(pwp[i]) == (int(pw[i]) << 32) // p for i in range(n - 1))


def test_inverse_roundtrip_and_invroot():
    rng = np.random.default_rng(5)
    for p in PRIMES:
        for L in (1, 2, 5, 10, 12):
            n = 1 << L
            x = rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32)
            y, w = orc.fft_direct(p, L, x)
            wi = orc.invroot(p, L, w)
            assert wi * w % p == 1
            z, _ = orc.fft_inverse(p, L, y, root=wi)
            assert np.array_equal(z.astype(np.uint64), x.astype(np.uint64) * n % p)


def test_ntt_errors():
    with pytest.raises(ValueError):
        orc.fft_direct(998244353, 0, [1])  # k must be positive
    with pytest.raises(ValueError):
        # -1 squares to 1 after one step: refused (fft-utils.h:65-78 checks
        # w^(2^i) != 1 for i < k only, so e.g. root=5 would be accepted)
        orc.fft_direct(998244353, 3, [1] * 8, root=998244352)


# ----------------------------------------------------------- product -----
def test_mul_golden_fixtures():
    for cs in load("polymat_mul.json"):
        st, c = orc.mul_fft_prime(cs["p"], cs["m"], cs["k"], cs["n"], cs["a"], cs["sa"], cs["b"], cs["sb"],
                                  cs["max_rowdeg"])
        if cs["sa"] + cs["sb"] <= 2:
            assert st == 3  # FFT(F,0): k must be positive; the public mul routes this to naive
        else:
            assert st == 0
            assert c.tolist() == cs["c"], (cs["p"], cs["m"], cs["k"], cs["n"], cs["sa"], cs["sb"])
        if cs["max_rowdeg"] == 0:
            cn = orc.mul_naive(cs["p"], cs["m"], cs["k"], cs["n"], cs["a"], cs["sa"], cs["b"], cs["sb"])
            assert cn.tolist() == cs["c"]


@pytest.mark.parametrize("p", PRIMES)
def test_mul_fft_vs_naive_and_evaluation(p):
    rng = np.random.default_rng(p)
    m, k, n, sa, sb = 5, 6, 4, 70, 45
    a = rng.integers(0, p, m * k * sa, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, p, k * n * sb, dtype=np.uint64).astype(np.uint32)
    st, c = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
    assert st == 0
    assert np.array_equal(c, orc.mul_naive(p, m, k, n, a, sa, b, sb))
    xs = [int(x) for x in rng.integers(2, p, 3)]
    assert check_by_evaluation(a, b, c, p, m, k, n, sa, sb, sa + sb - 1, xs)


def test_mul_errors():
    # 7-1 = 6 is not divisible by pts=8: bad FFT prime (wordsize-fast.inl:107-112)
    st, _ = orc.mul_fft_prime(7, 1, 1, 1, [1] * 4, 4, [1] * 4, 4)
    assert st == 1


# ------------------------------------------------------ primes and CRT ----
def test_generate_fft_primes_fixture():
    for case in load("fft_primes.json"):
        got = orc.generate_fft_primes(case["pbound"], int(case["prodbound_hex"], 16), case["kmin"])
        assert got == case["primes"]
    for q in load("fft_primes.json")[0]["primes"]:
        assert orc.lib().or_is_prime_u64(q) == 1


def test_miller_rabin_small_range():
    import sympy
    for x in list(range(0, 3000)) + [2013265921, 2013265923, 4294967291, 4294967295, (1 << 61) - 1]:
        assert bool(orc.lib().or_is_prime_u64(x)) == sympy.isprime(x), x


def test_crt_golden():
    crt = load("crt_mul.json")
    for cs in crt["cases"]:
        st, c, primes = orc.mul_crt(crt["P_limbs"], cs["m"], cs["k"], cs["n"], cs["a"], cs["sa"], cs["b"], cs["sb"])
        assert st == 0
        assert c.tolist() == cs["c"]


def test_crt_config4_prime_count():
    """32x32.32x32, degree 1023 over the 192-bit P needs 14 primes below 2^31."""
    crt = load("crt_mul.json")
    P = int(crt["P_hex"], 16)
    primes = orc.generate_fft_primes(1 << 31, 2 * P * P * 32 * 1024, 11)
    assert len(primes) == 14
    prod = 1
    for q in primes:
        assert (q - 1) % (1 << 12) == 0 and q < (1 << 31)
        prod *= q
    assert prod > 2 * P * P * 32 * 1024


def test_crt_vs_bigint_random():
    rng = random.Random(11)
    crt = load("crt_mul.json")
    P = int(crt["P_hex"], 16)
    nl = crt["nlimbs"]
    m, k, n, sa, sb = 2, 3, 2, 6, 5
    av = [rng.randrange(P) for _ in range(m * k * sa)]
    bv = [rng.randrange(P) for _ in range(k * n * sb)]
    limbs = lambda xs: [(x >> (32 * i)) & 0xFFFFFFFF for x in xs for i in range(nl)]
    st, c, _ = orc.mul_crt(crt["P_limbs"], m, k, n, limbs(av), sa, limbs(bv), sb)
    assert st == 0
    sc = sa + sb - 1
    for i in range(m):
        for j in range(n):
            ref = [0] * sc
            for q in range(k):
                for u in range(sa):
                    for v in range(sb):
                        ref[u + v] += av[(i * k + q) * sa + u] * bv[(q * n + j) * sb + v]
            got = c[(i * n + j) * sc * nl:(i * n + j + 1) * sc * nl].reshape(sc, nl)
            for t in range(sc):
                assert sum(int(got[t, z]) << (32 * z) for z in range(nl)) == ref[t] % P


# ------------------------------------------------------------ midproduct -
def test_midproduct_golden():
    """FFT-prime midproduct against the cyclic-correlation fixtures (Python big
    integers), including outputs where the correlation wraps."""
    for c in load("midproduct.json"):
        st, got = orc.midproduct_fft_prime(c["p"], c["m"], c["k"], c["n"], np.array(c["a"], dtype=np.uint32),
                                           c["sa"], np.array(c["b"], dtype=np.uint32), c["sb"], c["sc"],
                                           c["small_left"], c["n0"], c["n1"])
        assert st == c["status"], c
        if st == 0:
            assert got.tolist() == c["c"]


@pytest.mark.parametrize("small_left", [True, False])
def test_midproduct_valid_window_is_the_middle_product(small_left):
    """Inside the window [0, deg-hdeg] the FFT midproduct equals coefficients
    hdeg-1.. of the plain product (naive convolution), as does the naive
    midproduct, which leaves coefficients past min(sa+sb-1, deg) untouched."""
    rng = np.random.default_rng(5 + small_left)
    for p in PRIMES:
        for (m, k, n, hdeg, big, n1) in [(2, 3, 4, 5, 9, 0), (3, 2, 2, 8, 16, 16), (1, 4, 2, 16, 31, 0),
                                         (2, 2, 3, 7, 10, 13)]:
            deg = n1 or 2 * hdeg - 1
            sa, sb = (hdeg, big) if small_left else (big, hdeg)
            a = rng.integers(0, p, m * k * sa, dtype=np.uint64).astype(np.uint32)
            b = rng.integers(0, p, k * n * sb, dtype=np.uint64).astype(np.uint32)
            sc = deg - hdeg + 1
            st, c = orc.midproduct_fft_prime(p, m, k, n, a, sa, b, sb, sc, small_left, hdeg, n1)
            assert st == 0
            full = orc.mul_naive(p, m, k, n, a, sa, b, sb).reshape(m * n, sa + sb - 1)
            want = np.zeros((m * n, sc), dtype=np.uint32)
            top = min(sa + sb - 1, deg)
            want[:, :top - hdeg + 1] = full[:, hdeg - 1:top]
            assert np.array_equal(c.reshape(m * n, sc), want)
            sentinel = np.full(m * n * sc, 7, dtype=np.uint32)
            st, cn = orc.midproduct_naive(p, m, k, n, a, sa, b, sb, sentinel, small_left, hdeg, n1)
            assert st == 0
            cn = cn.reshape(m * n, sc)
            assert np.array_equal(cn[:, :top - hdeg + 1], want[:, :top - hdeg + 1])
            assert (cn[:, top - hdeg + 1:] == 7).all()


def test_midproduct_errors():
    p, a, b = 998244353, np.ones(4, dtype=np.uint32), np.ones(40, dtype=np.uint32)
    assert orc.midproduct_fft_prime(p, 1, 1, 1, a, 4, b, 40, 4, True, 0, 0)[0] == 5      # b.size >= hdeg+deg
    assert orc.midproduct_fft_prime(1000000007, 1, 1, 1, a, 4, b, 6, 4, True, 0, 0)[0] == 1  # not an FFT prime
    assert orc.midproduct_fft_prime(p, 1, 1, 1, a, 4, b, 6, 2, True, 4, 9)[0] == 5       # c.size < deg-hdeg


# ------------------------------------------------------------ mutations ---
def test_checks_reject_planted_mistakes():
    p = 998244353
    case = [c for c in load("polymat_mul.json") if c["p"] == p and c["m"] == 2 and c["k"] == 3][0]
    m, k, n, sa, sb = case["m"], case["k"], case["n"], case["sa"], case["sb"]
    a = np.array(case["a"], dtype=np.uint32)
    b = np.array(case["b"], dtype=np.uint32)
    gold = case["c"]
    # a wrong root (w^3 is also a primitive root of the same order) breaks the KAT
    kat = [x for x in load("ntt_kat.json") if x["p"] == p and x["L"] == 4][0]
    wrong_root = pow(kat["root"], 3, p)
    out_br, _ = orc.fft_direct(p, 4, kat["in"], root=wrong_root)
    assert out_br.tolist() != kat["out_br"]
    # a transposed operand (b read as n x k) changes the product
    bt = b.reshape(k, n, sb).transpose(1, 0, 2).copy().reshape(-1)
    if k == n:
        _, c = orc.mul_fft_prime(p, m, k, n, a, sa, bt, sb)
        assert c.tolist() != gold
    # a dropped term (one input coefficient zeroed) changes the product
    a2 = a.copy()
    a2[3] = 0
    _, c = orc.mul_fft_prime(p, m, k, n, a2, sa, b, sb)
    assert c.tolist() != gold
    # a permutation of output coefficients passes the x=1 check, not random points
    c_perm = np.array(gold, dtype=np.uint32).reshape(m * n, sa + sb - 1).copy()
    c_perm[0, [1, 2]] = c_perm[0, [2, 1]]
    assert check_by_evaluation(a, b, c_perm, p, m, k, n, sa, sb, sa + sb - 1, [1])
    assert not check_by_evaluation(a, b, c_perm, p, m, k, n, sa, sb, sa + sb - 1, [12345, 777])
    # midproduct with the roots of the two forward transforms exchanged computes
    # a convolution instead of the correlation: the golden fixture rejects it
    mc = [x for x in load("midproduct.json") if x["p"] == p and x["status"] == 0 and x["small_left"]][0]
    st, good = orc.midproduct_fft_prime(p, mc["m"], mc["k"], mc["n"], np.array(mc["a"], dtype=np.uint32), mc["sa"],
                                        np.array(mc["b"], dtype=np.uint32), mc["sb"], mc["sc"], True, mc["n0"],
                                        mc["n1"])
    assert good.tolist() == mc["c"]
    st, swapped = orc.midproduct_fft_prime(p, mc["m"], mc["k"], mc["n"], np.array(mc["a"], dtype=np.uint32),
                                           mc["sa"], np.array(mc["b"], dtype=np.uint32), mc["sb"], mc["sc"], False,
                                           mc["n0"], mc["n1"])
    assert swapped.tolist() != mc["c"]
