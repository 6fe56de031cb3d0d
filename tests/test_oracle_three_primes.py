"""The oracle's few-primes product (PolynomialMatrixThreePrimesFFTMulDomain,
three-primes.inl:57-191) pinned by the naive product, which it must equal
exactly (the product mod p is unique), and by the number of primes, which must
be the shortest prefix of the generated FFT primes whose product exceeds the
bound (p-1)^2 k min(sa,sb) of :75-76."""
import numpy as np
import pytest

from oracle import oracle as orc

NON_FFT = [97, 1048573, 268435399, 536870909]  # 2-valuations of p-1: 5, 2, 1, 2


def rand(rng, p, n):
    return rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32)


def expected_prime_count(p, k, sa, sb, lpts):
    bound = (p - 1) ** 2 * k * min(sa, sb)
    order = orc.generate_fft_primes(1 << 31, 1 << 300, lpts)  # a prefix of the generator's fixed order
    prod, r = 1, 0
    while prod <= bound:
        prod *= order[r]
        r += 1
    return r, order[:r]


def test_primes_are_prime():
    for p in NON_FFT:
        assert orc.lib().or_is_prime_u64(p)


@pytest.mark.parametrize("p", NON_FFT)
@pytest.mark.parametrize("m,k,n,sa,sb", [(2, 3, 2, 17, 9), (4, 4, 4, 64, 64), (3, 1, 5, 40, 1), (5, 6, 2, 33, 120)])
def test_equals_naive(p, m, k, n, sa, sb):
    rng = np.random.default_rng(p % 1000 + m * 7 + sa)
    a = rand(rng, p, m * k * sa)
    b = rand(rng, p, k * n * sb)
    st, c, primes = orc.mul_three_primes(p, m, k, n, a, sa, b, sb)
    assert st == 0
    assert np.array_equal(c, orc.mul_naive(p, m, k, n, a, sa, b, sb))
    pts = 1 << (sa + sb - 2).bit_length()
    if (p - 1) % pts:
        r, want = expected_prime_count(p, k, sa, sb, (pts - 1).bit_length())
        assert primes == want and len(primes) == r


def test_extreme_values_need_every_prime():
    """All coefficients p-1: the product reaches the bound, so a basis one prime short fails."""
    p, m, k, n, sa, sb = 268435399, 3, 8, 3, 100, 100
    a = np.full(m * k * sa, p - 1, np.uint32)
    b = np.full(k * n * sb, p - 1, np.uint32)
    st, c, primes = orc.mul_three_primes(p, m, k, n, a, sa, b, sb)
    assert st == 0 and len(primes) == 3
    assert np.array_equal(c, orc.mul_naive(p, m, k, n, a, sa, b, sb))
    # the exact integer coefficient at the middle is (p-1)^2 k sa, above the product of two primes
    assert (p - 1) ** 2 * k * sa > primes[0] * primes[1]


def test_fft_prime_and_errors():
    rng = np.random.default_rng(3)
    p = 469762049  # 7 * 2^26 + 1, an FFT prime below 2^29
    a = rand(rng, p, 2 * 2 * 30)
    b = rand(rng, p, 2 * 2 * 30)
    st, c, primes = orc.mul_three_primes(p, 2, 2, 2, a, 30, b, 30)
    assert st == 0 and primes == []
    assert np.array_equal(c, orc.mul_fft_prime(p, 2, 2, 2, a, 30, b, 30)[1])
    st, _, _ = orc.mul_three_primes(1 << 29 | 1, 1, 1, 1, a[:4], 4, b[:4], 4)
    assert st == 8  # more than 29 bits (three-primes.inl:59-63)
