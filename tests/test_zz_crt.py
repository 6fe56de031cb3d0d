"""The sign rule of the integer product's reconstruction (lbx_polymat_mul_zz), on the
CPU: for odd primes p_i with product M, the mixed-radix (Garner) digits of (M-1)/2
are all (p_i-1)/2, so X > (M-1)/2 exactly when X's digits, read from the most
significant, first exceed (p_i-1)/2; and X - M taken modulo 2^(32 w) is the w-word
two's complement of the symmetric representative.  Checked against direct integer
comparison on random and boundary values."""
import random


def garner(x, primes):
    v = []
    for i, p in enumerate(primes):
        t = x % p
        for j in range(i):
            t = (t - v[j]) * pow(primes[j], -1, p) % p
        v.append(t)
    return v


def negative_by_digits(v, primes):
    for i in range(len(primes) - 1, -1, -1):
        h = (primes[i] - 1) // 2
        if v[i] != h:
            return v[i] > h
    return False


def test_half_has_digits_p_minus_1_over_2():
    rnd = random.Random(1)
    for _ in range(200):
        primes = rnd.sample([97, 101, 103, 107, 109, 113, 127, 2013265921, 998244353, 469762049, 167772161], 4)
        M = 1
        for p in primes:
            M *= p
        assert garner((M - 1) // 2, primes) == [(p - 1) // 2 for p in primes]


def test_sign_rule_and_twos_complement():
    rnd = random.Random(2)
    primes = [2013265921, 998244353, 469762049, 167772161, 754974721]
    M = 1
    for p in primes:
        M *= p
    w = 6
    half = (M - 1) // 2
    cases = [0, 1, half - 1, half, half + 1, M - 1, M - 2] + [rnd.randrange(M) for _ in range(2000)]
    for x in cases:
        neg = negative_by_digits(garner(x, primes), primes)
        assert neg == (x > half)
        # X mod 2^(32w) built from digits and W_i = prod_{j<i} p_j mod 2^(32w)
        acc, Wi = 0, 1
        for vi, p in zip(garner(x, primes), primes):
            acc = (acc + vi * Wi) % (1 << (32 * w))
            Wi = Wi * p % (1 << (32 * w))
        if neg:
            acc = (acc - M) % (1 << (32 * w))
        sym = x - M if neg else x
        assert acc == sym % (1 << (32 * w))
