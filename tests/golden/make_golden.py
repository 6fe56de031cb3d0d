"""Generate the committed golden fixtures for the FFT polynomial-matrix path.

Nothing here imports or runs the reference (LinBox is header-only C++ on top of
Givaro/fflas-ffpack, neither of which exists in this image).  Every expected
output is computed from the mathematical definition with Python integers:

* ntt_kat.json     -- Horner evaluation at powers of the root, the known-answer
                      form of tests/test-fft.C:157-176 (DIF(in) == out at
                      bit-reversed positions, DIT(bitrev(in)) == out).  The root
                      is found by a brute-force search written from the
                      specification of compute_primitive_root (fft-utils.h:83-104).
* polymat_mul.json -- schoolbook convolution of small polynomial matrices over
                      Z/pZ for the four word-size primes (the semantics of
                      matpoly-mult-naive.h:56-82), including max_rowdeg truncation
                      (cyclic wrap, matpoly-mult-fft-wordsize-fast.inl:64-80).
* crt_mul.json     -- the same over Z/PZ for a seeded random 192-bit prime P.
* polymat_mul_wrap.json -- max_rowdeg (cyclic wrap and truncation) over word
                      moduli that are not FFT primes and over multiword moduli.
* midproduct.json  -- the FFT-prime midproduct written as the cyclic
                      correlation it computes (matpoly-mult-fft-wordsize-fast.inl:209-316):
                      with pts = 2^ceil(log2 deg) and the first hdeg
                      coefficients of the small operand reversed, c[t] =
                      sum_v a2[(v-t) mod pts] b[v] (small operand on the left)
                      or sum_u a[u] b2[(u-t) mod pts]; cases include outputs past
                      the middle product, where the correlation wraps.
* fft_primes.json  -- primes m*2^k+1 in the visiting order of
                      RandomFFTPrime::generatePrimes (random-fftprime.h:107-133),
                      primality from sympy.

Run:  python tests/golden/make_golden.py   (writes next to this file)
"""
import json
import os
import random

import sympy

HERE = os.path.dirname(os.path.abspath(__file__))
PRIMES = [998244353, 2013265921, 469762049, 167772161]


def v2(x):
    v = 0
    while x % 2 == 0:
        x //= 2
        v += 1
    return v


def primitive_root(p, k):
    """Smallest t >= 2 (in the order t=2,3,...) such that t^((p-1)/2^k) has
    multiplicative order exactly 2^k; that power is the root."""
    for t in range(2, p):
        w = pow(t, (p - 1) >> k, p)
        if pow(w, 1 << (k - 1), p) == p - 1:  # order exactly 2^k
            return w
    raise ValueError("no root")


def bitrev(i, L):
    r = 0
    for _ in range(L):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def horner(poly, x, p):
    acc = 0
    for c in reversed(poly):
        acc = (acc * x + c) % p
    return acc


def ntt_kats(rng):
    cases = []
    for p in PRIMES:
        for L in (1, 2, 3, 4, 5, 6, 8):
            n = 1 << L
            w = primitive_root(p, L)
            inp = [rng.randrange(p) for _ in range(n)]
            out = [horner(inp, pow(w, i, p), p) for i in range(n)]
            out_br = [0] * n
            in_br = [0] * n
            for i in range(n):
                out_br[bitrev(i, L)] = out[i]
                in_br[bitrev(i, L)] = inp[i]
            cases.append({"p": p, "L": L, "root": w, "in": inp, "in_br": in_br, "out": out, "out_br": out_br})
    return cases


def polymat_mul(p, m, k, n, a, sa, b, sb, deg):
    """c[e][t] for t <= deg of the product mod (x^pts - 1), pts = 2^ceil(log2(deg+1)),
    which for deg = sa+sb-2 is the plain product."""
    pts = 1
    while pts <= deg:
        pts <<= 1
    c = []
    for i in range(m):
        for j in range(n):
            acc = [0] * pts
            for q in range(k):
                pa = a[(i * k + q) * sa:(i * k + q + 1) * sa]
                pb = b[(q * n + j) * sb:(q * n + j + 1) * sb]
                for u, x in enumerate(pa):
                    if x:
                        for v, y in enumerate(pb):
                            acc[(u + v) % pts] += x * y
            c.extend(v % p for v in acc[:deg + 1])
    return c


def mul_cases(rng):
    shapes = [(1, 1, 1, 2, 2), (2, 3, 2, 5, 7), (3, 2, 4, 9, 4), (4, 4, 4, 16, 17), (2, 5, 3, 33, 31), (1, 3, 1, 1, 6)]
    cases = []
    for p in PRIMES:
        for (m, k, n, sa, sb) in shapes:
            a = [rng.randrange(p) for _ in range(m * k * sa)]
            b = [rng.randrange(p) for _ in range(k * n * sb)]
            deg = sa + sb - 2
            c = polymat_mul(p, m, k, n, a, sa, b, sb, deg)
            cases.append({"p": p, "m": m, "k": k, "n": n, "sa": sa, "sb": sb, "max_rowdeg": 0, "a": a, "b": b, "c": c})
        # extreme values: all p-1
        m, k, n, sa, sb = 3, 4, 2, 8, 8
        a = [p - 1] * (m * k * sa)
        b = [p - 1] * (k * n * sb)
        cases.append({"p": p, "m": m, "k": k, "n": n, "sa": sa, "sb": sb, "max_rowdeg": 0, "a": a, "b": b,
                      "c": polymat_mul(p, m, k, n, a, sa, b, sb, sa + sb - 2)})
        # max_rowdeg smaller than the full degree: pts=8, product wraps mod x^8-1
        m, k, n, sa, sb = 2, 2, 2, 6, 5
        a = [rng.randrange(p) for _ in range(m * k * sa)]
        b = [rng.randrange(p) for _ in range(k * n * sb)]
        cases.append({"p": p, "m": m, "k": k, "n": n, "sa": sa, "sb": sb, "max_rowdeg": 5, "a": a, "b": b,
                      "c": polymat_mul(p, m, k, n, a, sa, b, sb, 5)})
    return cases


def midproduct(p, m, k, n, a, sa, b, sb, sc, small_left, n0, n1):
    hdeg = n0 or sc
    deg = n1 or 2 * hdeg - 1
    pts = 1
    while pts < deg:
        pts <<= 1

    def padded(x, s, e):
        return x[e * s:(e + 1) * s] + [0] * (pts - s)

    def rev(v):  # the swaps of the first hdeg coefficients
        v = list(v)
        v[:hdeg] = v[:hdeg][::-1]
        return v
    c = []
    for i in range(m):
        for j in range(n):
            out = [0] * sc
            for q in range(k):
                pa = padded(a, sa, i * k + q)
                pb = padded(b, sb, q * n + j)
                if small_left:
                    pa = rev(pa)
                    for t in range(sc):
                        out[t] += sum(pa[(v - t) % pts] * pb[v] for v in range(pts) if pb[v])
                else:
                    pb = rev(pb)
                    for t in range(sc):
                        out[t] += sum(pa[u] * pb[(u - t) % pts] for u in range(pts) if pa[u])
            c.extend(x % p for x in out)
    return c


def midproduct_cases(rng):
    cases = []
    # (m, k, n, sa, sb, sc, small_left, n0, n1): default windows, explicit windows,
    # outputs longer than the middle product (wrapped terms), the big operand on the left
    shapes = [(2, 3, 2, 4, 7, 4, True, 0, 0), (1, 2, 3, 5, 9, 5, True, 0, 0), (2, 2, 2, 3, 8, 6, True, 4, 9),
              (3, 1, 2, 6, 3, 6, False, 0, 0), (2, 3, 1, 9, 5, 5, False, 5, 0), (2, 2, 2, 4, 12, 12, True, 4, 13),
              (1, 1, 1, 1, 1, 1, True, 0, 0), (2, 2, 2, 4, 8, 8, True, 4, 8), (2, 1, 2, 8, 3, 8, False, 3, 7)]
    for p in PRIMES:
        for (m, k, n, sa, sb, sc, sl, n0, n1) in shapes:
            a = [rng.randrange(p) for _ in range(m * k * sa)]
            b = [rng.randrange(p) for _ in range(k * n * sb)]
            deg = n1 or 2 * (n0 or sc) - 1
            case = {"p": p, "m": m, "k": k, "n": n, "sa": sa, "sb": sb, "sc": sc, "small_left": sl,
                    "n0": n0, "n1": n1, "a": a, "b": b, "status": 0,
                    "c": midproduct(p, m, k, n, a, sa, b, sb, sc, sl, n0, n1)}
            if deg <= 1:  # a 1-point transform: FFT's "k must be positive" (fft.h:136-139)
                case["status"], case["c"] = 3, None
            cases.append(case)
    return cases


def random_prime(bits, rng):
    while True:
        x = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        if sympy.isprime(x):
            return x


def to_limbs(x, nl):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(nl)]


def crt_cases(rng):
    P = random_prime(192, random.Random(0x4C494E424F58 + 3))
    nl = 6
    cases = []
    for (m, k, n, sa, sb) in [(2, 2, 2, 4, 3), (3, 2, 1, 7, 9), (1, 1, 1, 1, 1)]:
        a = [rng.randrange(P) for _ in range(m * k * sa)]
        b = [rng.randrange(P) for _ in range(k * n * sb)]
        c = polymat_mul(P, m, k, n, a, sa, b, sb, sa + sb - 2)
        cases.append({"m": m, "k": k, "n": n, "sa": sa, "sb": sb,
                      "a": [l for x in a for l in to_limbs(x, nl)],
                      "b": [l for x in b for l in to_limbs(x, nl)],
                      "c": [l for x in c for l in to_limbs(x, nl)]})
    return {"P_hex": hex(P), "P_limbs": to_limbs(P, nl), "nlimbs": nl, "cases": cases}


def fft_primes(pbound, prodbound, kmin):
    """Visit m*2^k+1 < pbound, k from bitsize(pbound-2)-1 down to kmin+1, m odd
    descending from the largest admissible value; stop once the product
    exceeds prodbound."""
    out, prod = [], 1
    for k in range((pbound - 2).bit_length() - 1, kmin, -1):
        B = pbound - 1
        m = (B >> k) if (B & ((1 << k) - 1)) else (B >> k) - 1
        if m % 2 == 0:
            m -= 1
        while m > 0:
            q = (m << k) + 1
            if sympy.isprime(q):
                out.append(q)
                prod *= q
                if prod > prodbound:
                    return out
            m -= 2
    return None


def prime_cases(crt):
    P = int(crt["P_hex"], 16)
    res = []
    for (k, minsz, lpts) in [(32, 1024, 11), (2, 4, 3), (3, 7, 4)]:
        bound = 2 * P * P * k * minsz
        res.append({"pbound": 1 << 31, "prodbound_hex": hex(bound), "kmin": lpts, "primes": fft_primes(1 << 31, bound, lpts)})
    res.append({"pbound": 1 << 24, "prodbound_hex": hex(1 << 100), "kmin": 10, "primes": fft_primes(1 << 24, 1 << 100, 10)})
    return res


def wrap_cases(rng):
    """max_rowdeg on the routes that do not use one FFT prime: word moduli that
    are not FFT primes for the length (the few-primes route of
    three-primes.inl:67-84) and multiword moduli (mul_crtla,
    multiprecision.inl:141-156): deg+1 coefficients of the product mod
    x^pts - 1, pts = 2^ceil(log2(deg+1)); also max_rowdeg above the full degree
    (zero coefficients past it) and constant operands."""
    shapes = [(2, 2, 2, 6, 5, 5), (2, 3, 1, 8, 8, 6), (3, 2, 2, 9, 7, 9), (2, 2, 2, 3, 3, 7), (2, 2, 3, 1, 1, 3),
              (1, 2, 1, 16, 16, 16)]
    moduli = [(97, 1), (1000000007, 1), (536870909, 1), (3221225473, 1), (4294967291, 1),
              (random_prime(61, random.Random(61)), 2), (random_prime(192, random.Random(192)), 6)]
    cases = []
    for (P, nl) in moduli:
        for (m, k, n, sa, sb, mr) in shapes:
            a = [rng.randrange(P) for _ in range(m * k * sa)]
            b = [rng.randrange(P) for _ in range(k * n * sb)]
            c = polymat_mul(P, m, k, n, a, sa, b, sb, mr)
            cases.append({"P_hex": hex(P), "P_limbs": to_limbs(P, nl), "nlimbs": nl, "m": m, "k": k, "n": n,
                          "sa": sa, "sb": sb, "max_rowdeg": mr,
                          "a": [l for x in a for l in to_limbs(x, nl)],
                          "b": [l for x in b for l in to_limbs(x, nl)],
                          "c": [l for x in c for l in to_limbs(x, nl)]})
    return cases


def main():
    rng = random.Random(20191)
    with open(os.path.join(HERE, "ntt_kat.json"), "w") as f:
        json.dump(ntt_kats(rng), f, separators=(",", ":"))
    with open(os.path.join(HERE, "polymat_mul.json"), "w") as f:
        json.dump(mul_cases(rng), f, separators=(",", ":"))
    crt = crt_cases(rng)
    with open(os.path.join(HERE, "crt_mul.json"), "w") as f:
        json.dump(crt, f, separators=(",", ":"))
    with open(os.path.join(HERE, "fft_primes.json"), "w") as f:
        json.dump(prime_cases(crt), f, separators=(",", ":"))
    with open(os.path.join(HERE, "midproduct.json"), "w") as f:
        json.dump(midproduct_cases(random.Random(20192)), f, separators=(",", ":"))
    with open(os.path.join(HERE, "polymat_mul_wrap.json"), "w") as f:
        json.dump(wrap_cases(random.Random(20193)), f, separators=(",", ":"))


if __name__ == "__main__":
    main()
