"""Seeded random sweep over the entry points, each case against the oracle.

Shapes, degrees, primes, layouts and pointwise forms are drawn from fixed seeds
so the sweep is reproducible; every drawn case prints its parameters before it
runs.  Products: any prime below 2^32 (FFT primes, word primes that take the CRT
route, primes of too small 2-valuation), both forms, all layouts.  Middle
products: both operand orders and explicit windows.  Order bases: M_Basis and
PM_Basis with random shifts.  Layout conversions: random layouts, windows,
offsets and views on either side.  Integer products (1..32-word operands,
random signed bit sizes, forms, spare output words, cyclic max_rowdeg) and
multiword moduli of 2..64 limbs against Python integers.
"""
import sys

import numpy as np
import pytest

import linbox_amd as lbx
from linbox_amd import PolynomialMatrix, PolynomialMatrixFFTMulDomain, OrderBasis
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
FFT_PRIMES = [998244353, 2013265921, 469762049, 167772161, 754974721, 1811939329]
WORD_PRIMES = [1000000007, 4294967291, 3221225473, 65537, 97, 2]
LAYOUTS = ["polfirst", "matfirst", "matrowfirst"]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = lbx.Context(0)
    yield c
    c.close()


def rand(rng, p, n):
    return rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32)


def ref_product(p, m, k, n, a, sa, b, sb):
    D = sa + sb - 2
    pts = 1
    while pts < D + 1:
        pts <<= 1
    if sa + sb > 2 and p < (1 << 31) and (p - 1) % pts == 0:
        st, c = orc.mul_fft_prime(p, m, k, n, a, sa, b, sb)
        assert st == 0
        return c
    if m * k * n * sa * sb <= 1 << 22:
        return orc.mul_naive(p, m, k, n, a, sa, b, sb)
    st, c, _ = orc.mul_crt(np.array([p], dtype=np.uint32), m, k, n, a, sa, b, sb)  # exact product mod p
    assert st == 0
    return c


def test_random_products(ctx):
    rng = np.random.default_rng(20261202)
    for case in range(120):
        p = int(rng.choice(FFT_PRIMES + WORD_PRIMES))
        m, k, n = (int(x) for x in rng.integers(1, 40, 3))
        sa = int(np.exp(rng.uniform(0, np.log(3000))))
        sb = max(1, int(sa * rng.uniform(0.2, 1.5)))
        method = int(rng.choice([lbx.METHOD_AUTO, lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA]))
        lay = [str(rng.choice(LAYOUTS)) for _ in range(3)]
        say("product", case, p, m, k, n, sa, sb, method, lay)
        a = rand(rng, p, m * k * sa)
        b = rand(rng, p, k * n * sb)
        A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k, lay[0])
        B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n, lay[1])
        C = PolynomialMatrix(m, n, 1, lay[2])
        PolynomialMatrixFFTMulDomain(p, ctx, method).mul(C, A, B)
        ref = ref_product(p, m, k, n, a, sa, b, sb)
        assert np.array_equal(C.to_polfirst().reshape(-1), ref), case


def test_random_midproducts(ctx):
    rng = np.random.default_rng(20261203)
    for case in range(60):
        p = int(rng.choice(FFT_PRIMES[:4] + [1000000007, 4294967291]))
        m, k, n = (int(x) for x in rng.integers(1, 20, 3))
        hdeg = int(rng.integers(2, 600))
        deg = hdeg + int(rng.integers(1, 600))
        small_left = bool(rng.integers(0, 2))
        ss = int(rng.integers(1, hdeg + 1))
        sbig = int(rng.integers(1, deg + 1))
        sc = deg - hdeg + 1
        sa, sb = (ss, sbig) if small_left else (sbig, ss)
        say("midproduct", case, p, m, k, n, hdeg, deg, small_left, sa, sb)
        a = rand(rng, p, m * k * sa)
        b = rand(rng, p, k * n * sb)
        A = PolynomialMatrix.from_polfirst(a.reshape(m * k, sa), m, k)
        B = PolynomialMatrix.from_polfirst(b.reshape(k * n, sb), k, n)
        C = PolynomialMatrix(m, n, sc)
        PolynomialMatrixFFTMulDomain(p, ctx).midproduct(C, A, B, small_left, hdeg, deg)
        # coefficients hdeg-1 .. deg-1 of a.b
        full = ref_product(p, m, k, n, a, sa, b, sb).reshape(m * n, sa + sb - 1)
        want = np.zeros((m * n, sc), dtype=np.uint32)
        top = min(sa + sb - 1, deg)
        want[:, :max(0, top - hdeg + 1)] = full[:, hdeg - 1:top]
        assert np.array_equal(C.to_polfirst(), want), case


def test_random_order_bases(ctx):
    rng = np.random.default_rng(20261204)
    for case in range(32):
        p = int(rng.choice([998244353, 2013265921, 1000000007, 469762049, 4294967291]))
        m = int(rng.integers(1, 24))
        n = int(rng.integers(1, min(m, 12) + 1))
        order = int(rng.integers(1, 160))
        shift = [int(x) for x in rng.integers(0, 6, m)]
        which = "pm" if rng.integers(0, 2) else "m"
        say("order basis", case, which, p, m, n, order, shift)
        serie = rand(rng, p, m * n * order)
        F = PolynomialMatrix.from_polfirst(serie.reshape(m * n, order), m, n)
        S = PolynomialMatrix(m, m, order + 1)
        sh = list(shift)
        ob = OrderBasis(p, ctx)
        d = (ob.PM_Basis if which == "pm" else ob.M_Basis)(S, F, order, sh)
        st, ref, dref, shref = (orc.pm_basis if which == "pm" else orc.m_basis)(p, m, n, serie, order, order, shift)
        assert st == 0 and d == dref and sh == [int(x) for x in shref], case
        assert np.array_equal(S.to_polfirst(), ref.reshape(m * m, order + 1)[:, :d + 1]), case


def test_random_copies(ctx):
    rng = np.random.default_rng(20261205)
    for case in range(100):
        r, c = (int(x) for x in rng.integers(1, 24, 2))
        if rng.integers(0, 3) == 0:  # entry counts that are multiples of four (the 16-byte kernel)
            r = 4 * int(rng.integers(1, 8))
        ss = int(rng.integers(1, 300))
        beg = int(rng.integers(0, ss))
        end = int(rng.integers(beg, ss))
        cnt = end - beg + 1
        ds = cnt + int(rng.integers(0, 40))
        start = int(rng.integers(0, ds - cnt + 1))
        ls, ld = str(rng.choice(LAYOUTS)), str(rng.choice(LAYOUTS))
        sview, dview = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
        say("copy", case, r, c, ss, ds, beg, end, start, ls, ld, sview, dview)
        s0 = int(rng.integers(0, 9)) if sview else 0  # views: extra coefficients before the window
        d0 = int(rng.integers(0, 9)) if dview else 0
        sdata = rng.integers(0, 1 << 32, (r * c, ss + s0 + 3), dtype=np.uint64).astype(np.uint32)
        S = PolynomialMatrix.from_polfirst(sdata, r, c, ls)
        src = S.at(s0, s0 + ss - 1) if sview else S
        if not sview:
            S = PolynomialMatrix.from_polfirst(sdata[:, :ss], r, c, ls)
            src = S
        ddata = rng.integers(0, 1 << 32, (r * c, ds + d0 + 2), dtype=np.uint64).astype(np.uint32)
        D = PolynomialMatrix.from_polfirst(ddata if dview else ddata[:, :ds], r, c, ld)
        dst = D.at(d0, d0 + ds - 1) if dview else D
        dst.copy(src, beg, end, start, ctx=ctx)
        want = (ddata if dview else ddata[:, :ds]).copy()
        want[:, d0 + start:d0 + start + cnt] = sdata[:, s0 + beg:s0 + end + 1]
        assert np.array_equal(D.to_polfirst(), want), case


def naive_int(A, B, m, k, n, sa, sb):
    """Product of integer polynomial matrices (polfirst lists of Python ints)."""
    sc = sa + sb - 1
    C = [0] * (m * n * sc)
    for i in range(m):
        for j in range(n):
            for q in range(k):
                a = A[(i * k + q) * sa:(i * k + q + 1) * sa]
                b = B[(q * n + j) * sb:(q * n + j + 1) * sb]
                base = (i * n + j) * sc
                for u, x in enumerate(a):
                    if x:
                        for v, y in enumerate(b):
                            C[base + u + v] += x * y
    return C


def test_random_integer_products(ctx):
    """lbx_polymat_mul_zz on random shapes, word counts (1..32 in), signed bit sizes,
    forms, spare output words and cyclic max_rowdeg wraps, against Python integers."""
    import random
    from linbox_amd.polymat import mul_zz, to_limbs, from_limbs
    rnd = random.Random(20261203)
    methods = [lbx.METHOD_AUTO, lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA]
    for case in range(60):
        m, k, n = rnd.randint(1, 5), rnd.randint(1, 6), rnd.randint(1, 5)
        sa, sb = rnd.randint(1, 40), rnd.randint(1, 40)
        if sa + sb <= 2:
            sb = 2
        wi = rnd.choice([1, 2, 3, 5, 8, 16, 32])
        bits = rnd.randint(1, 32 * wi - 1)
        A = [rnd.randrange(-(1 << bits) + 1, 1 << bits) for _ in range(m * k * sa)]
        B = [rnd.randrange(-(1 << bits) + 1, 1 << bits) for _ in range(k * n * sb)]
        la = max(abs(x).bit_length() for x in A)
        lb = max(abs(x).bit_length() for x in B)
        need = ((2 * (1 << la) * (1 << lb) * k * min(sa, sb)).bit_length() + 31) // 32
        wo = min(80, need + rnd.choice([0, 0, 1, 3]))
        full = naive_int(A, B, m, k, n, sa, sb)
        # a cyclic max_rowdeg product needs operands no longer than its transform
        # (the reference pads each operand into 2^ceil(log2(max_rowdeg+1)) words)
        mr = 0
        lo = max(sa, sb) - 1
        if rnd.random() < 0.25 and 1 <= lo <= sa + sb - 3:
            mr = rnd.randint(lo, sa + sb - 3)
        method = rnd.choice(methods)
        say("random zz", case, m, k, n, sa, sb, wi, bits, wo, mr, method)
        a = PolynomialMatrix(m, k, sa, data=to_limbs(A, wi), nlimbs=wi)
        b = PolynomialMatrix(k, n, sb, data=to_limbs(B, wi), nlimbs=wi)
        sc = mr + 1 if mr else sa + sb - 1
        c = PolynomialMatrix(m, n, sc, nlimbs=wo)
        mul_zz(c, a, b, ctx=ctx, method=method, max_rowdeg=mr)
        got = from_limbs(c.data, wo)
        if not mr:
            assert got == full, case
        else:
            pts = 1
            while pts < sc:
                pts <<= 1
            want = []
            for e in range(m * n):
                row = [0] * max(pts, sc)
                for t in range(sa + sb - 1):
                    row[t % pts] += full[e * (sa + sb - 1) + t]
                want += row[:sc]
            assert got == want, case
    # operands longer than the transform a max_rowdeg implies are refused
    a = PolynomialMatrix(2, 2, 9, data=to_limbs([1] * 36, 1), nlimbs=1)
    b = PolynomialMatrix(2, 2, 3, data=to_limbs([1] * 12, 1), nlimbs=1)
    c = PolynomialMatrix(2, 2, 3, nlimbs=2)
    with pytest.raises(lbx.LinboxError) as e:
        mul_zz(c, a, b, ctx=ctx, max_rowdeg=2)
    assert e.value.status == 5


def test_random_wide_moduli(ctx):
    """lbx_polymat_mul_mp for random moduli of 2..64 limbs (odd or even, any top word)
    on small random shapes, against the product reduced with Python integers."""
    import random
    rnd = random.Random(20261204)
    for case in range(40):
        nl = rnd.randint(2, 64)
        P = rnd.randrange(1 << (32 * nl - 32), 1 << (32 * nl))
        if P < 3:
            P = 3
        m, k, n = rnd.randint(1, 4), rnd.randint(1, 5), rnd.randint(1, 4)
        sa, sb = rnd.randint(1, 24), rnd.randint(2, 24)
        A = [rnd.randrange(P) for _ in range(m * k * sa)]
        B = [rnd.randrange(P) for _ in range(k * n * sb)]
        method = rnd.choice([lbx.METHOD_AUTO, lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA])
        say("random mp", case, nl, P.bit_length(), m, k, n, sa, sb, method)

        def words(vals):
            return np.array([(v >> (32 * i)) & 0xFFFFFFFF for v in vals for i in range(nl)], dtype=np.uint32)
        a = PolynomialMatrix(m, k, sa, data=words(A), nlimbs=nl)
        b = PolynomialMatrix(k, n, sb, data=words(B), nlimbs=nl)
        c = PolynomialMatrix(m, n, sa + sb - 1, nlimbs=nl)
        lbx.mul_mp(P, c, a, b, ctx=ctx, method=method)
        w = c.data.reshape(-1, nl)
        got = [sum(int(w[r, i]) << (32 * i) for i in range(nl)) for r in range(w.shape[0])]
        want = [x % P for x in naive_int(A, B, m, k, n, sa, sb)]
        assert got == want, case


def test_narrow_moduli_crt_widths(ctx):
    """lbx_polymat_mul_mp for moduli of one to four words, the sizes whose CRT
    runs in narrow register arithmetic.  With the bound 2 P^2 k min(sa,sb) of
    these shapes, P = 3, 2^32 - 5 and 2^32 + 15 need at most 4 primes (2-word
    kernel), 2^64 - 59 and the three-word P 5 to 8 (4-word kernel), the
    four-word P 9 (8-word kernel).  Word extremes of P, coefficients at P - 1
    (the largest Garner sums) and random ones, both pointwise forms, against
    Python integers."""
    import random
    rnd = random.Random(20261205)
    moduli = [(3, 1), ((1 << 32) - 5, 1), ((1 << 32) + 15, 2), ((1 << 64) - 59, 2),
              (rnd.randrange(1 << 80, 1 << 96), 3), (rnd.randrange(1 << 120, 1 << 128), 4)]
    for P, nl in moduli:
        for extreme in (False, True):
            for method in (lbx.METHOD_PLAIN_U64, lbx.METHOD_INT8_MFMA):
                m, k, n, sa, sb = 3, 17, 2, 40, 33
                A = [P - 1 if extreme else rnd.randrange(P) for _ in range(m * k * sa)]
                B = [P - 1 if extreme else rnd.randrange(P) for _ in range(k * n * sb)]
                say("narrow mp", nl, P.bit_length(), extreme, method)

                def words(vals):
                    return np.array([(v >> (32 * i)) & 0xFFFFFFFF for v in vals for i in range(nl)],
                                    dtype=np.uint32)
                a = PolynomialMatrix(m, k, sa, data=words(A), nlimbs=nl)
                b = PolynomialMatrix(k, n, sb, data=words(B), nlimbs=nl)
                c = PolynomialMatrix(m, n, sa + sb - 1, nlimbs=nl)
                lbx.mul_mp(P, c, a, b, ctx=ctx, method=method)
                w = c.data.reshape(-1, nl)
                got = [sum(int(w[r, i]) << (32 * i) for i in range(nl)) for r in range(w.shape[0])]
                want = [x % P for x in naive_int(A, B, m, k, n, sa, sb)]
                assert got == want, (P, nl, extreme, method)
