"""The oracle's order bases (OrderBasis::M_Basis / PM_Basis, order-basis.h:219-600)
pinned by properties that do not depend on how a basis is computed
(tests/basischeck.py), by known answers that follow from the reference's code,
and by a brute-force check that the rows generate every solution.

The reference's own test (tests/test-order-basis.C:19-101: m=8, n=4, d=80,
one FFT prime and one generic prime, M_Basis and PM_Basis, check_sigma) is
restated in test_reference_test_shape.  Which of the many order bases the
algorithm returns also depends on the row order of fflas-ffpack's PLUQ, which
is not vendored: the oracle takes the rank-profile order (pivot rows first,
both groups in increasing order), so that choice is not pinned beyond the
restatement."""
import numpy as np
import pytest

from oracle import oracle as orc
from basischeck import order_residual, det_is_monomial, generates_module

FFT_P = 998244353
GEN_P = 1000000007  # not an FFT prime for these lengths: products take the naive route


def rand_serie(rng, p, m, n, order):
    return rng.integers(0, p, m * n * order, dtype=np.uint64).astype(np.uint32)


def check_basis(fn, p, m, n, serie, order, shift0, points=(3, 1234567, 987654321)):
    st, sig, d, sh = fn(p, m, n, serie, order, order, shift0)
    assert st == 0
    so = order + 1
    res = order_residual(sig, so, serie, order, p, m, n, order)
    assert not res.any(), "sigma.serie != 0 mod x^order"
    assert np.asarray(sig).reshape(m * m, so)[:, d + 1:].sum() == 0
    assert np.asarray(sig).any()
    delta = int(sum(int(a) - int(b) for a, b in zip(sh, shift0)))
    assert det_is_monomial(sig, so, p, m, delta, [x % p for x in points])
    return sig, d, sh


@pytest.mark.parametrize("p", [FFT_P, GEN_P])
def test_reference_test_shape(p):
    """test-order-basis.C: 8 x 4 serie of degree 80, M_Basis and PM_Basis."""
    rng = np.random.default_rng(80)
    m, n, order = 8, 4, 80
    serie = rand_serie(rng, p, m, n, order)
    for fn in (orc.m_basis, orc.pm_basis):
        check_basis(fn, p, m, n, serie, order, [0] * m)


@pytest.mark.parametrize("m,n,order", [(4, 2, 33), (6, 3, 64), (5, 5, 40), (3, 4, 70), (1, 1, 45)])
def test_shapes_and_shifts(m, n, order):
    rng = np.random.default_rng(m * 1000 + n * 100 + order)
    serie = rand_serie(rng, FFT_P, m, n, order)
    shift = [int(x) for x in rng.integers(0, 5, m)]
    for fn in (orc.m_basis, orc.pm_basis):
        check_basis(fn, FFT_P, m, n, serie, order, shift)


def test_rank_deficient_serie():
    """Dependent columns and a zero row keep the properties."""
    rng = np.random.default_rng(5)
    m, n, order = 6, 3, 50
    F = rand_serie(rng, FFT_P, m, n, order).reshape(m, n, order).astype(np.uint64)
    F[:, 2] = (F[:, 0] * 3 + F[:, 1]) % FFT_P
    F[4] = 0
    serie = F.astype(np.uint32).reshape(-1)
    for fn in (orc.m_basis, orc.pm_basis):
        check_basis(fn, FFT_P, m, n, serie, order, [0] * m)


def test_known_answers():
    # zero serie: every step has rank 0, max_degree = min(order, max(.,0+1)) = 1
    # (order-basis.h:566), sigma stays a permutation of the identity (rows follow
    # the shift sorted by the first step's swaps) and is resized to 2
    m, n, order = 3, 2, 5
    st, sig, d, sh = orc.m_basis(FFT_P, m, n, np.zeros(m * n * order, np.uint32), order, order, [1, 0, 2])
    assert st == 0 and d == 1 and sh.tolist() == [0, 1, 2]
    S = sig.reshape(m, m, order + 1)
    assert (S[:, :, 0] == np.eye(m, dtype=np.uint32)[[1, 0, 2]]).all() and not S[:, :, 1:].any()
    # 1 x 1 serie with constant term 1: each step pivots, sigma = x^order
    for fn in (orc.m_basis, orc.pm_basis):
        order = 40
        serie = np.zeros(order, np.uint32)
        serie[0] = 1
        st, sig, d, sh = fn(FFT_P, 1, 1, serie, order, order, [7])
        assert st == 0 and d == order and sh.tolist() == [7 + order]
        assert sig[order] == 1 and not sig[:order].any()
    # serie e_1 (2 x 1): the basis is {x^order e_1, e_2} in some row order
    order = 6
    serie = np.zeros(2 * order, np.uint32)
    serie[0] = 1
    st, sig, d, sh = orc.m_basis(FFT_P, 2, 1, serie, order, order, [0, 0])
    S = sig.reshape(2, 2, order + 1)
    rows = sorted((tuple(S[i, 0]), tuple(S[i, 1])) for i in range(2))
    xo = tuple([0] * order + [1])
    one = tuple([1] + [0] * order)
    zero = tuple([0] * (order + 1))
    assert rows == sorted([(xo, zero), (zero, one)]) and d == order


@pytest.mark.parametrize("p,m,n,order,vdeg", [(2, 2, 1, 3, 3), (3, 2, 1, 3, 3), (2, 3, 1, 2, 2), (2, 3, 2, 2, 2),
                                              (5, 2, 1, 2, 2)])
def test_rows_generate_every_solution(p, m, n, order, vdeg):
    """Brute force over a tiny field: every solution is a polynomial combination of sigma's rows."""
    rng = np.random.default_rng(p * 100 + m * 10 + order)
    for trial in range(3):
        serie = rand_serie(rng, p, m, n, order)
        st, sig, d, sh = orc.m_basis(p, m, n, serie, order, order, [0] * m)
        assert st == 0
        assert not order_residual(sig, order + 1, serie, order, p, m, n, order).any()
        delta = int(sum(int(x) for x in sh))
        ok, found = generates_module(sig, order + 1, serie, order, p, m, n, order, delta, vdeg)
        assert ok and found > 1


def test_checks_reject_planted_mistakes():
    """Plausible slips in a basis are each caught by one of the checks."""
    rng = np.random.default_rng(11)
    m, n, order, p = 4, 2, 40, FFT_P
    serie = rand_serie(rng, p, m, n, order)
    st, sig, d, sh = orc.pm_basis(p, m, n, serie, order, order, [0] * m)
    so = order + 1
    S = sig.reshape(m * m, so)
    delta = int(sum(int(x) for x in sh))
    assert det_is_monomial(S, so, p, m, delta, [3, 5, 7])
    # a wrong coefficient breaks the order property
    bad = S.copy()
    bad[1, 0] = (int(bad[1, 0]) + 1) % p
    assert order_residual(bad, so, serie, order, p, m, n, order).any()
    # a row multiplied by x still annihilates, but det gains a factor x
    bad = S.reshape(m, m, so).copy()
    assert not bad[0, :, so - 1].any()
    bad[0, :, 1:] = bad[0, :, :-1].copy()
    bad[0, :, 0] = 0
    bad = bad.reshape(m * m, so)
    assert not order_residual(bad, so, serie, order, p, m, n, order).any()
    assert not det_is_monomial(bad, so, p, m, delta, [3, 5, 7])
    # a row added to another keeps both checks: the basis is one of many, which is
    # why the device result is compared with the oracle word for word
    alt = S.reshape(m, m, so).astype(np.uint64)
    alt[0] = (alt[0] + alt[1]) % p
    alt = alt.astype(np.uint32).reshape(m * m, so)
    assert not order_residual(alt, so, serie, order, p, m, n, order).any()
    assert det_is_monomial(alt, so, p, m, delta, [3, 5, 7])
    # a dropped pivot (row zeroed) makes det vanish
    bad = S.reshape(m, m, so).copy()
    bad[2] = 0
    assert not det_is_monomial(bad.reshape(m * m, so), so, p, m, delta, [3, 5, 7])


def test_pm_equals_m_for_small_order():
    """Below the threshold PM_Basis is M_Basis (order-basis.h:232)."""
    rng = np.random.default_rng(9)
    serie = rand_serie(rng, FFT_P, 5, 2, 32)
    a = orc.m_basis(FFT_P, 5, 2, serie, 32, 32, [0, 1, 2, 3, 4])
    b = orc.pm_basis(FFT_P, 5, 2, serie, 32, 32, [0, 1, 2, 3, 4])
    assert a[0] == b[0] == 0 and a[2] == b[2]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3])
