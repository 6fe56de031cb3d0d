"""Checks of order bases that do not depend on how a basis was computed.

An order basis of a serie F (m x n) at `order` is an m x m polynomial matrix
sigma whose rows generate every v with v.F = 0 mod x^order
(order-basis.h:219-326,346-600 compute one; tests/test-order-basis.C:19-45
checks the first half of that statement).

* order_residual:   coefficients 0..order-1 of sigma.F, exactly (numpy).
* det_is_monomial:  every step of M_Basis multiplies sigma by a permutation, a
                    unit triangular matrix and diag(x on the pivot rows), so
                    det(sigma) = c x^delta with delta = the total increase of
                    the shift; checked at random points.
* generates_module: over a tiny field, every v of bounded degree with
                    v.F = 0 mod x^order satisfies v.adj(sigma) = 0 mod x^delta,
                    i.e. v = u.sigma with u polynomial (brute force).
"""
import itertools

import numpy as np

from polycheck import eval_polymat, matmul_mod


def order_residual(sigma, ssig, serie, sser, p, m, n, order):
    """sigma: (m*m, ssig) coefficients, serie: (m*n, sser); returns (order, m, n)."""
    S = np.asarray(sigma, dtype=np.uint64).reshape(m, m, ssig)
    F = np.asarray(serie, dtype=np.uint64).reshape(m, n, sser)
    out = np.zeros((order, m, n), dtype=np.uint64)
    for t in range(order):
        acc = np.zeros((m, n), dtype=np.uint64)
        for l in range(min(t, ssig - 1) + 1):
            if t - l < sser:
                acc = (acc + matmul_mod(S[:, :, l], F[:, :, t - l], p)) % np.uint64(p)
        out[t] = acc
    return out


def det_mod(M, p):
    M = [[int(x) % p for x in row] for row in M]
    n, det = len(M), 1
    for c in range(n):
        r = next((i for i in range(c, n) if M[i][c]), None)
        if r is None:
            return 0
        if r != c:
            M[c], M[r] = M[r], M[c]
            det = -det
        det = det * M[c][c] % p
        inv = pow(M[c][c], p - 2, p)
        for i in range(c + 1, n):
            f = M[i][c] * inv % p
            if f:
                M[i] = [(a - f * b) % p for a, b in zip(M[i], M[c])]
    return det % p


def det_is_monomial(sigma, ssig, p, m, delta, points):
    """True if det(sigma(x)) x^-delta is the same nonzero constant at every point."""
    vals = []
    coeffs = np.asarray(sigma, dtype=np.uint32).reshape(m * m, ssig)
    for x in points:
        Mx = eval_polymat(coeffs, x, p).reshape(m, m)
        vals.append(det_mod(Mx, p) * pow(pow(x, delta, p), p - 2, p) % p)
    return vals[0] != 0 and all(v == vals[0] for v in vals)


# ---- tiny-field polynomial arithmetic (lists of ints, low degree first) ----
def _trim(a):
    while len(a) > 1 and a[-1] == 0:
        a = a[:-1]
    return a


def padd(a, b, p):
    r = [0] * max(len(a), len(b))
    for i, x in enumerate(a):
        r[i] = x
    for i, x in enumerate(b):
        r[i] = (r[i] + x) % p
    return _trim(r)


def pmul(a, b, p):
    r = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                r[i + j] = (r[i + j] + x * y) % p
    return _trim(r)


def pneg(a, p):
    return [(-x) % p for x in a]


def poly_det(M, p):
    """Determinant of a small polynomial matrix by cofactor expansion."""
    n = len(M)
    if n == 1:
        return M[0][0]
    acc = [0]
    for j in range(n):
        minor = [row[:j] + row[j + 1:] for row in M[1:]]
        term = pmul(M[0][j], poly_det(minor, p), p)
        acc = padd(acc, term if j % 2 == 0 else pneg(term, p), p)
    return acc


def adjugate(M, p):
    n = len(M)
    if n == 1:
        return [[[1]]]
    adj = [[None] * n for _ in range(n)]
    for i in range(n):
        for j in range(n):
            minor = [row[:j] + row[j + 1:] for k, row in enumerate(M) if k != i]
            c = poly_det(minor, p)
            adj[j][i] = c if (i + j) % 2 == 0 else pneg(c, p)
    return adj


def generates_module(sigma, ssig, serie, sser, p, m, n, order, delta, vdeg):
    """Brute force over all v in F_p[x]^(1 x m) of degree <= vdeg."""
    S = np.asarray(sigma).reshape(m, m, ssig)
    F = np.asarray(serie).reshape(m, n, sser)
    Sp = [[_trim([int(c) for c in S[i, j]]) for j in range(m)] for i in range(m)]
    Fp = [[[int(c) for c in F[i, j][:order]] for j in range(n)] for i in range(m)]
    adj = adjugate(Sp, p)
    found = 0
    for flat in itertools.product(range(p), repeat=m * (vdeg + 1)):
        v = [list(flat[i * (vdeg + 1):(i + 1) * (vdeg + 1)]) for i in range(m)]
        ok = True
        for j in range(n):
            acc = [0]
            for i in range(m):
                acc = padd(acc, pmul(v[i], Fp[i][j], p), p)
            if any(acc[t] for t in range(min(order, len(acc)))):
                ok = False
                break
        if not ok:
            continue
        found += 1
        for j in range(m):
            acc = [0]
            for i in range(m):
                acc = padd(acc, pmul(v[i], adj[i][j], p), p)
            if any(acc[t] for t in range(min(delta, len(acc)))):
                return False, found
    return True, found
