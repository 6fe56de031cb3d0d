"""Exact checks for polynomial-matrix products, independent of how they were computed.

* eval_polymat: Horner evaluation of every entry at a point x (mod p), vectorised.
* matmul_mod:   exact m x k . k x n matrix product mod p (< 2^32) in numpy uint64,
                splitting the right operand in 16-bit halves so no sum overflows.
* check_by_evaluation: Schwartz-Zippel check C(x) == A(x) B(x) at random x,
                which catches any wrong coefficient with probability >= 1 - deg/p
                per point (far stronger than the reference's x=1 check,
                matpoly-mult-fft.h:122-163, which a coefficient permutation passes).
"""
import numpy as np


def eval_polymat(coeffs, x, p):
    """coeffs: (entries, size) uint32 array; returns (entries,) uint64 values mod p."""
    coeffs = np.asarray(coeffs, dtype=np.uint64)
    acc = np.zeros(coeffs.shape[0], dtype=np.uint64)
    x = np.uint64(x)
    pp = np.uint64(p)
    for t in range(coeffs.shape[1] - 1, -1, -1):
        acc = (acc * x + coeffs[:, t]) % pp
    return acc


def matmul_mod(A, B, p):
    A = np.asarray(A, dtype=np.uint64) % np.uint64(p)
    B = np.asarray(B, dtype=np.uint64) % np.uint64(p)
    assert A.shape[1] <= 1 << 16
    lo = B & np.uint64(0xFFFF)
    hi = B >> np.uint64(16)
    pp = np.uint64(p)
    # A < 2^32, lo/hi < 2^16 -> products < 2^48, k <= 2^16 terms -> < 2^64
    r_lo = (A @ lo) % pp
    r_hi = (A @ hi) % pp
    return (r_hi * np.uint64(1 << 16) % pp + r_lo) % pp


def check_by_evaluation(a, b, c, p, m, k, n, sa, sb, sc, xs):
    """a: (m*k, sa), b: (k*n, sb), c: (m*n, sc).  True if C(x)==A(x)B(x) for all x in xs."""
    a = np.asarray(a).reshape(m * k, sa)
    b = np.asarray(b).reshape(k * n, sb)
    c = np.asarray(c).reshape(m * n, sc)
    for x in xs:
        Ax = eval_polymat(a, x, p).reshape(m, k)
        Bx = eval_polymat(b, x, p).reshape(k, n)
        Cx = eval_polymat(c, x, p).reshape(m, n)
        if not np.array_equal(matmul_mod(Ax, Bx, p), Cx):
            return False
    return True
