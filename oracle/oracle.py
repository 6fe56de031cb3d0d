"""ctypes wrapper over oracle/_build/liblbx_oracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblbx_oracle.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sz, u32, u64, i32 = ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.or_fill_uniform.argtypes = [u64, u32, _u32p, sz]
        L.or_powmod.argtypes = [u32, u64, u32]
        L.or_powmod.restype = u32
        L.or_invmod.argtypes = [u32, u32]
        L.or_invmod.restype = u32
        L.or_is_prime_u64.argtypes = [u64]
        L.or_compute_primitive_root.argtypes = [u32, ctypes.c_uint, _u32p]
        L.or_is_primitive_root.argtypes = [u32, u32, ctypes.c_uint]
        L.or_init_powers.argtypes = [u32, ctypes.c_uint, u32, _u32p, _u32p]
        L.or_fft_dif.argtypes = [u32, ctypes.c_uint, _u32p, _u32p]
        L.or_fft_dit.argtypes = [u32, ctypes.c_uint, _u32p, _u32p]
        L.or_fft_direct.argtypes = [u32, ctypes.c_uint, u32, _u32p, _u32p]
        L.or_fft_inverse.argtypes = [u32, ctypes.c_uint, u32, _u32p, _u32p]
        L.or_fft_invroot.argtypes = [u32, ctypes.c_uint, u32]
        L.or_fft_invroot.restype = u32
        L.or_polymat_mul_fft_prime.argtypes = [u32, sz, sz, sz, _u32p, sz, _u32p, sz, _u32p, sz]
        L.or_polymat_mul_naive.argtypes = [u32, sz, sz, sz, _u32p, sz, _u32p, sz, _u32p]
        L.or_generate_fft_primes.argtypes = [u64, _u32p, i32, ctypes.c_uint, _u32p, i32]
        L.or_polymat_mul_crt.argtypes = [_u32p, i32, sz, sz, sz, _u32p, sz, _u32p, sz, _u32p, _u32p,
                                         ctypes.POINTER(i32)]
        L.or_polymat_mul_three_primes.argtypes = [u32, sz, sz, sz, _u32p, sz, _u32p, sz, _u32p, _u32p,
                                                  ctypes.POINTER(ctypes.c_int)]
        L.or_polymat_midproduct_fft_prime.argtypes = [u32, sz, sz, sz, _u32p, sz, _u32p, sz, _u32p, sz, i32, sz, sz]
        L.or_polymat_midproduct_naive.argtypes = [u32, sz, sz, sz, _u32p, sz, _u32p, sz, _u32p, sz, i32, sz, sz]
        u64p = ctypes.POINTER(u64)
        L.or_m_basis.argtypes = [u32, sz, sz, _u32p, sz, sz, u64p, _u32p, ctypes.POINTER(sz)]
        L.or_pm_basis.argtypes = [u32, sz, sz, _u32p, sz, sz, u64p, _u32p, ctypes.POINTER(sz)]
        L.or_stage_forward.argtypes = [u32, ctypes.c_uint, _u32p, _u32p, sz]
        L.or_stage_inverse.argtypes = [u32, ctypes.c_uint, _u32p, _u32p, sz]
        L.or_stage_pointwise.argtypes = [u32, sz, sz, sz, sz, _u32p, _u32p, _u32p, sz, sz]
        L.or_stage_forward_mt.argtypes = [u32, ctypes.c_uint, _u32p, _u32p, sz, i32]
        L.or_stage_inverse_mt.argtypes = [u32, ctypes.c_uint, _u32p, _u32p, sz, i32]
        L.or_stage_pointwise_mt.argtypes = [u32, sz, sz, sz, sz, _u32p, _u32p, _u32p, sz, sz, i32]
        _lib = L
    return _lib


def _p(arr):
    assert arr.dtype == np.uint32 and arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(_u32p)


def fill_uniform(seed, p, n):
    out = np.empty(n, dtype=np.uint32)
    lib().or_fill_uniform(seed, p, _p(out), n)
    return out


def primitive_root(p, k):
    w = ctypes.c_uint32(0)
    st = lib().or_compute_primitive_root(p, k, ctypes.byref(w))
    if st:
        raise ValueError("oracle status %d" % st)
    return w.value


def init_powers(p, L, w):
    n = 1 << L
    pw = np.empty(n - 1, dtype=np.uint32)
    pwp = np.empty(n - 1, dtype=np.uint32)
    lib().or_init_powers(p, L, w, _p(pw), _p(pwp))
    return pw, pwp


def fft_direct(p, L, coeffs, root=0):
    c = np.ascontiguousarray(coeffs, dtype=np.uint32).copy()
    w = ctypes.c_uint32(0)
    st = lib().or_fft_direct(p, L, root, _p(c), ctypes.byref(w))
    if st:
        raise ValueError("oracle status %d" % st)
    return c, w.value


def fft_inverse(p, L, coeffs, root=0):
    c = np.ascontiguousarray(coeffs, dtype=np.uint32).copy()
    w = ctypes.c_uint32(0)
    st = lib().or_fft_inverse(p, L, root, _p(c), ctypes.byref(w))
    if st:
        raise ValueError("oracle status %d" % st)
    return c, w.value


def invroot(p, L, w):
    return lib().or_fft_invroot(p, L, w)


def mul_fft_prime(p, m, k, n, a, sa, b, sb, max_rowdeg=0):
    """Polfirst product; returns (status, c) with c of shape (m*n, deg+1)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    deg = max_rowdeg if max_rowdeg else sa + sb - 2
    c = np.zeros(m * n * (deg + 1), dtype=np.uint32)
    st = lib().or_polymat_mul_fft_prime(p, m, k, n, _p(a), sa, _p(b), sb, _p(c), max_rowdeg)
    return st, c


def mul_naive(p, m, k, n, a, sa, b, sb):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    c = np.zeros(m * n * (sa + sb - 1), dtype=np.uint32)
    st = lib().or_polymat_mul_naive(p, m, k, n, _p(a), sa, _p(b), sb, _p(c))
    assert st == 0
    return c


def midproduct_fft_prime(p, m, k, n, a, sa, b, sb, sc, small_left=True, n0=0, n1=0):
    """FFT-prime midproduct (polfirst); returns (status, c) with c of shape (m*n*sc,)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    c = np.zeros(m * n * sc, dtype=np.uint32)
    st = lib().or_polymat_midproduct_fft_prime(p, m, k, n, _p(a), sa, _p(b), sb, _p(c), sc, 1 if small_left else 0,
                                               n0, n1)
    return st, c


def midproduct_naive(p, m, k, n, a, sa, b, sb, c, small_left=True, n0=0, n1=0):
    """Naive midproduct into a copy of c (coefficients it does not compute keep c's values)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    c = np.array(c, dtype=np.uint32, copy=True)
    sc = c.size // (m * n)
    st = lib().or_polymat_midproduct_naive(p, m, k, n, _p(a), sa, _p(b), sb, _p(c), sc, 1 if small_left else 0,
                                           n0, n1)
    return st, c


def _order_basis(fn, p, m, n, serie, ssz, order, shift):
    serie = np.ascontiguousarray(serie, dtype=np.uint32)
    assert serie.size == m * n * ssz
    sh = np.ascontiguousarray(shift, dtype=np.uint64).copy()
    sigma = np.zeros(m * m * (order + 1), dtype=np.uint32)
    deg = ctypes.c_size_t(0)
    st = fn(p, m, n, _p(serie), ssz, order, sh.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _p(sigma),
            ctypes.byref(deg))
    return st, sigma, int(deg.value), sh


def m_basis(p, m, n, serie, ssz, order, shift):
    """OrderBasis::M_Basis: (status, sigma polfirst m*m*(order+1), degree, shift out)."""
    return _order_basis(lib().or_m_basis, p, m, n, serie, ssz, order, shift)


def pm_basis(p, m, n, serie, ssz, order, shift):
    """OrderBasis::PM_Basis: (status, sigma polfirst m*m*(order+1), degree, shift out)."""
    return _order_basis(lib().or_pm_basis, p, m, n, serie, ssz, order, shift)


def generate_fft_primes(pbound, prodbound, kmin, max_primes=64):
    nl = max(1, (prodbound.bit_length() + 31) // 32)
    limbs = np.array([(prodbound >> (32 * i)) & 0xFFFFFFFF for i in range(nl)], dtype=np.uint32)
    out = np.zeros(max_primes, dtype=np.uint32)
    r = lib().or_generate_fft_primes(pbound, _p(limbs), nl, kmin, _p(out), max_primes)
    return None if r < 0 else [int(x) for x in out[:r]]


def mul_crt(P_limbs, m, k, n, a, sa, b, sb):
    P_limbs = np.ascontiguousarray(P_limbs, dtype=np.uint32)
    nl = len(P_limbs)
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    c = np.zeros(m * n * (sa + sb - 1) * nl, dtype=np.uint32)
    primes = np.zeros(64, dtype=np.uint32)
    npr = ctypes.c_int(0)
    st = lib().or_polymat_mul_crt(_p(P_limbs), nl, m, k, n, _p(a), sa, _p(b), sb, _p(c), _p(primes),
                                  ctypes.byref(npr))
    return st, c, [int(x) for x in primes[:npr.value]]


def mul_three_primes(p, m, k, n, a, sa, b, sb):
    """PolynomialMatrixThreePrimesFFTMulDomain::mul (three-primes.inl:57-191); returns
    (status, c polfirst of stride sa+sb-1, primes used)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    c = np.zeros(m * n * (sa + sb - 1), dtype=np.uint32)
    primes = np.zeros(64, dtype=np.uint32)
    npr = ctypes.c_int(0)
    st = lib().or_polymat_mul_three_primes(p, m, k, n, _p(a), sa, _p(b), sb, _p(c), _p(primes), ctypes.byref(npr))
    return st, c, [int(x) for x in primes[:npr.value]]


def stage_forward_mt(p, L, pow_w, polys, count, threads):
    lib().or_stage_forward_mt(p, L, _p(pow_w), _p(polys), count, threads)


def stage_inverse_mt(p, L, pow_w, polys, count, threads):
    lib().or_stage_inverse_mt(p, L, _p(pow_w), _p(polys), count, threads)


def stage_pointwise_mt(p, m, k, n, pts, ahat, bhat, chat, pt0, npts, threads):
    lib().or_stage_pointwise_mt(p, m, k, n, pts, _p(ahat), _p(bhat), _p(chat), pt0, npts, threads)


def stage_forward(p, L, pow_w, polys, count):
    lib().or_stage_forward(p, L, _p(pow_w), _p(polys), count)


def stage_inverse(p, L, pow_w, polys, count):
    lib().or_stage_inverse(p, L, _p(pow_w), _p(polys), count)


def stage_pointwise(p, m, k, n, pts, ahat, bhat, chat, pt0, npts):
    lib().or_stage_pointwise(p, m, k, n, pts, _p(ahat), _p(bhat), _p(chat), pt0, npts)
