"""Host-side mirror of LinBox's polynomial-matrix interface over the C ABI.

* PolynomialMatrix            -- PolynomialMatrix<Modular<uint32_t,uint64_t>,PMType>
                                 (linbox/matrix/polynomial-matrix.h:54-910): the three
                                 storage layouts, get/ref, resize, views (at()).
* PolynomialMatrixMulDomain   -- PolynomialMatrixMulDomain::mul (polynomial-matrix-domain.h:56-75);
                                 like the reference it does not forward max_rowdeg (:61).
* PolynomialMatrixFFTMulDomain / PolynomialMatrixFFTPrimeMulDomain
                              -- the FFT domains (matpoly-mult-fft-wordsize.inl:57-88,
                                 matpoly-mult-fft-wordsize-fast.inl:62-190).
* PolynomialMatrixThreePrimesFFTMulDomain
                              -- the few-primes domain for p < 2^29 (three-primes.inl:40-328).
* FFT                         -- FFT<Field> with FFT_direct / FFT_inverse (fft.h:57-144).

Storage is a numpy uint32 array (host) or any device tensor exposing
data_ptr() (e.g. a torch int32 CUDA tensor holding the same bits).
"""
import ctypes

import numpy as np

from . import (LinboxError, Polmat, Job, check, default_context, lib, POLFIRST, MATFIRST, MATROWFIRST,
               METHOD_AUTO)

LAYOUT_NAMES = {"polfirst": POLFIRST, "matfirst": MATFIRST, "matrowfirst": MATROWFIRST}


def _layout(x):
    return LAYOUT_NAMES[x] if isinstance(x, str) else int(x)


def _ptr(data):
    if isinstance(data, np.ndarray):
        return data.ctypes.data
    return int(data.data_ptr())


class PolynomialMatrix:
    """rows x cols matrix of polynomials with `size` coefficients each."""

    def __init__(self, rows, cols, size, layout=POLFIRST, data=None, nlimbs=1):
        self.rows, self.cols = int(rows), int(cols)
        self.layout = _layout(layout)
        self.nlimbs = nlimbs
        self._size = int(size)
        self.storage_size = int(size)
        self.shift = 0
        self._parent = None
        n = self.rows * self.cols * self.storage_size * nlimbs
        if data is None:
            data = np.zeros(n, dtype=np.uint32)
        self.data = data

    # --- reference accessors -------------------------------------------
    def rowdim(self):
        return self.rows

    def coldim(self):
        return self.cols

    def size(self):
        return self._size

    def degree(self):
        return self._size - 1

    def index(self, e, t):
        """Storage word of coefficient t of entry e = i*cols+j (polynomial-matrix.h:131-137,422-428,660-683)."""
        tt = self.shift + t
        if self.layout == POLFIRST:
            return e * self.storage_size + tt
        if self.layout == MATFIRST:
            return e + tt * self.rows * self.cols
        i, j = divmod(e, self.cols)
        return i * self.cols * self.storage_size + j + tt * self.cols

    def get(self, i, j, k):
        return int(self.data[self.index(i * self.cols + j, k)])

    def ref_set(self, i, j, k, v):
        self.data[self.index(i * self.cols + j, k)] = v

    # --- bulk conversion (host numpy) ----------------------------------
    def _index_grid(self):
        e = np.arange(self.rows * self.cols, dtype=np.int64)[:, None]
        t = np.arange(self._size, dtype=np.int64)[None, :] + self.shift
        if self.layout == POLFIRST:
            return e * self.storage_size + t
        if self.layout == MATFIRST:
            return e + t * self.rows * self.cols
        i, j = e // self.cols, e % self.cols
        return i * self.cols * self.storage_size + j + t * self.cols

    def to_polfirst(self):
        """(rows*cols, size) array of the visible coefficients."""
        return np.asarray(self.data)[self._index_grid()]

    def set_polfirst(self, arr):
        self.data[self._index_grid()] = np.asarray(arr, dtype=np.uint32).reshape(self.rows * self.cols, self._size)

    @classmethod
    def from_polfirst(cls, arr, rows, cols, layout=POLFIRST):
        arr = np.asarray(arr, dtype=np.uint32).reshape(rows * cols, -1)
        M = cls(rows, cols, arr.shape[1], layout)
        M.set_polfirst(arr)
        return M

    # --- size changes ---------------------------------------------------
    def resize(self, s):
        """Whole matrices reallocate keeping coefficients [0,min(s,size)) (polynomial-matrix.h:184-203);
        views can only shrink (SubPolynomialMatrix::resize, :983-989)."""
        s = int(s)
        if self._parent is not None:
            if s > self._size:
                raise LinboxError(5, "SubPolynomialMatrix: resizing with larger size than available")
            self._size = s
            return
        if s == self._size:
            return
        old = self.to_polfirst()
        keep = min(s, self._size)
        self.storage_size = s
        self._size = s
        self.data = np.zeros(self.rows * self.cols * s, dtype=np.uint32)
        new = np.zeros((self.rows * self.cols, s), dtype=np.uint32)
        new[:, :keep] = old[:, :keep]
        self.set_polfirst(new)

    def at(self, i, j):
        """View of coefficients i..j (SubPolynomialMatrix(M,i,j), polynomial-matrix.h:954-967)."""
        if not (i < self._size and i <= j < self._size):
            raise LinboxError(7, "invalid view range")
        v = PolynomialMatrix.__new__(PolynomialMatrix)
        v.__dict__.update(self.__dict__)
        v.shift = self.shift + i
        v._size = j - i + 1
        v._parent = self
        return v

    def polmat(self):
        return Polmat(_ptr(self.data), self.layout, self.rows, self.cols, self._size, self.storage_size, self.shift)

    def copy(self, M, beg=0, end=None, start=0, ctx=None, stream=None):
        """copy(M, beg, end, start): coefficients beg..end of every entry of M into
        start.. of this matrix or view, any layouts (polynomial-matrix.h:206-243,508-532,
        763-790; a view writes at its shift, :988-992); copy(M) copies all of M.  Runs on
        the device (lbx_polymat_copy_u32); the rest of this matrix is left as it is."""
        end = M.size() - 1 if end is None else end
        ctx = ctx or default_context()
        pd, ps = self.polmat(), M.polmat()
        check(lib().lbx_polymat_copy_u32(ctx.handle, ctypes.byref(pd), ctypes.byref(ps), beg, end, start, stream))
        return self


class PolynomialMatrixFFTMulDomain:
    """PolynomialMatrixFFTMulDomain<Modular<uint32_t,uint64_t>> (matpoly-mult-fft-wordsize.inl:39-88)."""

    def __init__(self, p, ctx=None, method=METHOD_AUTO, stream=None):
        self.p = int(p)
        self.ctx = ctx or default_context()
        self.method = method
        self.stream = stream

    def mul(self, c, a, b, max_rowdeg=0):
        deg = max_rowdeg if max_rowdeg else a.size() + b.size() - 2
        if a.coldim() != b.rowdim():
            raise LinboxError(5, "a.coldim() != b.rowdim()")
        c.resize(deg + 1)
        pc, pa, pb = c.polmat(), a.polmat(), b.polmat()
        check(lib().lbx_polymat_mul_u32(self.ctx.handle, self.p, ctypes.byref(pc), ctypes.byref(pa),
                                        ctypes.byref(pb), max_rowdeg, self.method, self.stream))
        return c

    def midproduct(self, c, a, b, small_left=True, n0=0, n1=0):
        """c = (a*b x^-(hdeg-1)) mod x^pts on c.size() coefficients, hdeg = n0 or c.size(),
        deg = n1 or 2*hdeg-1 (matpoly-mult-fft-wordsize.inl:90-127); c is not resized."""
        if a.coldim() != b.rowdim():
            raise LinboxError(5, "a.coldim() != b.rowdim()")
        pc, pa, pb = c.polmat(), a.polmat(), b.polmat()
        check(lib().lbx_polymat_midproduct_u32(self.ctx.handle, self.p, ctypes.byref(pc), ctypes.byref(pa),
                                               ctypes.byref(pb), 1 if small_left else 0, n0, n1, self.method,
                                               self.stream))
        return c


class PolynomialMatrixNaiveMulDomain:
    """PolynomialMatrixNaiveMulDomain::midproduct (matpoly-mult-naive.h:85-125), on the device."""

    def __init__(self, p, ctx=None, stream=None):
        self.p = int(p)
        self.ctx = ctx or default_context()
        self.stream = stream

    def midproduct(self, c, a, b, small_left=True, n0=0, n1=0):
        if a.coldim() != b.rowdim():
            raise LinboxError(5, "a.coldim() != b.rowdim()")
        pc, pa, pb = c.polmat(), a.polmat(), b.polmat()
        check(lib().lbx_polymat_midproduct_naive_u32(self.ctx.handle, self.p, ctypes.byref(pc), ctypes.byref(pa),
                                                     ctypes.byref(pb), n0, n1, self.stream))
        return c


class PolynomialMatrixMulDomain(PolynomialMatrixFFTMulDomain):
    """PolynomialMatrixMulDomain (polynomial-matrix-domain.h:40-104).

    mul: d = a.size()+b.size(); FFT when d > 2, naive otherwise -- both run on
    the device here.  max_rowdeg is accepted and, as in the reference (:61), not
    forwarded.  midproduct (:77-91): FFT when b.size() > 2, naive otherwise.
    midproductgen (:94-103): naive when c.size() <= 4, FFT otherwise."""

    def mul(self, c, a, b, max_rowdeg=0):
        return PolynomialMatrixFFTMulDomain.mul(self, c, a, b, 0)

    def midproduct(self, c, a, b):
        if b.size() > 2:
            return PolynomialMatrixFFTMulDomain.midproduct(self, c, a, b)
        return PolynomialMatrixNaiveMulDomain(self.p, self.ctx, self.stream).midproduct(c, a, b)

    def midproductgen(self, c, a, b, small_left=True, n0=0, n1=0):
        if c.size() <= 4:
            return PolynomialMatrixNaiveMulDomain(self.p, self.ctx, self.stream).midproduct(c, a, b, small_left, n0,
                                                                                          n1)
        return PolynomialMatrixFFTMulDomain.midproduct(self, c, a, b, small_left, n0, n1)


class OrderBasis:
    """OrderBasis<Field> (order-basis.h:181-600): PM_Basis and M_Basis on the device.

    serie: rowdim m x coldim n PolynomialMatrix with at least `order`
    coefficients; sigma: m x m PolynomialMatrix with at least order+1
    coefficients; shift: list of m non-negative ints, updated in place like the
    reference's std::vector<size_t>&.  Both return the degree d of sigma and
    leave sigma.size() == d+1 (sigma.resize(d+1), :592-597 and :297-298)."""

    def __init__(self, p, ctx=None, method=METHOD_AUTO, stream=None):
        self.p = int(p)
        self.ctx = ctx or default_context()
        self.method = method
        self.stream = stream

    def _run(self, fn, sigma, serie, order, shift, *extra):
        m = serie.rowdim()
        if len(shift) != m:
            raise LinboxError(5, "shift must have serie.rowdim() entries")
        sh = (ctypes.c_size_t * m)(*[int(x) for x in shift])
        deg = ctypes.c_size_t(0)
        ps, pz = sigma.polmat(), serie.polmat()
        check(fn(self.ctx.handle, self.p, ctypes.byref(ps), ctypes.byref(pz), int(order), sh, ctypes.byref(deg),
                 *extra, self.stream))
        d = int(deg.value)
        shift[:] = [int(x) for x in sh]
        if sigma._parent is None and isinstance(sigma.data, np.ndarray):
            sigma.resize(d + 1)
        else:
            sigma._size = d + 1  # device storage keeps its allocation, like a view
        return d

    def PM_Basis(self, sigma, serie, order, shift):
        """order-basis.h:219-326 (M_Basis below order 32, :232)."""
        return self._run(lib().lbx_order_basis_pm_u32, sigma, serie, order, shift, self.method)

    def M_Basis(self, sigma, serie, order, shift):
        """order-basis.h:346-600."""
        return self._run(lib().lbx_order_basis_m_u32, sigma, serie, order, shift)


class PolynomialMatrixFFTPrimeMulDomain:
    """PolynomialMatrixFFTPrimeMulDomain (matpoly-mult-fft-wordsize-fast.inl:40-190)."""

    def __init__(self, p, ctx=None, method=METHOD_AUTO, stream=None):
        self.p = int(p)
        self.ctx = ctx or default_context()
        self.method = method
        self.stream = stream

    def mul(self, c, a, b, max_rowdeg=0):
        # mul (:62-82) pads to pts and calls mul_fft; the C ABI does both
        return PolynomialMatrixFFTMulDomain(self.p, self.ctx, self.method, self.stream).mul(c, a, b, max_rowdeg)

    def mul_fft(self, lpts, c, a, b):
        """a, b, c: polfirst PolynomialMatrix of size 2^lpts (a and b are not modified)."""
        check(lib().lbx_polymat_mul_fft_u32(self.ctx.handle, self.p, lpts, a.rowdim(), a.coldim(), b.coldim(),
                                            _ptr(a.data), _ptr(b.data), _ptr(c.data), self.method, self.stream))
        return c


class PolynomialMatrixThreePrimesFFTMulDomain(PolynomialMatrixFFTMulDomain):
    """PolynomialMatrixThreePrimesFFTMulDomain (matpoly-mult-fft-wordsize-three-primes.inl:40-328):
    word primes below 2^29 (the constructor raises for larger ones, :57-64).  mul and
    midproduct take the single-prime route when (p-1) % pts == 0 (:107-111, :238-243),
    else the few-primes route: products modulo 31-bit FFT primes whose product exceeds
    (p-1)^2 k min(sa,sb), and the mixed-radix reconstruction mod p (:164-181)."""

    def __init__(self, p, ctx=None, method=METHOD_AUTO, stream=None):
        if int(p).bit_length() > 29:
            raise LinboxError(8, "too large FFT Prime (more than 29 bits)")
        super().__init__(p, ctx, method, stream)


def _ntt(fn, p, lpts, data, count, root, ctx, stream):
    ctx = ctx or default_context()
    used = ctypes.c_uint32(0)
    check(fn(ctx.handle, p, lpts, root, _ptr(data), count, ctypes.byref(used), stream))
    return used.value


def ntt_forward(p, lpts, data, count=None, root=0, ctx=None, stream=None):
    """In-place FFT_direct of `count` polynomials of 2^lpts words; returns the root used."""
    count = (len(data) >> lpts) if count is None else count
    return _ntt(lib().lbx_ntt_forward_u32, p, lpts, data, count, root, ctx, stream)


def ntt_inverse(p, lpts, data, count=None, root=0, ctx=None, stream=None):
    """In-place FFT_inverse (DIT, no scaling) of `count` polynomials; returns the root used."""
    count = (len(data) >> lpts) if count is None else count
    return _ntt(lib().lbx_ntt_inverse_u32, p, lpts, data, count, root, ctx, stream)


class FFT:
    """FFT<Modular<uint32_t,uint64_t>>(F, k, w) (fft.h:57-144)."""

    def __init__(self, p, k, w=0, ctx=None):
        self.p, self.k, self.ctx = int(p), int(k), ctx
        probe = np.zeros(1 << k, dtype=np.uint32)
        self.w = ntt_forward(self.p, self.k, probe, count=0, root=w, ctx=ctx)

    def root(self):
        return self.p - 1 if self.k == 1 else self.w

    def invroot(self):
        if self.k == 1:
            return self.p - 1
        return (self.p - pow(self.w, (1 << (self.k - 1)) - 1, self.p)) % self.p

    def FFT_direct(self, coeffs):
        ntt_forward(self.p, self.k, coeffs, count=1, root=self.w, ctx=self.ctx)
        return coeffs

    def FFT_inverse(self, coeffs):
        ntt_inverse(self.p, self.k, coeffs, count=1, root=self.w, ctx=self.ctx)
        return coeffs


def mul_mp(P, c, a, b, ctx=None, method=METHOD_AUTO, stream=None, max_rowdeg=0):
    """Product over Z/PZ for a multiprecision P (int); a, b, c hold nlimbs words per coefficient.
    With max_rowdeg, max_rowdeg+1 coefficients of the cyclic product (mul_crtla)."""
    ctx = ctx or default_context()
    nl = a.nlimbs
    limbs = np.array([(P >> (32 * i)) & 0xFFFFFFFF for i in range(nl)], dtype=np.uint32)
    pc, pa, pb = c.polmat(), a.polmat(), b.polmat()
    check(lib().lbx_polymat_mul_mp(ctx.handle, limbs.ctypes.data, nl, ctypes.byref(pc), ctypes.byref(pa),
                                   ctypes.byref(pb), max_rowdeg, method, stream))
    c._size = pc.size
    return c


def mul_zz(c, a, b, ctx=None, method=METHOD_AUTO, stream=None, max_rowdeg=0):
    """PolynomialMatrixFFTMulDomain<ZRing<integer>>::mul (matpoly-mult-fft-multiprecision.inl:74-89):
    the product over the integers.  a and b hold a.nlimbs words per coefficient, c
    c.nlimbs, all two's complement; c.nlimbs must hold the product's bound."""
    ctx = ctx or default_context()
    if a.nlimbs != b.nlimbs:
        raise LinboxError(7, "a and b must have the same limb count")
    pc, pa, pb = c.polmat(), a.polmat(), b.polmat()
    check(lib().lbx_polymat_mul_zz(ctx.handle, ctypes.byref(pc), c.nlimbs, ctypes.byref(pa), ctypes.byref(pb),
                                   a.nlimbs, max_rowdeg, method, stream))
    c._size = pc.size
    return c


def to_limbs(values, nlimbs):
    """Signed Python ints -> uint32 array of nlimbs two's complement words each."""
    out = np.zeros((len(values), nlimbs), dtype=np.uint32)
    mask = (1 << (32 * nlimbs)) - 1
    for i, v in enumerate(values):
        u = int(v) & mask
        for l in range(nlimbs):
            out[i, l] = (u >> (32 * l)) & 0xFFFFFFFF
    return out.reshape(-1)


def from_limbs(words, nlimbs):
    """uint32 array of nlimbs two's complement words each -> signed Python ints."""
    w = np.asarray(words, dtype=np.uint32).reshape(-1, nlimbs)
    out = []
    for row in w:
        u = 0
        for l in range(nlimbs - 1, -1, -1):
            u = (u << 32) | int(row[l])
        if u >> (32 * nlimbs - 1):
            u -= 1 << (32 * nlimbs)
        out.append(u)
    return out


def mul_batch(jobs, ctx=None, method=METHOD_AUTO, stream=None):
    """jobs: list of dicts with p, m, k, n, sa, sb and device pointers a, b, c."""
    ctx = ctx or default_context()
    arr = (Job * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = Job(j["p"], j["m"], j["k"], j["n"], j["sa"], j["sb"], j["a"], j["b"], j["c"])
    check(lib().lbx_polymat_mul_batch(ctx.handle, arr, len(jobs), method, stream))


def check_midproduct(p, c, a, b, n0=0, n1=0, ctx=None, stream=None):
    """check_midproduct(c, a, b, smallLeft, n0, n1) (matpoly-mult-fft.h:168-262) on the
    device by random row and column projections: True when coefficients
    0..min(c.size(), deg-hdeg+1)-1 of c equal coefficients hdeg-1.. of a.b mod p
    (hdeg = n0 or c.size(), deg = n1 or 2 hdeg - 1)."""
    ctx = ctx or default_context()
    if a.coldim() != b.rowdim():
        raise LinboxError(5, "a.coldim() != b.rowdim()")
    pc, pa, pb = c.polmat(), a.polmat(), b.polmat()
    eq = ctypes.c_int(0)
    check(lib().lbx_polymat_check_midproduct_u32(ctx.handle, int(p), ctypes.byref(pc), ctypes.byref(pa),
                                                 ctypes.byref(pb), int(n0), int(n1), ctypes.byref(eq), stream))
    return bool(eq.value)


def check_mul(p, c, a, b, deg=0, x=1, ctx=None, stream=None):
    """check_mul(c, a, b, deg) (matpoly-mult-fft.h:87-163) on the device at the point x
    (the reference evaluates at 1): True when C(x) = A(x).B(x) mod p, with c taken over
    its first deg coefficients (c.size() when 0)."""
    ctx = ctx or default_context()
    if a.coldim() != b.rowdim():
        raise LinboxError(5, "a.coldim() != b.rowdim()")
    pc, pa, pb = c.polmat(), a.polmat(), b.polmat()
    eq = ctypes.c_int(0)
    check(lib().lbx_polymat_check_mul_u32(ctx.handle, int(p), ctypes.byref(pc), ctypes.byref(pa), ctypes.byref(pb),
                                          int(deg), int(x) % int(p), ctypes.byref(eq), stream))
    return bool(eq.value)
