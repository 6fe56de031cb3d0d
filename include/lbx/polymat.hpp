// polymat.hpp -- header-only C++ mirror of LinBox's polynomial-matrix
// multiplication interface over the C ABI of lbx.h.
//
// Mirrors, with the reference's names and argument meaning:
//   Modular<uint32_t,uint64_t>               -> lbx::Modular (Givaro::Modular subset used by the path)
//   PolynomialMatrix<Field, PMType>          -> linbox/matrix/polynomial-matrix.h:54-910
//   SubPolynomialMatrix (views, at(i,j))     -> polynomial-matrix.h:940-1036
//   PolynomialMatrixMulDomain::mul           -> polynomial-matrix-domain.h:56-75
//   PolynomialMatrixFFTMulDomain::mul        -> matpoly-mult-fft-wordsize.inl:57-88
//   PolynomialMatrixFFTPrimeMulDomain        -> matpoly-mult-fft-wordsize-fast.inl:40-190
//   FFT<Field>::FFT_direct / FFT_inverse     -> fft.h:57-144
// Errors are thrown as lbx::LinboxError carrying the lbx_status, with the
// reference's message text (lbx_strerror).
#ifndef LBX_POLYMAT_HPP
#define LBX_POLYMAT_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbx.h"

namespace lbx {

class LinboxError : public std::runtime_error {
public:
    LinboxError(lbx_status s, const std::string &what) : std::runtime_error(what), status(s) {}
    lbx_status status;
};

inline void check(lbx_status s)
{
    if (s != LBX_OK) throw LinboxError(s, lbx_strerror(s));
}

// One context per process by default; pass an explicit one for other devices.
inline lbx_ctx *default_ctx()
{
    static lbx_ctx *ctx = [] {
        lbx_ctx *c = nullptr;
        check(lbx_init(&c, 0));
        return c;
    }();
    return ctx;
}

// Z/pZ with uint32_t elements (the Givaro::Modular<uint32_t,uint64_t> operations the path uses).
class Modular {
public:
    typedef uint32_t Element;
    explicit Modular(uint32_t p) : p_(p) {}
    uint64_t cardinality() const { return p_; }
    uint64_t characteristic() const { return p_; }
    Element zero = 0, one = 1;
    Element &init(Element &x, uint64_t v) const { return x = (Element)(v % p_); }
    Element &mul(Element &r, Element a, Element b) const { return r = (Element)((uint64_t)a * b % p_); }
    Element &add(Element &r, Element a, Element b) const { return r = (Element)(((uint64_t)a + b) % p_); }

private:
    uint32_t p_;
};

enum PMType { polfirst = LBX_POLFIRST, matfirst = LBX_MATFIRST, matrowfirst = LBX_MATROWFIRST };

template <class MatPoly>
class SubPolynomialMatrix;

template <PMType T>
class PolynomialMatrix {
public:
    typedef Modular Field;
    typedef uint32_t Element;
    typedef SubPolynomialMatrix<PolynomialMatrix> view;

    // polynomial matrix in F[x]^(r x c) of size s (degree s-1)
    PolynomialMatrix(const Field &F, size_t r, size_t c, size_t s)
        : fld_(&F), rep_(r * c * s, 0), row_(r), col_(c), size_(s) {}

    size_t rowdim() const { return row_; }
    size_t coldim() const { return col_; }
    size_t size() const { return size_; }
    size_t degree() const { return size_ - 1; }
    const Field &field() const { return *fld_; }
    Element *getPointer() { return rep_.data(); }
    const Element *getPointer() const { return rep_.data(); }

    // storage index of coefficient k of entry i (polynomial-matrix.h:131-137, 422-428, 660-683)
    size_t index(size_t i, size_t k) const
    {
        if (T == polfirst) return i * size_ + k;
        if (T == matfirst) return i + k * row_ * col_;
        return (i / col_) * col_ * size_ + (i % col_) + k * col_;
    }
    Element &ref(size_t i, size_t k) { return rep_[index(i, k)]; }
    Element get(size_t i, size_t k) const { return rep_[index(i, k)]; }
    Element &ref(size_t i, size_t j, size_t k) { return ref(i * col_ + j, k); }
    Element get(size_t i, size_t j, size_t k) const { return get(i * col_ + j, k); }

    // resize keeping coefficients [0, min(s,size)) (polynomial-matrix.h:184-203)
    void resize(size_t s)
    {
        if (s == size_) return;
        std::vector<Element> old(rep_);
        const size_t keep = s < size_ ? s : size_;
        PolynomialMatrix tmp(*fld_, row_, col_, s);
        for (size_t i = 0; i < row_ * col_; ++i)
            for (size_t k = 0; k < keep; ++k) tmp.rep_[tmp.index(i, k)] = old[index(i, k)];
        rep_.swap(tmp.rep_);
        size_ = s;
    }

    view at(size_t i, size_t j) { return view(*this, i, j); }

    // copy(M, beg, end, start): coefficients beg..end of M into start.. of this
    // matrix, any pair of layouts (polynomial-matrix.h:206-243, 508-532, 763-790),
    // through the device (lbx_polymat_copy_u32); copy(M) copies all of M.
    template <class Mat>
    void copy(const Mat &M, size_t beg, size_t end, size_t start = 0)
    {
        lbx_polmat d = desc(), s = M.desc();
        check(lbx_polymat_copy_u32(default_ctx(), &d, &s, beg, end, start, nullptr));
    }
    template <class Mat>
    void copy(const Mat &M)
    {
        copy(M, 0, M.size() - 1);
    }

    lbx_polmat desc() const
    {
        lbx_polmat d;
        d.data = (void *)rep_.data();
        d.layout = (int)T;
        d.rows = row_;
        d.cols = col_;
        d.size = size_;
        d.storage_size = size_;
        d.shift = 0;
        return d;
    }
    void after_mul(const lbx_polmat &) {}

private:
    const Field *fld_;
    std::vector<Element> rep_;
    size_t row_, col_, size_;
};

// Degree window [i, j] of a polynomial matrix; it can only shrink.
template <class MatPoly>
class SubPolynomialMatrix {
public:
    typedef typename MatPoly::Element Element;
    SubPolynomialMatrix(MatPoly &M, size_t i, size_t j) : ptr_(&M), size_(j - i + 1), shift_(i)
    {
        if (!(i < M.size() && i <= j && j < M.size())) throw LinboxError(LBX_ERR_INVALID_ARG, "invalid view range");
    }
    size_t rowdim() const { return ptr_->rowdim(); }
    size_t coldim() const { return ptr_->coldim(); }
    size_t size() const { return size_; }
    size_t degree() const { return size_ - 1; }
    Element get(size_t i, size_t k) const { return ptr_->get(i, k + shift_); }
    Element get(size_t i, size_t j, size_t k) const { return get(i * coldim() + j, k); }
    void resize(size_t s)
    {
        if (s <= size_)
            size_ = s;
        else
            throw LinboxError(LBX_ERR_SIZE_MISMATCH,
                              "SubPolynomialMatrix: resizing with larger size than available, exiting ...");
    }
    lbx_polmat desc() const
    {
        lbx_polmat d = ptr_->desc();
        d.storage_size = ptr_->size();
        d.shift = shift_;
        d.size = size_;
        return d;
    }
    void after_mul(const lbx_polmat &d) { size_ = d.size; }
    // SubPolynomialMatrix::copy(M, beg, end) writes at the view's shift (polynomial-matrix.h:988-992)
    template <class Mat>
    void copy(const Mat &M, size_t beg, size_t end)
    {
        lbx_polmat d = desc(), s = M.desc();
        check(lbx_polymat_copy_u32(default_ctx(), &d, &s, beg, end, 0, nullptr));
    }
    template <class Mat>
    void copy(const Mat &M)
    {
        copy(M, 0, M.size() - 1);
    }

private:
    MatPoly *ptr_;
    size_t size_, shift_;
};

class PolynomialMatrixFFTMulDomain {
public:
    typedef Modular Field;
    explicit PolynomialMatrixFFTMulDomain(const Field &F, lbx_ctx *ctx = nullptr, int method = LBX_METHOD_AUTO,
                                          void *stream = nullptr)
        : fld_(&F), ctx_(ctx ? ctx : default_ctx()), method_(method), stream_(stream) {}
    const Field &field() const { return *fld_; }

    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void mul(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, size_t max_rowdeg = 0) const
    {
        if (a.coldim() != b.rowdim()) throw LinboxError(LBX_ERR_SIZE_MISMATCH, "a.coldim() != b.rowdim()");
        const size_t deg = max_rowdeg ? max_rowdeg : a.size() + b.size() - 2;
        c.resize(deg + 1);  // matpoly-mult-fft-wordsize.inl:60
        lbx_polmat dc = c.desc(), da = a.desc(), db = b.desc();
        check(lbx_polymat_mul_u32(ctx_, (uint32_t)fld_->cardinality(), &dc, &da, &db, max_rowdeg, method_, stream_));
        c.after_mul(dc);
    }

    // c = (a*b x^-(hdeg-1)) mod x^pts on c.size() coefficients, hdeg = n0 ? n0 : c.size(),
    // deg = n1 ? n1 : 2*hdeg-1 (matpoly-mult-fft-wordsize.inl:90-127); c keeps its size.
    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void midproduct(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, bool smallLeft = true, size_t n0 = 0,
                    size_t n1 = 0) const
    {
        if (a.coldim() != b.rowdim()) throw LinboxError(LBX_ERR_SIZE_MISMATCH, "a.coldim() != b.rowdim()");
        lbx_polmat dc = c.desc(), da = a.desc(), db = b.desc();
        check(lbx_polymat_midproduct_u32(ctx_, (uint32_t)fld_->cardinality(), &dc, &da, &db, smallLeft ? 1 : 0, n0,
                                         n1, method_, stream_));
    }

private:
    const Field *fld_;
    lbx_ctx *ctx_;
    int method_;
    void *stream_;
};

// PolynomialMatrixNaiveMulDomain::midproduct (matpoly-mult-naive.h:85-125).
class PolynomialMatrixNaiveMulDomain {
public:
    typedef Modular Field;
    explicit PolynomialMatrixNaiveMulDomain(const Field &F, lbx_ctx *ctx = nullptr, void *stream = nullptr)
        : fld_(&F), ctx_(ctx ? ctx : default_ctx()), stream_(stream) {}
    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void midproduct(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, bool smallLeft = true, size_t n0 = 0,
                    size_t n1 = 0) const
    {
        (void)smallLeft;  // both branches of :91-123 sum the same coefficient
        if (a.coldim() != b.rowdim()) throw LinboxError(LBX_ERR_SIZE_MISMATCH, "a.coldim() != b.rowdim()");
        lbx_polmat dc = c.desc(), da = a.desc(), db = b.desc();
        check(lbx_polymat_midproduct_naive_u32(ctx_, (uint32_t)fld_->cardinality(), &dc, &da, &db, n0, n1, stream_));
    }

private:
    const Field *fld_;
    lbx_ctx *ctx_;
    void *stream_;
};

// PolynomialMatrixMulDomain (polynomial-matrix-domain.h:40-104).  mul:
// d = a.size()+b.size() > 2 takes the FFT domain, otherwise the naive one;
// both run on the device here.  Like the reference (:61), max_rowdeg is not
// forwarded.  midproduct: FFT when b.size() > 2, naive otherwise (:77-91);
// midproductgen: naive when c.size() <= 4, FFT otherwise (:94-103).
class PolynomialMatrixMulDomain {
public:
    typedef Modular Field;
    explicit PolynomialMatrixMulDomain(const Field &F, lbx_ctx *ctx = nullptr, int method = LBX_METHOD_AUTO,
                                       void *stream = nullptr)
        : fft_(F, ctx, method, stream), naive_(F, ctx, stream) {}
    const Field &field() const { return fft_.field(); }
    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void mul(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, size_t max_rowdeg = 0) const
    {
        (void)max_rowdeg;
        fft_.mul(c, a, b);
    }
    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void midproduct(Matrix1 &c, const Matrix2 &a, const Matrix3 &b) const
    {
        if (b.size() > 2)
            fft_.midproduct(c, a, b);
        else
            naive_.midproduct(c, a, b);
    }
    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void midproductgen(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, bool smallLeft = true, size_t n0 = 0,
                       size_t n1 = 0) const
    {
        if (c.size() <= 4)
            naive_.midproduct(c, a, b, smallLeft, n0, n1);
        else
            fft_.midproduct(c, a, b, smallLeft, n0, n1);
    }

private:
    PolynomialMatrixFFTMulDomain fft_;
    PolynomialMatrixNaiveMulDomain naive_;
};

// OrderBasis<Field> (order-basis.h:181-600): PM_Basis and M_Basis on the
// device.  serie: m x n with at least `order` coefficients; sigma: m x m,
// resized by the call to degree+1 (sigma.resize, :297-298, :592-597); shift:
// m entries, updated in place.  Both return the degree of sigma.
class OrderBasis {
public:
    typedef Modular Field;
    explicit OrderBasis(const Field &F, lbx_ctx *ctx = nullptr, int method = LBX_METHOD_AUTO, void *stream = nullptr)
        : fld_(&F), ctx_(ctx ? ctx : default_ctx()), method_(method), stream_(stream) {}
    const Field &field() const { return *fld_; }

    template <typename PMatrix1, typename PMatrix2>
    size_t PM_Basis(PMatrix1 &sigma, const PMatrix2 &serie, size_t order, std::vector<size_t> &shift)
    {
        return run(true, sigma, serie, order, shift);
    }
    template <typename PMatrix1, typename PMatrix2>
    size_t M_Basis(PMatrix1 &sigma, const PMatrix2 &serie, size_t order, std::vector<size_t> &shift)
    {
        return run(false, sigma, serie, order, shift);
    }

private:
    template <typename PMatrix1, typename PMatrix2>
    size_t run(bool divide, PMatrix1 &sigma, const PMatrix2 &serie, size_t order, std::vector<size_t> &shift)
    {
        if (shift.size() != serie.rowdim()) throw LinboxError(LBX_ERR_SIZE_MISMATCH, "shift size != serie.rowdim()");
        if (sigma.size() < order + 1) sigma.resize(order + 1);
        lbx_polmat ds = sigma.desc(), dz = serie.desc();
        size_t deg = 0;
        const uint32_t p = (uint32_t)fld_->cardinality();
        if (divide)
            check(lbx_order_basis_pm_u32(ctx_, p, &ds, &dz, order, shift.data(), &deg, method_, stream_));
        else
            check(lbx_order_basis_m_u32(ctx_, p, &ds, &dz, order, shift.data(), &deg, stream_));
        sigma.resize(deg + 1);
        return deg;
    }
    const Field *fld_;
    lbx_ctx *ctx_;
    int method_;
    void *stream_;
};

class PolynomialMatrixFFTPrimeMulDomain {
public:
    typedef Modular Field;
    typedef PolynomialMatrix<polfirst> MatrixP;
    explicit PolynomialMatrixFFTPrimeMulDomain(const Field &F, lbx_ctx *ctx = nullptr, int method = LBX_METHOD_AUTO,
                                               void *stream = nullptr)
        : fld_(&F), ctx_(ctx ? ctx : default_ctx()), method_(method), stream_(stream) {}

    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void mul(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, size_t max_rowdeg = 0) const
    {
        PolynomialMatrixFFTMulDomain(*fld_, ctx_, method_, stream_).mul(c, a, b, max_rowdeg);
    }

    // a, b, c must have size 2^lpts (wordsize-fast.inl:86); a and b are left unchanged
    void mul_fft(size_t lpts, MatrixP &c, MatrixP &a, MatrixP &b) const
    {
        check(lbx_polymat_mul_fft_u32(ctx_, (uint32_t)fld_->cardinality(), (unsigned)lpts, a.rowdim(), a.coldim(),
                                      b.coldim(), a.getPointer(), b.getPointer(), c.getPointer(), method_, stream_));
    }

private:
    const Field *fld_;
    lbx_ctx *ctx_;
    int method_;
    void *stream_;
};

// PolynomialMatrixThreePrimesFFTMulDomain (matpoly-mult-fft-wordsize-three-primes.inl:40-328):
// word primes below 2^29.  The constructor throws for larger primes (:57-64);
// mul and midproduct take the single-prime route when (p-1) % pts == 0
// (:107-111, :238-243) and the few-primes route otherwise, which is the route
// lbx_polymat_mul_u32 / lbx_polymat_midproduct_u32 choose for such primes.
class PolynomialMatrixThreePrimesFFTMulDomain {
public:
    typedef Modular Field;
    explicit PolynomialMatrixThreePrimesFFTMulDomain(const Field &F, lbx_ctx *ctx = nullptr,
                                                     int method = LBX_METHOD_AUTO, void *stream = nullptr)
        : dom_(F, ctx, method, stream)
    {
        if (F.cardinality() >= (1ull << 29))
            throw LinboxError(LBX_ERR_PRIME_TOO_LARGE, "too large FFT Prime (more than 29 bits)");
    }
    const Field &field() const { return dom_.field(); }

    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void mul(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, size_t max_rowdeg = 0) const
    {
        dom_.mul(c, a, b, max_rowdeg);
    }

    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void midproduct(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, bool smallLeft = true, size_t n0 = 0,
                    size_t n1 = 0) const
    {
        dom_.midproduct(c, a, b, smallLeft, n0, n1);
    }

private:
    PolynomialMatrixFFTMulDomain dom_;
};

// FFT<Modular<uint32_t,uint64_t>> (fft.h:57-144).
class FFT {
public:
    typedef Modular Field;
    typedef uint32_t Element;
    FFT(const Field &F, size_t k, Element w = 0, lbx_ctx *ctx = nullptr)
        : fld_(&F), k_(k), ctx_(ctx ? ctx : default_ctx())
    {
        // check_args (fft.h:133-144): computes the default root or validates w
        Element dummy = 0;
        check(lbx_ntt_forward_u32(ctx_, (uint32_t)F.cardinality(), (unsigned)k, w, &dummy, 0, &w_, nullptr));
    }
    // natural order in, bit-reversed order out
    void FFT_direct(Element *coeffs) const
    {
        check(lbx_ntt_forward_u32(ctx_, (uint32_t)fld_->cardinality(), (unsigned)k_, w_, coeffs, 1, nullptr, nullptr));
    }
    // bit-reversed order in, natural order out (no scaling)
    void FFT_inverse(Element *coeffs) const
    {
        check(lbx_ntt_inverse_u32(ctx_, (uint32_t)fld_->cardinality(), (unsigned)k_, w_, coeffs, 1, nullptr, nullptr));
    }
    size_t size() const { return (size_t)1 << k_; }
    size_t log2_size() const { return k_; }
    Element root() const { return k_ == 1 ? (Element)(fld_->cardinality() - 1) : w_; }
    Element invroot() const
    {
        const uint64_t p = fld_->cardinality();
        if (k_ == 1) return (Element)(p - 1);
        uint64_t r = 1, b = w_, e = ((uint64_t)1 << (k_ - 1)) - 1;
        while (e) {
            if (e & 1) r = r * b % p;
            b = b * b % p;
            e >>= 1;
        }
        return (Element)((p - r) % p);
    }

private:
    const Field *fld_;
    size_t k_;
    lbx_ctx *ctx_;
    Element w_ = 0;
};

}  // namespace lbx

#endif  // LBX_POLYMAT_HPP
