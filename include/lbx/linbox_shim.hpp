// linbox_shim.hpp -- drop-in specialisations that put LinBox's FFT
// polynomial-matrix product over Givaro::Modular<uint32_t,uint64_t> on the GPU
// through the C ABI of lbx/lbx.h.
//
// Include it after LinBox's headers (or after lbx/linbox_mirror.hpp, which
// stands in for them) and before the first use of these classes:
//
//   #include <linbox/algorithms/polynomial-matrix/polynomial-matrix-domain.h>
//   #include <linbox/algorithms/polynomial-matrix/order-basis.h>   // optional
//   #include <lbx/linbox_shim.hpp>
//
// It specialises, for Field = Givaro::Modular<uint32_t,uint64_t>:
//   PolynomialMatrixFFTMulDomain<Field>    matpoly-mult-fft-wordsize.inl:39-127 (mul, midproduct)
//   PolynomialMatrixFFTPrimeMulDomain<Field> matpoly-mult-fft-wordsize-fast.inl:40-190 (mul, mul_fft)
//   PolynomialMatrixNaiveMulDomain<Field>  matpoly-mult-naive.h:39-125 (the d <= 2 products and
//                                          the c.size() <= 4 middle products of the dispatcher)
//   OrderBasis<Field>                      order-basis.h:179-600 (PM_Basis, M_Basis, PM_Basis2),
//                                          when order-basis.h was included first
//   FFT<Field, S>                          fft.h:48-144 (FFT_direct, FFT_inverse), when fft.h
//                                          was included first
// so PolynomialMatrixMulDomain / PolynomialMatrixDomain (polynomial-matrix-domain.h:40-119),
// OrderBasis::PM_Basis and their callers (block Wiedemann through
// block-massey-domain.h:604-605) run on the GPU without source changes.
//
// Operands: a whole PolynomialMatrix<Field, T> of any of the three layouts is
// passed to the device by its storage pointer (getPointer()) and its layout T,
// with no copy.  A SubPolynomialMatrix<MatPoly> view is read through
// PolynomialMatrix::copy(view, beg, end) into a temporary matrix of its own
// layout, and an output view is written back with SubPolynomialMatrix::copy
// (polynomial-matrix.h:300-303, 985-992); both are public, so LinBox needs no
// change.  (With a public accessor of the view's _ptr and _shift -- one line,
// e.g. `const MatPoly *base() const { return _ptr; } size_t shift() const {
// return _shift; }` at polynomial-matrix.h:1028 -- views of matfirst and
// matrowfirst matrices could be passed without the copy.)
//
// Errors are thrown as LinBox::LinboxError with the reference's message text.
// All calls go through one lbx_ctx on device 0 (lbx_shim::context()), in
// order on the default stream; lbx_shim::set_context() selects another.
#ifndef LBX_LINBOX_SHIM_HPP
#define LBX_LINBOX_SHIM_HPP

#if !defined(LBX_LINBOX_MIRROR) && !defined(__LINBOX_POLYNOMIAL_MATRIX_DOMAIN_H)
#error "include LinBox's polynomial-matrix-domain.h (or lbx/linbox_mirror.hpp) before lbx/linbox_shim.hpp"
#endif

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <type_traits>
#include <vector>

#include "lbx.h"

namespace lbx_shim {

inline lbx_ctx *&context_slot()
{
    static lbx_ctx *ctx = nullptr;
    return ctx;
}

// The context every specialisation uses (created on device 0 at first use).
inline lbx_ctx *context()
{
    static std::once_flag once;
    std::call_once(once, [] {
        if (!context_slot()) {
            lbx_ctx *c = nullptr;
            const lbx_status st = lbx_init(&c, 0);
            if (st != LBX_OK) throw LinBox::LinboxError(lbx_strerror(st));
#if defined(CHECK_MATPOL_MUL) || defined(CHECK_MATPOL_MIDP)
            // the reference's product and middle-product checks
            // (polynomial-matrix-domain.h:72-74,88-90) on the device
            lbx_set_check(c, 1);
#endif
            context_slot() = c;
        }
    });
    return context_slot();
}
inline void set_context(lbx_ctx *ctx) { context_slot() = ctx; }

inline void check(lbx_status st)
{
    if (st != LBX_OK) throw LinBox::LinboxError(lbx_strerror(st));
}

// Descriptor of a whole polynomial matrix: its storage, with no copy.
template <class Field, LinBox::PMType T>
lbx_polmat describe(const LinBox::PolynomialMatrix<Field, T> &M)
{
    static_assert(sizeof(typename Field::Element) == sizeof(uint32_t), "32-bit elements");
    lbx_polmat d;
    d.data = (void *)M.getPointer();
    d.layout = (int)T;  // polfirst, matfirst, matrowfirst = LBX_POLFIRST, LBX_MATFIRST, LBX_MATROWFIRST
    d.rows = M.rowdim();
    d.cols = M.coldim();
    d.size = M.size();
    d.storage_size = M.size();
    d.shift = 0;
    return d;
}

template <class M>
struct WholeOf {
    typedef M type;
};
template <class MatPoly>
struct WholeOf<LinBox::SubPolynomialMatrix<MatPoly>> {
    typedef typename std::remove_const<MatPoly>::type type;
};

// An input operand: a whole matrix as it is, a view through a copy of its window.
template <class M>
class In {
public:
    explicit In(const M &m) : _m(m) {}
    lbx_polmat desc() const { return describe(_m); }

private:
    const M &_m;
};
template <class MatPoly>
class In<LinBox::SubPolynomialMatrix<MatPoly>> {
public:
    typedef typename WholeOf<LinBox::SubPolynomialMatrix<MatPoly>>::type Whole;
    explicit In(const LinBox::SubPolynomialMatrix<MatPoly> &v) : _tmp(v.field(), v.rowdim(), v.coldim(), v.size())
    {
        _tmp.copy(v, 0, v.size() - 1);
    }
    lbx_polmat desc() const { return describe(_tmp); }

private:
    Whole _tmp;
};

// An output operand of its current size; a view is computed into a temporary
// (holding the view's words first when `keep`, for outputs the call only
// partly writes) and copied back by finish().
template <class M>
class Out {
public:
    Out(M &m, bool) : _m(m) {}
    lbx_polmat desc() const { return describe(_m); }
    void finish() {}

private:
    M &_m;
};
template <class MatPoly>
class Out<LinBox::SubPolynomialMatrix<MatPoly>> {
public:
    typedef LinBox::SubPolynomialMatrix<MatPoly> View;
    typedef typename WholeOf<View>::type Whole;
    Out(View &v, bool keep) : _v(v), _tmp(v.field(), v.rowdim(), v.coldim(), v.size())
    {
        if (keep) _tmp.copy(v, 0, v.size() - 1);
    }
    lbx_polmat desc() const { return describe(_tmp); }
    void finish() { _v.copy(_tmp, 0, _v.size() - 1); }

private:
    View &_v;
    Whole _tmp;
};

}  // namespace lbx_shim

namespace LinBox {

// PolynomialMatrixFFTMulDomain<Modular<uint32_t,uint64_t>> (matpoly-mult-fft-wordsize.inl:39-127):
// mul: deg = max_rowdeg ? max_rowdeg : a.size()+b.size()-2, c.resize(deg+1),
// c = a.b (mod x^pts - 1 when max_rowdeg truncates) through lbx_polymat_mul_u32
// for every p < 2^32 (FFT primes below 2^31 on one prime, the others through
// FFT primes and CRT, as the reference's three routes); midproduct through
// lbx_polymat_midproduct_u32.
template <>
class PolynomialMatrixFFTMulDomain<Givaro::Modular<uint32_t, uint64_t>> {
public:
    typedef Givaro::Modular<uint32_t, uint64_t> Field;

    explicit PolynomialMatrixFFTMulDomain(const Field &F) : _field(&F), _p((uint64_t)F.cardinality()) {}
    const Field &field() const { return *_field; }

    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void mul(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, size_t max_rowdeg = 0) const
    {
        if (a.coldim() != b.rowdim()) lbx_shim::check(LBX_ERR_SIZE_MISMATCH);
        const size_t deg = max_rowdeg ? max_rowdeg : a.size() + b.size() - 2;
        c.resize(deg + 1);
        lbx_shim::In<Matrix2> A(a);
        lbx_shim::In<Matrix3> B(b);
        lbx_shim::Out<Matrix1> C(c, false);
        lbx_polmat dc = C.desc(), da = A.desc(), db = B.desc();
        lbx_shim::check(lbx_polymat_mul_u32(lbx_shim::context(), (uint32_t)_p, &dc, &da, &db, max_rowdeg,
                                            LBX_METHOD_AUTO, nullptr));
        C.finish();
    }

    // c = (a.b x^-(hdeg-1)) mod x^pts on c.size() coefficients, hdeg = n0 ? n0 : c.size(),
    // deg = n1 ? n1 : 2 hdeg - 1, pts = 2^ceil(log2 deg); c keeps its size
    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void midproduct(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, bool smallLeft = true, size_t n0 = 0,
                    size_t n1 = 0) const
    {
        if (a.coldim() != b.rowdim()) lbx_shim::check(LBX_ERR_SIZE_MISMATCH);
        lbx_shim::In<Matrix2> A(a);
        lbx_shim::In<Matrix3> B(b);
        lbx_shim::Out<Matrix1> C(c, false);
        lbx_polmat dc = C.desc(), da = A.desc(), db = B.desc();
        lbx_shim::check(lbx_polymat_midproduct_u32(lbx_shim::context(), (uint32_t)_p, &dc, &da, &db,
                                                   smallLeft ? 1 : 0, n0, n1, LBX_METHOD_AUTO, nullptr));
        C.finish();
    }

private:
    const Field *_field;
    uint64_t _p;
};

// PolynomialMatrixFFTPrimeMulDomain<Modular<uint32_t,uint64_t>> (matpoly-mult-fft-wordsize-fast.inl:40-190):
// mul writes coefficients 0..deg of the pre-sized c (deg = max_rowdeg ? max_rowdeg :
// a.size()+b.size()-2; c is not resized, :62-82); mul_fft is the narrower seam on
// polfirst matrices of 2^lpts coefficients (a and b are left unchanged here; the
// reference transforms them in place).  Both throw "bad FFT Prime" when 2^lpts
// does not divide p-1 (:107-112).
template <>
class PolynomialMatrixFFTPrimeMulDomain<Givaro::Modular<uint32_t, uint64_t>> {
public:
    typedef Givaro::Modular<uint32_t, uint64_t> Field;
    typedef PolynomialMatrix<Field, PMType::polfirst> MatrixP;
    typedef PolynomialMatrix<Field, PMType::matfirst> PMatrix;

    explicit PolynomialMatrixFFTPrimeMulDomain(const Field &F) : _field(&F), _p((uint64_t)F.cardinality()) {}
    const Field &field() const { return *_field; }

    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void mul(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, size_t max_rowdeg = 0) const
    {
        if (a.coldim() != b.rowdim()) lbx_shim::check(LBX_ERR_SIZE_MISMATCH);
        const size_t deg = max_rowdeg ? max_rowdeg : a.size() + b.size() - 2;
        size_t pts = 1;
        while (pts <= deg) pts <<= 1;
        if (_p >= ((uint64_t)1 << 31) || (_p - 1) % pts) lbx_shim::check(LBX_ERR_BAD_FFT_PRIME);
        lbx_shim::In<Matrix2> A(a);
        lbx_shim::In<Matrix3> B(b);
        lbx_shim::Out<Matrix1> C(c, true);
        lbx_polmat dc = C.desc(), da = A.desc(), db = B.desc();
        lbx_shim::check(lbx_polymat_mul_u32(lbx_shim::context(), (uint32_t)_p, &dc, &da, &db, deg, LBX_METHOD_AUTO,
                                            nullptr));
        C.finish();
    }

    void mul_fft(size_t lpts, MatrixP &c, MatrixP &a, MatrixP &b) const
    {
        const size_t pts = (size_t)1 << lpts;
        if (a.size() != pts || b.size() != pts || c.size() != pts || a.coldim() != b.rowdim())
            lbx_shim::check(LBX_ERR_SIZE_MISMATCH);
        lbx_shim::check(lbx_polymat_mul_fft_u32(lbx_shim::context(), (uint32_t)_p, (unsigned)lpts, a.rowdim(),
                                                a.coldim(), b.coldim(), a.getPointer(), b.getPointer(),
                                                c.getPointer(), LBX_METHOD_AUTO, nullptr));
    }

private:
    const Field *_field;
    uint64_t _p;
};

// PolynomialMatrixNaiveMulDomain<Modular<uint32_t,uint64_t>> (matpoly-mult-naive.h:39-125):
// mul writes a.size()+b.size()-1 coefficients of the pre-sized c; midproduct
// writes c[k-hdeg+1] = coefficient k of a.b for hdeg-1 <= k < deg and leaves
// the rest of c as it was.
template <>
class PolynomialMatrixNaiveMulDomain<Givaro::Modular<uint32_t, uint64_t>> {
public:
    typedef Givaro::Modular<uint32_t, uint64_t> Field;

    explicit PolynomialMatrixNaiveMulDomain(const Field &F) : _field(&F), _p((uint64_t)F.cardinality()) {}
    const Field &field() const { return *_field; }

    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void mul(Matrix1 &c, const Matrix2 &a, const Matrix3 &b) const
    {
        if (a.coldim() != b.rowdim()) lbx_shim::check(LBX_ERR_SIZE_MISMATCH);
        lbx_shim::In<Matrix2> A(a);
        lbx_shim::In<Matrix3> B(b);
        lbx_shim::Out<Matrix1> C(c, true);
        lbx_polmat dc = C.desc(), da = A.desc(), db = B.desc();
        lbx_shim::check(lbx_polymat_mul_u32(lbx_shim::context(), (uint32_t)_p, &dc, &da, &db, 0, LBX_METHOD_AUTO,
                                            nullptr));
        C.finish();
    }

    template <typename Matrix1, typename Matrix2, typename Matrix3>
    void midproduct(Matrix1 &c, const Matrix2 &a, const Matrix3 &b, bool smallLeft = true, size_t n0 = 0,
                    size_t n1 = 0) const
    {
        (void)smallLeft;  // both branches of :91-123 sum the same coefficient
        if (a.coldim() != b.rowdim()) lbx_shim::check(LBX_ERR_SIZE_MISMATCH);
        lbx_shim::In<Matrix2> A(a);
        lbx_shim::In<Matrix3> B(b);
        lbx_shim::Out<Matrix1> C(c, true);
        lbx_polmat dc = C.desc(), da = A.desc(), db = B.desc();
        lbx_shim::check(lbx_polymat_midproduct_naive_u32(lbx_shim::context(), (uint32_t)_p, &dc, &da, &db, n0, n1,
                                                         nullptr));
        C.finish();
    }

private:
    const Field *_field;
    uint64_t _p;
};

}  // namespace LinBox

#if defined(LBX_LINBOX_MIRROR) || defined(__LINBOX_ORDER_BASIS)
namespace LinBox {

// OrderBasis<Modular<uint32_t,uint64_t>> (order-basis.h:179-600) with the
// default early termination: sigma (m x m) with sigma.serie = 0 mod x^order
// and the shifted row degrees in shift, on the device.  PM_Basis below
// order 32 (MBASIS_THRESHOLD) is M_Basis; above, the halves are joined by a
// middle product and a product on device buffers (:219-326).  sigma is resized
// to order+1 coefficients if it is smaller and to degree+1 on return; both
// return the degree.  The discrepancy's PLUQ is taken in its rank-profile
// form (pivot rows = the row rank profile, placed first, in row order).
template <>
class OrderBasis<Givaro::Modular<uint32_t, uint64_t>, EarlyTerm<(size_t)-1>> {
public:
    typedef Givaro::Modular<uint32_t, uint64_t> Field;
    typedef PolynomialMatrix<Field, PMType::polfirst> MatrixP;
    typedef PolynomialMatrix<Field, PMType::matfirst> PMatrix;

    explicit OrderBasis(const Field &f) : _field(&f) {}
    const Field &field() const { return *_field; }

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
    // :204-215: PM_Basis on a copy of the serie (the device call never writes it)
    template <typename PMatrix1, typename PMatrix2>
    size_t PM_Basis2(PMatrix1 &sigma, const PMatrix2 &serie, size_t order, std::vector<size_t> &shift)
    {
        return run(true, sigma, serie, order, shift);
    }

private:
    template <typename PMatrix1, typename PMatrix2>
    size_t run(bool divide, PMatrix1 &sigma, const PMatrix2 &serie, size_t order, std::vector<size_t> &shift)
    {
        if (shift.size() != serie.rowdim() || sigma.rowdim() != serie.rowdim() || sigma.coldim() != serie.rowdim())
            lbx_shim::check(LBX_ERR_SIZE_MISMATCH);
        if (sigma.size() < order + 1) sigma.resize(order + 1);
        size_t deg = 0;
        lbx_shim::In<PMatrix2> Z(serie);
        lbx_shim::Out<PMatrix1> S(sigma, false);
        lbx_polmat ds = S.desc(), dz = Z.desc();
        const uint32_t p = (uint32_t)field().cardinality();
        if (divide)
            lbx_shim::check(lbx_order_basis_pm_u32(lbx_shim::context(), p, &ds, &dz, order, shift.data(), &deg,
                                                   LBX_METHOD_AUTO, nullptr));
        else
            lbx_shim::check(lbx_order_basis_m_u32(lbx_shim::context(), p, &ds, &dz, order, shift.data(), &deg,
                                                  nullptr));
        sigma.resize(deg + 1);  // then a view's first deg+1 coefficients are written back
        S.finish();
        return deg;
    }

    const Field *_field;
};

}  // namespace LinBox
#endif

#if defined(LBX_LINBOX_MIRROR) || defined(__LINBOX_fft_H)
namespace LinBox {

// FFT<Modular<uint32_t,uint64_t>, S> (fft.h:48-144) for every SIMD choice S:
// FFT_direct is the DIF (natural order in, bit-reversed out) and FFT_inverse
// the DIT (bit-reversed in, natural out, no scaling) with this object's root,
// on one polynomial of 2^k words in host or device memory; w = 0 selects
// compute_primitive_root (fft-utils.h:83-104), and the constructor throws as
// check_args does (fft.h:133-144).
template <typename S>
class FFT<Givaro::Modular<uint32_t, uint64_t>, S> {
public:
    typedef Givaro::Modular<uint32_t, uint64_t> Field;
    typedef typename Field::Element Element;

    FFT(const Field &F, size_t k, Element w = 0) : _fld(&F), _k(k)
    {
        Element none = 0;
        lbx_shim::check(lbx_ntt_forward_u32(lbx_shim::context(), (uint32_t)F.cardinality(), (unsigned)k, w, &none, 0,
                                            &_w, nullptr));
    }
    void FFT_direct(Element *coeffs) const
    {
        lbx_shim::check(lbx_ntt_forward_u32(lbx_shim::context(), (uint32_t)_fld->cardinality(), (unsigned)_k, _w,
                                            coeffs, 1, nullptr, nullptr));
    }
    void FFT_inverse(Element *coeffs) const
    {
        lbx_shim::check(lbx_ntt_inverse_u32(lbx_shim::context(), (uint32_t)_fld->cardinality(), (unsigned)_k, _w,
                                            coeffs, 1, nullptr, nullptr));
    }
    const Field &field() const { return *_fld; }
    size_t size() const { return (size_t)1 << _k; }
    size_t log2_size() const { return _k; }
    const Element &root() const { return _k == 1 ? _fld->mOne : _w; }
    // w^-1 = -w^(n/2-1) (fft.h:119-130)
    Element invroot() const
    {
        if (_k == 1) return _fld->mOne;
        Element r = _fld->one, b = _w;
        for (uint64_t e = ((uint64_t)1 << (_k - 1)) - 1; e; e >>= 1) {
            if (e & 1) _fld->mulin(r, b);
            _fld->mulin(b, b);
        }
        return _fld->negin(r);
    }

private:
    const Field *_fld;
    size_t _k;
    Element _w = 0;
};

}  // namespace LinBox
#endif

#endif  // LBX_LINBOX_SHIM_HPP
