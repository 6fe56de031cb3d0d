// linbox_mirror.hpp -- the part of LinBox's (and Givaro's) C++ interface that
// the FFT polynomial-matrix product and its callers use, with the reference's
// names, template shapes and storage layouts.  It stands in for the LinBox
// headers where they are not installed, so that code written against LinBox
// -- the drop-in specialisations of lbx/linbox_shim.hpp and test programs
// shaped like LinBox's own tests -- compiles against either unchanged.
//
//   Givaro::Modular<uint32_t,uint64_t>          the field (Givaro 4.2.1 interface subset)
//   LinBox::PMType, PolynomialMatrix<F,PMType>  linbox/matrix/polynomial-matrix.h:54-910
//   LinBox::SubPolynomialMatrix<MatPoly>        polynomial-matrix.h:940-1036
//   PolynomialMatrixFFTMulDomain<F>,
//   PolynomialMatrixFFTPrimeMulDomain<F>,
//   PolynomialMatrixNaiveMulDomain<F>           declared only: lbx/linbox_shim.hpp specialises
//                                               them for Givaro::Modular<uint32_t,uint64_t>
//   PolynomialMatrixMulDomain<F>, PolynomialMatrixDomain<F>
//                                               polynomial-matrix-domain.h:40-119 (dispatch)
//   EarlyTerm<K>, OrderBasis<F,ET>              order-basis.h:152-181 (OrderBasis declared only)
//   FFT<F,Simd>                                 fft.h:48-144 (declared only)
//
// Storage, view windows, resize and copy behave as in the reference; all of
// it is host code.  Products and order bases come from the shim, which calls
// the C ABI of lbx/lbx.h.
#ifndef LBX_LINBOX_MIRROR_HPP
#define LBX_LINBOX_MIRROR_HPP

#define LBX_LINBOX_MIRROR 1

#include <cstddef>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

// fflas-ffpack's SIMD tags (fflas_simd.h), the second parameter of LinBox::FFT.
template <class T>
struct NoSimd {};
template <class T>
struct Simd {};

namespace Givaro {

template <class Storage_t, class Compute_t = Storage_t>
class Modular;

// Z/pZ with uint32_t elements and 64-bit intermediate products.
template <>
class Modular<uint32_t, uint64_t> {
public:
    typedef uint32_t Element;
    typedef uint32_t Residu_t;
    typedef Modular<uint32_t, uint64_t> Self_t;

    explicit Modular(Residu_t p) : _p(p), zero(0), one(1 % p), mOne(p - 1) {}
    Residu_t cardinality() const { return _p; }
    Residu_t characteristic() const { return _p; }
    Element &init(Element &x, uint64_t v) const { return x = (Element)(v % _p); }
    Element &assign(Element &x, Element y) const { return x = y; }
    Element &add(Element &r, Element a, Element b) const { return r = (Element)(((uint64_t)a + b) % _p); }
    Element &sub(Element &r, Element a, Element b) const { return r = (Element)(((uint64_t)a + _p - b) % _p); }
    Element &mul(Element &r, Element a, Element b) const { return r = (Element)((uint64_t)a * b % _p); }
    Element &addin(Element &r, Element a) const { return add(r, r, a); }
    Element &mulin(Element &r, Element a) const { return mul(r, r, a); }
    Element &negin(Element &r) const { return r = r ? _p - r : 0; }
    Element &axpyin(Element &r, Element a, Element x) const
    {
        return r = (Element)(((uint64_t)a * x + r) % _p);
    }
    bool areEqual(Element a, Element b) const { return a == b; }
    bool isZero(Element a) const { return a == 0; }
    std::ostream &write(std::ostream &os) const { return os << "Modular<uint32_t,uint64_t> modulo " << _p; }

    // uniform elements in [0, min(p, size)) from a seeded splitmix64 stream
    class RandIter {
    public:
        typedef Self_t::Residu_t Residu_t;
        RandIter(const Self_t &F, uint64_t seed = 0, Residu_t size = 0)
            : _F(&F), _state(seed ^ 0x9E3779B97F4A7C15ull), _size(size && size < F.cardinality() ? size : F.cardinality())
        {
        }
        Element &random(Element &x) const
        {
            uint64_t z = (_state += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return x = (Element)(z % _size);
        }
        const Self_t &ring() const { return *_F; }

    private:
        const Self_t *_F;
        mutable uint64_t _state;
        Residu_t _size;
    };

private:
    Residu_t _p;

public:
    const Element zero, one, mOne;
};

}  // namespace Givaro

namespace LinBox {

// linbox/util/error.h: the reference throws LinboxError(const char*).
class LinboxError : public std::runtime_error {
public:
    explicit LinboxError(const char *msg) : std::runtime_error(msg) {}
    explicit LinboxError(const std::string &msg) : std::runtime_error(msg) {}
};

enum PMType { polfirst, matfirst, matrowfirst };

template <class MatPoly>
class SubPolynomialMatrix;

// A polynomial matrix in F[x]^(r x c) of size s (degree s-1) in one of the
// three storage layouts.  Coefficient k of entry i = row*c + col lives at
//   polfirst     rep[i*s + k]                                (:131-137)
//   matfirst     rep[i + k*r*c]                              (:422-428)
//   matrowfirst  rep[row*c*s + col + k*c]                    (:660-683)
template <class _Field, PMType T>
class PolynomialMatrix {
public:
    typedef _Field Field;
    typedef typename Field::Element Element;
    typedef PolynomialMatrix<Field, T> Self_t;
    typedef SubPolynomialMatrix<Self_t> view;
    typedef SubPolynomialMatrix<const Self_t> const_view;

    PolynomialMatrix(const Field &f, size_t r, size_t c, size_t s)
        : _fld(&f), _rep(r * c * s, f.zero), _row(r), _col(c), _size(s)
    {
    }

    size_t rowdim() const { return _row; }
    size_t coldim() const { return _col; }
    size_t degree() const { return _size - 1; }
    size_t size() const { return _size; }
    const Field &field() const { return *_fld; }
    Element *getPointer() { return _rep.data(); }
    const Element *getPointer() const { return _rep.data(); }

    Element &ref(size_t i, size_t k) { return _rep[index(i, k)]; }
    Element get(size_t i, size_t k) const { return _rep[index(i, k)]; }
    Element &ref(size_t i, size_t j, size_t k) { return ref(i * _col + j, k); }
    Element get(size_t i, size_t j, size_t k) const { return get(i * _col + j, k); }

    view at(size_t i, size_t j) { return view(*this, i, j); }
    const_view at(size_t i, size_t j) const { return const_view(*this, i, j); }

    template <class RandIter>
    void random(RandIter &G)
    {
        for (auto &x : _rep) G.random(x);
    }

    // new size s, coefficients [0, min(s, size())) kept, new ones zero (:184-203)
    void resize(size_t s)
    {
        if (s == _size) return;
        std::vector<Element> rep(_row * _col * s, _fld->zero);
        const size_t keep = s < _size ? s : _size;
        for (size_t i = 0; i < _row * _col; ++i)
            for (size_t k = 0; k < keep; ++k) rep[index_for(i, k, s)] = _rep[index(i, k)];
        _rep.swap(rep);
        _size = s;
    }

    // coefficients beg..end of M into start.. of this matrix, any layout
    // (:206-243, :508-532, :763-790); through a view: at the view's shift (:300-303)
    template <PMType U>
    void copy(const PolynomialMatrix<Field, U> &M, size_t beg, size_t end, size_t start = 0)
    {
        for (size_t i = 0; i < _row * _col; ++i)
            for (size_t k = beg; k <= end; ++k) ref(i, start + k - beg) = M.get(i, k);
    }
    template <class MatPoly>
    void copy(const SubPolynomialMatrix<MatPoly> &M, size_t beg, size_t end, size_t start = 0)
    {
        copy(*M._ptr, beg + M._shift, end + M._shift, start);
    }
    template <class Mat>
    void copy(const Mat &M)
    {
        copy(M, 0, M.size() - 1);
    }

private:
    size_t index_for(size_t i, size_t k, size_t s) const
    {
        if (T == polfirst) return i * s + k;
        if (T == matfirst) return i + k * _row * _col;
        return (i / _col) * _col * s + (i % _col) + k * _col;
    }
    size_t index(size_t i, size_t k) const { return index_for(i, k, _size); }

    const Field *_fld;
    std::vector<Element> _rep;
    size_t _row, _col, _size;
};

// The degree window [i, j] of a polynomial matrix (:940-1036): it reads and
// writes the underlying storage and can only shrink.
template <class MatPoly>
class SubPolynomialMatrix {
public:
    typedef SubPolynomialMatrix<MatPoly> Self_t;
    typedef typename MatPoly::Field Field;
    typedef typename MatPoly::Element Element;
    typedef SubPolynomialMatrix<MatPoly> view;
    typedef SubPolynomialMatrix<const MatPoly> const_view;

    SubPolynomialMatrix(MatPoly &M, size_t i, size_t j) : _ptr(&M), _size(j - i + 1), _shift(i)
    {
        if (!(i < M.size() && i <= j && j < M.size())) throw LinboxError("SubPolynomialMatrix: invalid window");
    }
    SubPolynomialMatrix(Self_t &M, size_t i, size_t j) : _ptr(M._ptr), _size(j - i + 1), _shift(i + M._shift)
    {
        if (!(i < M.size() && i <= j && j < M.size())) throw LinboxError("SubPolynomialMatrix: invalid window");
    }

    void resize(size_t s)
    {
        if (s <= _size)
            _size = s;
        else
            throw LinboxError("SubPolynomialMatrix: resizing with larger size than available, exiting ...");
    }
    template <typename Mat>
    void copy(const Mat &M, size_t beg, size_t end)
    {
        _ptr->copy(M, beg, end, _shift);
    }
    template <typename Mat>
    void copy(const Mat &M)
    {
        copy(M, 0, M.size() - 1);
    }

    Element get(size_t i, size_t k) const { return _ptr->get(i, k + _shift); }
    Element get(size_t i, size_t j, size_t k) const { return get(i * coldim() + j, k); }
    size_t rowdim() const { return _ptr->rowdim(); }
    size_t coldim() const { return _ptr->coldim(); }
    size_t degree() const { return _size - 1; }
    size_t size() const { return _size; }
    const Field &field() const { return _ptr->field(); }
    const_view at(size_t i, size_t j) const { return const_view(*_ptr, i + _shift, j + _shift); }
    view at(size_t i, size_t j) { return view(*_ptr, i + _shift, j + _shift); }

    template <class F, PMType>
    friend class PolynomialMatrix;

protected:
    MatPoly *_ptr;
    size_t _size;
    size_t _shift;
};

// Domains (the shim specialises the two transform-level ones for
// Givaro::Modular<uint32_t,uint64_t>).
template <class Field>
class PolynomialMatrixFFTMulDomain;
template <class Field>
class PolynomialMatrixFFTPrimeMulDomain;
template <class Field>
class PolynomialMatrixNaiveMulDomain;

#define LBX_MIRROR_FFT_DEG_THRESHOLD 2  // polynomial-matrix-domain.h:29-30

// PolynomialMatrixMulDomain (polynomial-matrix-domain.h:40-104): mul takes the
// FFT domain when a.size()+b.size() > 2 (max_rowdeg is not forwarded, :61),
// the naive one otherwise; midproduct the FFT domain when b.size() > 2;
// midproductgen the naive domain when c.size() <= 4.
template <class _Field>
class PolynomialMatrixMulDomain {
public:
    typedef _Field Field;
    PolynomialMatrixFFTMulDomain<Field> _fft;
    PolynomialMatrixNaiveMulDomain<Field> _naive;
    const Field *_field;

    explicit PolynomialMatrixMulDomain(const Field &F) : _fft(F), _naive(F), _field(&F) {}
    const Field &field() const { return *_field; }

    template <class PMatrix1, class PMatrix2, class PMatrix3>
    void mul(PMatrix1 &c, const PMatrix2 &a, const PMatrix3 &b, size_t max_rowdeg = 0) const
    {
        (void)max_rowdeg;
        if (a.size() + b.size() > LBX_MIRROR_FFT_DEG_THRESHOLD)
            _fft.mul(c, a, b);
        else
            _naive.mul(c, a, b);
    }
    template <class PMatrix1, class PMatrix2, class PMatrix3>
    void midproduct(PMatrix1 &c, const PMatrix2 &a, const PMatrix3 &b) const
    {
        if (b.size() > LBX_MIRROR_FFT_DEG_THRESHOLD)
            _fft.midproduct(c, a, b);
        else
            _naive.midproduct(c, a, b);
    }
    template <class PMatrix1, class PMatrix2, class PMatrix3>
    void midproductgen(PMatrix1 &c, const PMatrix2 &a, const PMatrix3 &b, bool smallLeft = true, size_t n0 = 0,
                       size_t n1 = 0) const
    {
        if (c.size() <= 4)
            _naive.midproduct(c, a, b, smallLeft, n0, n1);
        else
            _fft.midproduct(c, a, b, smallLeft, n0, n1);
    }
};

// polynomial-matrix-domain.h:106-119 (the addition part is not on this path).
template <class Field>
class PolynomialMatrixDomain : public PolynomialMatrixMulDomain<Field> {
public:
    explicit PolynomialMatrixDomain(const Field &F) : PolynomialMatrixMulDomain<Field>(F) {}
};

// order-basis.h:152-176: the default early-termination policy never stops early.
template <size_t K>
class EarlyTerm {
public:
    void reset() {}
};

template <class Field, class ET = EarlyTerm<(size_t)-1>>
class OrderBasis;

// fft.h:48-144 (declared only; the shim specialises it).
template <typename Field, typename S = ::Simd<typename Field::Element>>
class FFT;

}  // namespace LinBox

#endif  // LBX_LINBOX_MIRROR_HPP
