// The drop-in boundary in C++: code written against LinBox's interface,
// compiled against lbx/linbox_mirror.hpp (which stands in for the LinBox
// headers) with the GPU specialisations of lbx/linbox_shim.hpp, checked
// against the CPU oracle.  The test functions have the shape of LinBox's own
// tests: tests/test-matpoly-mult.C:77-144 (mul, midproduct, midproductgen and
// the 48x48x22 / 48x32x42 DLP shape through PolynomialMatrixDomain, here in
// all three storage layouts), tests/test-order-basis.C:72-105 (M_Basis and
// PM_Basis of a matfirst serie) and tests/test-fft.C:157-196 (FFT_direct /
// FFT_inverse against Horner evaluation); plus SubPolynomialMatrix views as
// operands and results, max_rowdeg, and the reference's error behaviour.
// Exit status 0 and "PASS" on success.
#include <cstdio>
#include <vector>

#include "lbx/linbox_mirror.hpp"  // in a LinBox build: the LinBox headers
#include "lbx/linbox_shim.hpp"
#include "oracle_checks.hpp"

using namespace LinBox;

template <typename MatrixP, typename Field, typename RandIter>
bool check_matpol_mul(const Field &fld, RandIter &Gen, size_t n, size_t d)
{
    MatrixP A(fld, n, n, d), B(fld, n, n, d), C(fld, n, n, 2 * d - 1);
    A.random(Gen);
    B.random(Gen);
    typedef PolynomialMatrixDomain<Field> PolMatDom;
    PolMatDom PMD(fld);
    PMD.mul(C, A, B);
    return check_mul(C, A, B, C.size());
}

template <typename MatrixP, typename Field, typename RandIter>
bool check_matpol_midp(const Field &fld, RandIter &Gen, size_t n, size_t d)
{
    MatrixP A(fld, n, n, d), C(fld, n, n, 2 * d - 1);
    MatrixP B(fld, n, n, d);
    A.random(Gen);
    C.random(Gen);
    typedef PolynomialMatrixDomain<Field> PolMatDom;
    PolMatDom PMD(fld);
    PMD.midproduct(B, A, C);
    return check_midproduct(B, A, C);
}

template <typename MatrixP, typename Field, typename RandIter>
bool check_matpol_midpgen(const Field &fld, RandIter &Gen, size_t n, size_t d)
{
    const size_t d1 = d / 2, d0 = d - d1;
    MatrixP A(fld, n, n, d0 + 1), B(fld, n, n, d1), C(fld, n, n, d);
    A.random(Gen);
    C.random(Gen);
    typedef PolynomialMatrixDomain<Field> PolMatDom;
    PolMatDom PMD(fld);
    PMD.midproductgen(B, A, C, true, d0 + 1, d);
    return check_midproduct(B, A, C, true, d0 + 1, d);
}

template <typename MatrixP, typename Field, typename RandIter>
bool debug_midpgen_dlp(const Field &fld, RandIter &Gen)
{
    const size_t n0 = 22, n1 = 42;
    MatrixP A(fld, 48, 48, 22), B(fld, 48, 32, 21), C(fld, 48, 32, 42);
    A.random(Gen);
    C.random(Gen);
    typedef PolynomialMatrixDomain<Field> PolMatDom;
    PolMatDom PMD(fld);
    PMD.midproductgen(B, A, C, true, n0, n1);
    return check_midproduct(B, A, C, true, n0, n1);
}

template <typename Field>
bool launchTest(const Field &F, size_t n, long d, long seed)
{
    bool ok = true;
    typename Field::RandIter G(F, seed);
    std::printf("Polynomial matrix testing over ");
    F.write(std::cout) << std::endl;
    ok &= check_matpol_mul<PolynomialMatrix<Field, PMType::polfirst>>(F, G, n, d);
    ok &= check_matpol_mul<PolynomialMatrix<Field, PMType::matfirst>>(F, G, n, d);
    ok &= check_matpol_mul<PolynomialMatrix<Field, PMType::matrowfirst>>(F, G, n, d);
    ok &= check_matpol_midp<PolynomialMatrix<Field, PMType::polfirst>>(F, G, n, d);
    ok &= check_matpol_midp<PolynomialMatrix<Field, PMType::matfirst>>(F, G, n, d);
    ok &= check_matpol_midpgen<PolynomialMatrix<Field, PMType::polfirst>>(F, G, n, d);
    ok &= check_matpol_midpgen<PolynomialMatrix<Field, PMType::matrowfirst>>(F, G, n, d);
    ok &= debug_midpgen_dlp<PolynomialMatrix<Field, PMType::polfirst>>(F, G);
    return ok;
}

template <typename Field, typename RandIter>
bool check_sigma_all(const Field &F, RandIter &Gen, size_t m, size_t n, size_t d)
{
    typedef PolynomialMatrix<Field, PMType::matfirst> MatrixP;
    MatrixP Serie(F, m, n, d);
    MatrixP Sigma1(F, m, m, d + 1), Sigma3(F, m, m, d + 1);
    for (size_t k = 0; k < d; ++k)
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) Gen.random(Serie.ref(i, j, k));
    std::vector<size_t> shift(m, 0);
    std::vector<size_t> shift0(shift), shift3(shift);
    OrderBasis<Field> SB(F);
    bool passed = true;
    SB.M_Basis(Sigma3, Serie, d, shift3);
    passed &= check_sigma(Sigma3, Serie, d, shift0, shift3, false);
    SB.PM_Basis(Sigma1, Serie, d, shift);
    passed &= check_sigma(Sigma1, Serie, d, shift0, shift, true);
    // a polfirst sigma, a serie read through a view, and a nonzero starting shift
    PolynomialMatrix<Field, PMType::polfirst> Sigma2(F, m, m, 1), Big(F, m, n, d + 7);
    Big.random(Gen);
    auto V = Big.at(3, 3 + d - 1);
    std::vector<size_t> sh2(m), sh20;
    for (size_t i = 0; i < m; ++i) sh2[i] = (i * 7) % 5;
    sh20 = sh2;
    SB.PM_Basis(Sigma2, V, d, sh2);
    passed &= check_sigma(Sigma2, V, d, sh20, sh2, true);
    return passed;
}

template <typename Field, typename RandIter>
bool check_views(const Field &F, RandIter &Gen)
{
    bool ok = true;
    // operands are windows of larger matrices of each layout, the result a window too
    PolynomialMatrix<Field, PMType::matfirst> Abig(F, 5, 6, 60);
    PolynomialMatrix<Field, PMType::matrowfirst> Bbig(F, 6, 4, 50);
    PolynomialMatrix<Field, PMType::polfirst> Cbig(F, 5, 4, 100);
    Abig.random(Gen);
    Bbig.random(Gen);
    Cbig.random(Gen);
    const PolynomialMatrix<Field, PMType::polfirst> Cold(Cbig);
    auto A = Abig.at(7, 36);  // 30 coefficients
    auto B = Bbig.at(2, 41);  // 40 coefficients
    auto C = Cbig.at(20, 99);  // room for 80 >= 69
    PolynomialMatrixDomain<Field> PMD(F);
    PMD.mul(C, A, B);
    ok &= C.size() == 69;
    ok &= check_mul(C, A, B, C.size());
    // storage outside the window is untouched
    bool kept = true;
    for (size_t i = 0; i < 20; ++i)
        for (size_t k = 0; k < 20; ++k) kept = kept && Cbig.get(i, k) == Cold.get(i, k);
    for (size_t i = 0; i < 20; ++i)
        for (size_t k = 89; k < 100; ++k) kept = kept && Cbig.get(i, k) == Cold.get(i, k);
    ok &= lbx_test::report(kept, "view output", "storage outside the window");
    // a view of a view as operand of a middle product
    auto A2 = A.at(5, 24);  // 20 coefficients
    PolynomialMatrix<Field, PMType::matfirst> Mid(F, 5, 4, 20);
    PMD.midproduct(Mid, A2, B);
    ok &= check_midproduct(Mid, A2, B);
    // a view cannot grow: mul into a window smaller than the product throws
    auto small = Cbig.at(0, 9);
    bool threw = false;
    try {
        PMD.mul(small, A, B);
    } catch (const LinboxError &) {
        threw = true;
    }
    ok &= lbx_test::report(threw, "view resize", "larger than available throws");
    return ok;
}

template <typename Field, typename RandIter>
bool check_max_rowdeg_and_errors(const Field &F, RandIter &Gen)
{
    bool ok = true;
    PolynomialMatrix<Field, PMType::polfirst> A(F, 3, 4, 40), B(F, 4, 2, 30), C(F, 3, 2, 1);
    A.random(Gen);
    B.random(Gen);
    PolynomialMatrixFFTMulDomain<Field> FD(F);
    FD.mul(C, A, B, 50);  // 51 coefficients of the 69-coefficient product mod x^64 - 1
    // reference: the oracle's full product folded mod x^64 - 1
    PolynomialMatrix<Field, PMType::polfirst> Full(F, 3, 2, 69);
    FD.mul(Full, A, B);
    bool same = C.size() == 51;
    for (size_t i = 0; same && i < 6; ++i)
        for (size_t k = 0; k < 51; ++k) {
            uint32_t v = Full.get(i, k);
            if (k + 64 < 69) F.addin(v, Full.get(i, k + 64));
            same = same && C.get(i, k) == v;
        }
    ok &= lbx_test::report(same, "max_rowdeg", "3x4[40].4x2[30] deg 50 (wrap mod x^64-1)");
    ok &= check_mul(Full, A, B, Full.size());
    bool threw = false;
    try {
        PolynomialMatrix<Field, PMType::polfirst> Bad(F, 3, 2, 40);
        FD.mul(C, A, Bad);  // a.coldim() != b.rowdim()
    } catch (const LinboxError &) {
        threw = true;
    }
    ok &= lbx_test::report(threw, "dimension mismatch", "throws");
    return ok;
}

// PolynomialMatrixFFTPrimeMulDomain::mul_fft on padded polfirst operands of
// 2^lpts coefficients: the cyclic product mod x^pts - 1 (the oracle's product
// truncated by max_rowdeg = pts - 1); the "bad FFT Prime" error otherwise.
template <typename Field, typename RandIter>
bool check_mul_fft(const Field &F, RandIter &Gen, size_t lpts)
{
    typedef PolynomialMatrixFFTPrimeMulDomain<Field> Dom;
    typedef typename Dom::MatrixP MatrixP;
    const size_t pts = (size_t)1 << lpts;
    MatrixP A(F, 3, 5, pts), B(F, 5, 2, pts), C(F, 3, 2, pts);
    A.random(Gen);
    B.random(Gen);
    Dom D(F);
    D.mul_fft(lpts, C, A, B);
    const uint32_t p = (uint32_t)F.cardinality();
    const std::vector<uint32_t> a = lbx_test::polfirst(A), b = lbx_test::polfirst(B);
    std::vector<uint32_t> ref(6 * pts);
    bool ok = or_polymat_mul_fft_prime(p, 3, 5, 2, a.data(), pts, b.data(), pts, ref.data(), pts - 1) == 0;
    ok = ok && lbx_test::polfirst(C) == ref;
    bool threw = false;
    try {
        Field G(1000000007u);  // (p-1) has 2-valuation 1
        PolynomialMatrixFFTPrimeMulDomain<Field> Bad(G);
        MatrixP X(G, 1, 1, 8), Y(G, 1, 1, 8), Z(G, 1, 1, 8);
        Bad.mul_fft(3, Z, X, Y);
    } catch (const LinboxError &) {
        threw = true;
    }
    char shape[32];
    std::snprintf(shape, sizeof shape, "2^%zu points", lpts);
    return lbx_test::report(ok && threw, "mul_fft", shape);
}

// Horner evaluation at w^i: every point up to 2^10 points, 64 sampled points above.
template <typename Field, typename RandIter>
bool check_fft(const Field &F, RandIter &Gen, size_t k)
{
    typedef typename Field::Element Elt;
    const size_t n = (size_t)1 << k;
    std::vector<Elt> in(n), in_br(n);
    FFT<Field, NoSimd<Elt>> fft(F, k);
    for (auto &v : in) Gen.random(v);
    const Elt w = fft.root();
    auto br = [&](size_t i) {
        size_t ib = 0;
        for (size_t j = 0, t = i; j < k; ++j, t >>= 1) ib = (ib << 1) | (t & 1);
        return ib;
    };
    for (size_t i = 0; i < n; ++i) in_br[br(i)] = in[i];
    std::vector<size_t> pts;
    for (size_t i = 0; i < n && (k <= 10 || pts.size() < 64); ++i) {
        Elt r;
        pts.push_back(k <= 10 ? i : (size_t)(Gen.random(r) % n));
    }
    std::vector<Elt> y(in), z(in_br);
    fft.FFT_direct(y.data());
    fft.FFT_inverse(z.data());
    bool ok = true;
    for (size_t i : pts) {
        Elt x = F.one, b = w;  // x = w^i
        for (size_t e = i; e; e >>= 1) {
            if (e & 1) F.mulin(x, b);
            F.mulin(b, b);
        }
        Elt acc = F.zero;
        for (size_t t = n; t-- > 0;) {
            F.mulin(acc, x);
            F.addin(acc, in[t]);
        }
        ok = ok && y[br(i)] == acc && z[i] == acc;
    }
    Elt wi = fft.invroot(), one = F.one;
    F.mulin(wi, w);
    ok = ok && wi == one;
    char shape[32];
    std::snprintf(shape, sizeof shape, "2^%zu points", k);
    return lbx_test::report(ok, "FFT_direct / FFT_inverse", shape);
}

int main()
{
    bool ok = true;
    typedef Givaro::Modular<uint32_t, uint64_t> Field;
    for (uint32_t p : {998244353u, 2013265921u, 469762049u}) {
        Field F(p);
        ok &= launchTest(F, 16, 300, 11);
        typename Field::RandIter G(F, 7);
        ok &= check_sigma_all(F, G, 6, 3, 80);
        ok &= check_views(F, G);
        ok &= check_max_rowdeg_and_errors(F, G);
        for (size_t k : {1, 5, 10, 14, 17}) ok &= check_fft(F, G, k);
        ok &= check_mul_fft(F, G, 9);
    }
    {  // a word prime that is not an FFT prime for these lengths (the few-primes route)
        Field F(1000000007u);
        ok &= launchTest(F, 8, 100, 5);
        typename Field::RandIter G(F, 9);
        ok &= check_sigma_all(F, G, 4, 2, 40);
    }
    {  // FFT: k must be positive; w must be a primitive root
        Field F(998244353u);
        int thrown = 0;
        try {
            FFT<Field, NoSimd<uint32_t>> f(F, 0);
        } catch (const LinboxError &) {
            ++thrown;
        }
        try {
            FFT<Field, NoSimd<uint32_t>> f(F, 3, 998244352u);
        } catch (const LinboxError &) {
            ++thrown;
        }
        ok &= lbx_test::report(thrown == 2, "FFT check_args", "throws");
    }
    std::printf(ok ? "PASS\n" : "FAIL\n");
    return ok ? 0 : 1;
}
