// C++ counterpart of the reference's hot-path tests, written against the
// header-only mirror include/lbx/polymat.hpp:
//  * tests/test-fft.C:157-196 -- FFT_direct / FFT_inverse against Horner
//    evaluation at powers of the root (bit-reversed orders);
//  * tests/test-matpoly-mult.C:77-88,171-244 -- mul on 16x16 polynomial
//    matrices of size 512, checked at x = 1 (check_mul, matpoly-mult-fft.h:122-163)
//    and at random points (a check a coefficient permutation cannot pass);
//  * tests/test-polynomial-matrix.C:224-366 -- the same product in every
//    layout and through views gives the same result;
//  * error behaviour: dimension mismatch and enlarging a view throw.
// Exit status 0 and "PASS" on success.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "lbx/polymat.hpp"

using namespace lbx;

static int failures = 0;
#define EXPECT(cond, msg)                                  \
    do {                                                   \
        if (!(cond)) {                                     \
            std::printf("FAIL: %s (line %d)\n", msg, __LINE__); \
            ++failures;                                    \
        }                                                  \
    } while (0)


static size_t bitrev(size_t i, size_t n)
{
    size_t r = 0;
    for (size_t j = 0; j < n; ++j, i >>= 1) r = (r << 1) | (i & 1);
    return r;
}

static void test_fft(uint32_t p, std::mt19937_64 &g)
{
    Modular F(p);
    for (size_t k = 1; k <= 10; ++k) {
        const size_t n = (size_t)1 << k;
        FFT fft(F, k);
        const uint64_t w = fft.root();
        std::vector<uint32_t> in(n), in_br(n), out(n), out_br(n);
        for (auto &v : in) v = (uint32_t)(g() % p);
        uint64_t x = 1;
        for (size_t i = 0; i < n; ++i, x = x * w % p) {
            uint64_t acc = 0;
            for (size_t t = n; t-- > 0;) acc = (acc * x + in[t]) % p;
            out[i] = (uint32_t)acc;
            out_br[bitrev(i, k)] = out[i];
            in_br[bitrev(i, k)] = in[i];
        }
        std::vector<uint32_t> y(in);
        fft.FFT_direct(y.data());
        EXPECT(y == out_br, "FFT_direct vs Horner");
        y = in_br;
        fft.FFT_inverse(y.data());
        EXPECT(y == out, "FFT_inverse (DIT with the same root) vs Horner");
        FFT inv(F, k, fft.invroot());
        y = in;
        fft.FFT_direct(y.data());
        inv.FFT_inverse(y.data());
        bool ok = true;
        for (size_t i = 0; i < n; ++i) ok &= y[i] == (uint64_t)in[i] * n % p;
        EXPECT(ok, "inverse(direct(x)) == n x");
    }
}

template <class M>
static std::vector<uint64_t> eval(const M &A, uint64_t x, uint64_t p)
{
    std::vector<uint64_t> r(A.rowdim() * A.coldim());
    for (size_t e = 0; e < r.size(); ++e) {
        uint64_t acc = 0;
        for (size_t t = A.size(); t-- > 0;) acc = (acc * x + A.get(e, t)) % p;
        r[e] = acc;
    }
    return r;
}

template <class MC, class MA, class MB>
static bool check_at(const MC &C, const MA &A, const MB &B, uint64_t x, uint64_t p)
{
    const size_t m = A.rowdim(), k = A.coldim(), n = B.coldim();
    auto a = eval(A, x, p), b = eval(B, x, p), c = eval(C, x, p);
    for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < n; ++j) {
            uint64_t acc = 0;
            for (size_t q = 0; q < k; ++q) acc = (acc + a[i * k + q] * b[q * n + j]) % p;
            if (acc != c[i * n + j]) return false;
        }
    return true;
}

template <PMType T>
static void fill(PolynomialMatrix<T> &A, uint32_t p, std::mt19937_64 &g)
{
    for (size_t e = 0; e < A.rowdim() * A.coldim(); ++e)
        for (size_t t = 0; t < A.size(); ++t) A.ref(e, t) = (uint32_t)(g() % p);
}

template <PMType TA, PMType TB, PMType TC>
static void test_mul(uint32_t p, size_t n, size_t d, int method, std::mt19937_64 &g)
{
    Modular F(p);
    PolynomialMatrix<TA> A(F, n, n, d);
    PolynomialMatrix<TB> B(F, n, n, d);
    PolynomialMatrix<TC> C(F, n, n, 1);
    fill(A, p, g);
    fill(B, p, g);
    PolynomialMatrixMulDomain PMD(F, nullptr, method);
    PMD.mul(C, A, B);
    EXPECT(C.size() == 2 * d - 1, "c resized to deg+1");
    EXPECT(check_at(C, A, B, 1, p), "check_mul at x=1");
    EXPECT(check_at(C, A, B, g() % p, p), "check at a random point");
    EXPECT(check_at(C, A, B, g() % p, p), "check at a random point");
    // the same product in polfirst must agree coefficient by coefficient
    PolynomialMatrix<polfirst> A2(F, n, n, d), B2(F, n, n, d), C2(F, n, n, 1);
    for (size_t e = 0; e < n * n; ++e)
        for (size_t t = 0; t < d; ++t) {
            A2.ref(e, t) = A.get(e, t);
            B2.ref(e, t) = B.get(e, t);
        }
    PMD.mul(C2, A2, B2);
    bool same = true;
    for (size_t e = 0; e < n * n; ++e)
        for (size_t t = 0; t < C.size(); ++t) same &= C.get(e, t) == C2.get(e, t);
    EXPECT(same, "layouts give the same product");
}

static void test_views_and_errors(uint32_t p, std::mt19937_64 &g)
{
    Modular F(p);
    PolynomialMatrix<matfirst> A(F, 3, 4, 50);
    PolynomialMatrix<polfirst> B(F, 4, 2, 40);
    fill(A, p, g);
    fill(B, p, g);
    auto Av = A.at(5, 24);
    auto Bv = B.at(3, 33);
    PolynomialMatrix<matrowfirst> Cbig(F, 3, 2, 80);
    auto Cv = Cbig.at(10, 79);
    PolynomialMatrixMulDomain PMD(F);
    PMD.mul(Cv, Av, Bv);
    EXPECT(Cv.size() == 50, "view resized down to deg+1");
    EXPECT(check_at(Cv, Av, Bv, g() % p, p), "view product at a random point");
    bool threw = false;
    try {
        auto small = Cbig.at(0, 20);
        PMD.mul(small, Av, Bv);
    } catch (const LinboxError &e) {
        threw = e.status == LBX_ERR_SIZE_MISMATCH;
    }
    EXPECT(threw, "enlarging a view throws");
    threw = false;
    try {
        PolynomialMatrix<polfirst> X(F, 2, 3, 4), Y(F, 2, 2, 4), Z(F, 2, 2, 7);
        PMD.mul(Z, X, Y);
    } catch (const LinboxError &e) {
        threw = e.status == LBX_ERR_SIZE_MISMATCH;
    }
    EXPECT(threw, "dimension mismatch throws");
}

// midproduct(c, a, b) with the defaults (hdeg = c.size(), deg = 2*hdeg-1) is
// coefficients hdeg-1 .. 2*hdeg-2 of a*b (the check of matpoly-mult-fft.h:168-215
// in closed form); midproductgen with c.size() <= 4 leaves the coefficients
// past the middle product untouched (matpoly-mult-naive.h:85-125).
static void test_midproduct(uint32_t p, std::mt19937_64 &g)
{
    Modular F(p);
    const size_t hdeg = 40;
    PolynomialMatrix<polfirst> A(F, 3, 5, hdeg), B(F, 5, 4, 2 * hdeg - 1), C(F, 3, 4, hdeg), P(F, 3, 4, 1);
    fill(A, p, g);
    fill(B, p, g);
    PolynomialMatrixMulDomain PMD(F);
    PMD.midproduct(C, A, B);
    PMD.mul(P, A, B);
    bool same = C.size() == hdeg;
    for (size_t e = 0; e < 12; ++e)
        for (size_t t = 0; t < hdeg; ++t) same &= C.get(e, t) == P.get(e, t + hdeg - 1);
    EXPECT(same, "midproduct = middle coefficients of the product");
    PolynomialMatrix<matfirst> Cs(F, 3, 4, 4);
    for (size_t e = 0; e < 12; ++e)
        for (size_t t = 0; t < 4; ++t) Cs.ref(e, t) = 7;
    PolynomialMatrix<polfirst> As(F, 3, 5, 2), Bs(F, 5, 4, 3), Ps(F, 3, 4, 1);
    fill(As, p, g);
    fill(Bs, p, g);
    PMD.midproductgen(Cs, As, Bs, true, 2, 0);  // hdeg 2, deg 3: coefficients 1..2 of a*b
    PMD.mul(Ps, As, Bs);
    bool ok = true;
    for (size_t e = 0; e < 12; ++e) {
        ok &= Cs.get(e, 0) == Ps.get(e, 1) && Cs.get(e, 1) == Ps.get(e, 2);
        ok &= Cs.get(e, 2) == 7 && Cs.get(e, 3) == 7;
    }
    EXPECT(ok, "naive midproduct writes only its window");
}

// check_sigma of tests/test-order-basis.C:19-45: sigma.serie = 0 mod x^order
// and sigma != 0, for M_Basis and PM_Basis on an 8 x 4 serie of order 80.
static void test_order_basis(uint32_t p, std::mt19937_64 &g)
{
    Modular F(p);
    const size_t m = 8, n = 4, d = 80;
    PolynomialMatrix<matfirst> Serie(F, m, n, d);
    fill(Serie, p, g);
    PolynomialMatrixMulDomain PMD(F);
    OrderBasis SB(F);
    for (int variant = 0; variant < 2; ++variant) {
        PolynomialMatrix<matfirst> Sigma(F, m, m, d + 1);
        std::vector<size_t> shift(m, 0);
        const size_t deg = variant ? SB.PM_Basis(Sigma, Serie, d, shift) : SB.M_Basis(Sigma, Serie, d, shift);
        PolynomialMatrix<polfirst> T(F, m, n, 1);
        PMD.mul(T, Sigma, Serie);
        bool zero = Sigma.size() == deg + 1, nonzero = false;
        for (size_t e = 0; e < m * n; ++e)
            for (size_t t = 0; t < d; ++t) zero &= T.get(e, t) == 0;
        for (size_t e = 0; e < m * m; ++e)
            for (size_t t = 0; t <= deg; ++t) nonzero |= Sigma.get(e, t) != 0;
        EXPECT(zero && nonzero, variant ? "PM_Basis: sigma.serie = 0 mod x^order" : "M_Basis: sigma.serie = 0 mod x^order");
    }
}

// PolynomialMatrixThreePrimesFFTMulDomain: primes above 29 bits are refused
// (three-primes.inl:57-64); a non-FFT prime takes the few-primes route.
static void test_three_primes(std::mt19937_64 &g)
{
    bool threw = false;
    try {
        Modular big(1000000007u);
        PolynomialMatrixThreePrimesFFTMulDomain D(big);
    } catch (const LinboxError &e) {
        threw = e.status == LBX_ERR_PRIME_TOO_LARGE;
    }
    EXPECT(threw, "three-primes domain refuses a 30-bit prime");
    const uint32_t p = 1048573u;  // p-1 = 2^2 * 262143
    Modular F(p);
    PolynomialMatrix<matfirst> A(F, 6, 5, 300);
    PolynomialMatrix<polfirst> B(F, 5, 4, 200);
    PolynomialMatrix<matrowfirst> C(F, 6, 4, 1);
    fill(A, p, g);
    fill(B, p, g);
    PolynomialMatrixThreePrimesFFTMulDomain D(F);
    D.mul(C, A, B);
    EXPECT(C.size() == 499, "c resized to deg+1");
    EXPECT(check_at(C, A, B, 1, p), "few primes: check_mul at x=1");
    EXPECT(check_at(C, A, B, g() % p, p), "few primes: check at a random point");
}

// PolynomialMatrix::copy across layouts and through a view (polynomial-matrix.h:206-243,
// 508-532, 763-790, 988-992): coefficient by coefficient against the source.
static void test_copy(std::mt19937_64 &g)
{
    const uint32_t p = 998244353u;
    Modular F(p);
    PolynomialMatrix<polfirst> A(F, 5, 7, 90);
    fill(A, p, g);
    PolynomialMatrix<matfirst> B(F, 5, 7, 90);
    B.copy(A);
    PolynomialMatrix<matrowfirst> C(F, 5, 7, 60);
    C.copy(B, 20, 69, 5);  // coefficients 20..69 of B into 5..54 of C
    bool ok = true;
    for (size_t e = 0; e < 35; ++e) {
        for (size_t t = 0; t < 90; ++t) ok &= B.get(e, t) == A.get(e, t);
        for (size_t t = 0; t < 60; ++t) ok &= C.get(e, t) == ((t >= 5 && t < 55) ? A.get(e, t + 15) : 0u);
    }
    EXPECT(ok, "copy polfirst -> matfirst -> matrowfirst window");
    PolynomialMatrix<polfirst> D(F, 5, 7, 40);
    auto V = D.at(10, 29);
    V.copy(C, 5, 24);  // into coefficients 10..29 of D
    ok = true;
    for (size_t e = 0; e < 35; ++e)
        for (size_t t = 0; t < 40; ++t) ok &= D.get(e, t) == ((t >= 10 && t < 30) ? A.get(e, t + 10) : 0u);
    EXPECT(ok, "copy into a view");
}

int main(int argc, char **argv)
{
    const uint64_t seed = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 20191;
    std::mt19937_64 g(seed);
    const uint32_t primes[4] = {998244353u, 2013265921u, 469762049u, 167772161u};
    for (uint32_t p : primes) {
        test_fft(p, g);
        test_mul<polfirst, polfirst, polfirst>(p, 16, 512, LBX_METHOD_PLAIN_U64, g);
        test_mul<polfirst, polfirst, polfirst>(p, 16, 512, LBX_METHOD_INT8_MFMA, g);
        test_mul<matfirst, matrowfirst, polfirst>(p, 16, 100, LBX_METHOD_AUTO, g);
        test_mul<matrowfirst, matfirst, matfirst>(p, 5, 33, LBX_METHOD_AUTO, g);
        test_views_and_errors(p, g);
        test_midproduct(p, g);
        test_order_basis(p, g);
    }
    test_mul<polfirst, polfirst, polfirst>(1000000007u, 8, 64, LBX_METHOD_AUTO, g);  // CRT route
    test_three_primes(g);
    test_copy(g);
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("PASS\n");
    return 0;
}
