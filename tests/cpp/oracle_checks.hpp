// Checks for tests/cpp/test_linbox_shim.cpp, in the role LinBox's own check
// functions play for its tests (check_mul and check_midproduct of
// matpoly-mult-fft.h:84-260, check_sigma of tests/test-order-basis.C:19-45):
// every result computed through the shim is compared word for word with the
// CPU oracle (oracle/lbx_oracle.h), which is test infrastructure only.
#ifndef LBX_TESTS_ORACLE_CHECKS_HPP
#define LBX_TESTS_ORACLE_CHECKS_HPP

#include <cstdio>
#include <utility>
#include <vector>

#include "../../oracle/lbx_oracle.h"

namespace lbx_test {

// Entries of a polynomial matrix (or view) as a polfirst array of `size` coefficients.
template <class Mat>
std::vector<uint32_t> polfirst(const Mat &M, size_t size = 0)
{
    if (!size) size = M.size();
    std::vector<uint32_t> v(M.rowdim() * M.coldim() * size, 0);
    for (size_t i = 0; i < M.rowdim() * M.coldim(); ++i)
        for (size_t k = 0; k < size && k < M.size(); ++k) v[i * size + k] = M.get(i, k);
    return v;
}

inline bool report(bool ok, const char *what, const char *shape)
{
    std::printf("Checking %s %s ... %s\n", what, shape, ok ? "done" : "error");
    return ok;
}

}  // namespace lbx_test

// c = a.b on `deg` coefficients (the product's first deg coefficients).
template <class MatC, class MatA, class MatB>
bool check_mul(const MatC &c, const MatA &a, const MatB &b, size_t deg)
{
    const uint32_t p = (uint32_t)c.field().cardinality();
    const std::vector<uint32_t> A = lbx_test::polfirst(a), B = lbx_test::polfirst(b);
    const size_t full = a.size() + b.size() - 1;
    std::vector<uint32_t> ref(c.rowdim() * c.coldim() * full);
    bool ok = or_polymat_mul_naive(p, a.rowdim(), a.coldim(), b.coldim(), A.data(), a.size(), B.data(), b.size(),
                                   ref.data()) == 0;
    for (size_t i = 0; ok && i < c.rowdim() * c.coldim(); ++i)
        for (size_t k = 0; k < deg; ++k)
            if (c.get(i, k) != (k < full ? ref[i * full + k] : 0u)) {
                ok = false;
                break;
            }
    char shape[96];
    std::snprintf(shape, sizeof shape, "%zux%zu[%zu].%zux%zu[%zu]", a.rowdim(), a.coldim(), a.size(), b.rowdim(),
                  b.coldim(), b.size());
    return lbx_test::report(ok, "polynomial matrix mul", shape);
}

// c = the FFT middle product of a and b (the route PolynomialMatrixMulDomain
// takes for these sizes), all c.size() coefficients; hdeg = n0 ? n0 : c.size(),
// deg = n1 ? n1 : 2 hdeg - 1, pts = 2^ceil(log2 deg).  For an FFT prime the
// oracle's restatement (wordsize-fast.inl:209-316); for any other modulus the
// cyclic correlation it computes, written out: with the first hdeg
// coefficients of the small operand reversed (a2 or b2, zero-padded to pts),
// c[t] = sum_v a2[(v-t) mod pts] b[v] (smallLeft) or sum_u a[u] b2[(u-t) mod pts].
template <class MatC, class MatA, class MatB>
bool check_midproduct(const MatC &c, const MatA &a, const MatB &b, bool smallLeft = true, size_t n0 = 0,
                      size_t n1 = 0)
{
    const uint32_t p = (uint32_t)c.field().cardinality();
    const std::vector<uint32_t> A = lbx_test::polfirst(a), B = lbx_test::polfirst(b);
    const size_t m = a.rowdim(), k = a.coldim(), n = b.coldim(), sc = c.size();
    std::vector<uint32_t> ref(m * n * sc);
    bool ok = or_polymat_midproduct_fft_prime(p, m, k, n, A.data(), a.size(), B.data(), b.size(), ref.data(), sc,
                                              smallLeft ? 1 : 0, n0, n1) == 0;
    if (!ok) {  // not an FFT prime for this length
        const size_t hdeg = n0 ? n0 : sc, deg = n1 ? n1 : 2 * hdeg - 1;
        size_t pts = 1;
        while (pts < deg) pts <<= 1;
        auto padded = [&](const std::vector<uint32_t> &X, size_t s, size_t e, bool rev) {
            std::vector<uint64_t> v(pts, 0);
            for (size_t t = 0; t < s && t < pts; ++t) v[t] = X[e * s + t];
            if (rev)
                for (size_t t = 0; t < hdeg / 2; ++t) std::swap(v[t], v[hdeg - 1 - t]);
            return v;
        };
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) {
                std::vector<unsigned __int128> acc(sc, 0);
                for (size_t q = 0; q < k; ++q) {
                    const std::vector<uint64_t> pa = padded(A, a.size(), i * k + q, smallLeft);
                    const std::vector<uint64_t> pb = padded(B, b.size(), q * n + j, !smallLeft);
                    for (size_t t = 0; t < sc; ++t)
                        for (size_t v = 0; v < pts; ++v)
                            acc[t] += smallLeft ? (unsigned __int128)pa[(v + pts - t % pts) % pts] * pb[v]
                                                : (unsigned __int128)pa[v] * pb[(v + pts - t % pts) % pts];
                }
                for (size_t t = 0; t < sc; ++t) ref[(i * n + j) * sc + t] = (uint32_t)(acc[t] % p);
            }
        ok = true;
    }
    ok = ok && lbx_test::polfirst(c) == ref;
    char shape[96];
    std::snprintf(shape, sizeof shape, "%zux%zu[%zu].%zux%zu[%zu] -> [%zu]", a.rowdim(), a.coldim(), a.size(),
                  b.rowdim(), b.coldim(), b.size(), c.size());
    return lbx_test::report(ok, "polynomial matrix midproduct", shape);
}

// sigma and shift equal the oracle's order basis of serie at `order` from the
// starting shift shift0 (PM_Basis when divide, else M_Basis), and
// sigma.serie = 0 mod x^order with sigma nonzero (the reference's check_sigma).
template <class MatS, class MatZ>
bool check_sigma(const MatS &sigma, const MatZ &serie, size_t order, const std::vector<size_t> &shift0,
                 const std::vector<size_t> &shift, bool divide)
{
    const uint32_t p = (uint32_t)sigma.field().cardinality();
    const size_t m = serie.rowdim(), n = serie.coldim();
    const std::vector<uint32_t> Z = lbx_test::polfirst(serie);
    std::vector<uint64_t> sh(shift0.begin(), shift0.end());
    std::vector<uint32_t> ref(m * m * (order + 1));
    size_t deg = 0;
    bool ok = (divide ? or_pm_basis : or_m_basis)(p, m, n, Z.data(), serie.size(), order, sh.data(), ref.data(),
                                                  &deg) == 0;
    ok = ok && sigma.size() == deg + 1;
    for (size_t i = 0; ok && i < m; ++i) ok = shift[i] == sh[i];
    for (size_t i = 0; ok && i < m * m; ++i)
        for (size_t k = 0; ok && k <= deg; ++k) ok = sigma.get(i, k) == ref[i * (order + 1) + k];
    // order property
    const std::vector<uint32_t> S = lbx_test::polfirst(sigma);
    std::vector<uint32_t> T(m * n * (sigma.size() + serie.size() - 1));
    ok = ok && or_polymat_mul_naive(p, m, m, n, S.data(), sigma.size(), Z.data(), serie.size(), T.data()) == 0;
    bool nonzero = false;
    for (size_t i = 0; ok && i < m * n; ++i)
        for (size_t k = 0; k < order; ++k) ok = ok && T[i * (sigma.size() + serie.size() - 1) + k] == 0;
    for (uint32_t v : S) nonzero = nonzero || v;
    ok = ok && nonzero;
    char shape[96];
    std::snprintf(shape, sizeof shape, "%zux%zu order %zu", m, n, order);
    return lbx_test::report(ok, divide ? "PM-Basis" : "M-Basis", shape);
}

#endif  // LBX_TESTS_ORACLE_CHECKS_HPP
