/*
 * lbx_oracle.c -- CPU restatement of LinBox's FFT polynomial-matrix product.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product path).  See
 * lbx_oracle.h for what each function restates and how it is pinned.
 */
#include "lbx_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Input generator: splitmix64 -> xoshiro256** -> Lemire rejection.          */
/* ------------------------------------------------------------------------ */
uint64_t or_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s[4]; } or_rng;

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void rng_seed(or_rng *r, uint64_t seed)
{
    uint64_t st = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = or_splitmix64(&st);
}

static uint64_t rng_next(or_rng *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

static uint32_t rng_uniform(or_rng *r, uint32_t bound)
{
    uint32_t x = (uint32_t)(rng_next(r) >> 32);
    uint64_t m = (uint64_t)x * bound;
    uint32_t l = (uint32_t)m;
    if (l < bound) {
        uint32_t t = (uint32_t)(-bound) % bound;
        while (l < t) {
            x = (uint32_t)(rng_next(r) >> 32);
            m = (uint64_t)x * bound;
            l = (uint32_t)m;
        }
    }
    return (uint32_t)(m >> 32);
}

void or_fill_uniform(uint64_t seed, uint32_t p, uint32_t *out, size_t n)
{
    or_rng r;
    rng_seed(&r, seed);
    for (size_t i = 0; i < n; ++i) out[i] = rng_uniform(&r, p);
}

/* ------------------------------------------------------------------------ */
/* Prime-field helpers.                                                      */
/* ------------------------------------------------------------------------ */
uint32_t or_powmod(uint32_t a, uint64_t e, uint32_t p)
{
    uint64_t r = 1 % p, b = a % p;
    while (e) {
        if (e & 1) r = r * b % p;
        b = b * b % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

uint32_t or_invmod(uint32_t a, uint32_t p) { return or_powmod(a, (uint64_t)p - 2, p); }

static uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t n) { return (uint64_t)((u128)a * b % n); }

static uint64_t powmod64(uint64_t a, uint64_t e, uint64_t n)
{
    uint64_t r = 1 % n;
    a %= n;
    while (e) {
        if (e & 1) r = mulmod64(r, a, n);
        a = mulmod64(a, a, n);
        e >>= 1;
    }
    return r;
}

/* Deterministic Miller-Rabin for n < 2^64 (bases: first 12 primes).  The
 * reference uses Givaro's probab_prime with 25 rounds; for the candidates
 * generatePrimes visits both decide primality identically. */
int or_is_prime_u64(uint64_t n)
{
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; ++i) {
        if (n == bases[i]) return 1;
        if (n % bases[i] == 0) return 0;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    for (int i = 0; i < 12; ++i) {
        uint64_t x = powmod64(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int composite = 1;
        for (int r = 1; r < s; ++r) {
            x = mulmod64(x, x, n);
            if (x == n - 1) { composite = 0; break; }
        }
        if (composite) return 0;
    }
    return 1;
}

/* fft-utils.h:65-78 is_primitive_root: w, w^2, ..., w^(2^(k-1)) all != 1. */
int or_is_primitive_root(uint32_t p, uint32_t w, unsigned k)
{
    uint64_t x = w;
    for (unsigned i = 0; i < k; ++i) {
        if (x == 1) return 0;
        x = x * x % p;
    }
    return 1;
}

/* fft-utils.h:83-104 compute_primitive_root. */
int or_compute_primitive_root(uint32_t p, unsigned k, uint32_t *w_out)
{
    unsigned val2p = 0;
    uint32_t m = p - 1;
    for (; (m & 1) == 0; m >>= 1, val2p++)
        ;
    if (k > val2p) return OR_ERR_K_TOO_LARGE;
    for (uint32_t t = 2;; ++t) {
        uint64_t w = or_powmod(t, m, p); /* w <- t^m */
        for (unsigned i = 0; i < val2p - k; ++i) w = w * w % p;
        if (or_is_primitive_root(p, (uint32_t)w, k)) {
            *w_out = (uint32_t)w;
            return OR_OK;
        }
    }
}

/* fft-integral.inl:1614-1650 init_powers (the first subarray and its
 * strided copies; the bit-reversed copies are not needed by DIF/DIT). */
void or_init_powers(uint32_t p, unsigned L, uint32_t w, uint32_t *pow_w, uint32_t *pow_wp)
{
    size_t n = (size_t)1 << L;
    pow_w[0] = 1 % p;
    for (size_t i = 1; i < n / 2; ++i) pow_w[i] = (uint32_t)((uint64_t)pow_w[i - 1] * w % p);
    size_t idx = n / 2;
    for (size_t kk = 2; kk <= n / 2; kk <<= 1)
        for (size_t i = 0; i < n / 2; i += kk, idx++) pow_w[idx] = pow_w[i];
    if (pow_wp)
        for (size_t i = 0; i + 1 < n; ++i) pow_wp[i] = (uint32_t)(((uint64_t)pow_w[i] << 32) / p);
}

/* NoSimd DIF_core (fft-integral.inl:204-217) with Butterfly_DIF (:378-386):
 * A <- A+B, B <- (A-B)*alpha, families f, width w, 2fw == n. */
void or_fft_dif(uint32_t p, unsigned L, const uint32_t *pow_w, uint32_t *c)
{
    size_t n = (size_t)1 << L;
    const uint32_t *pw = pow_w;
    for (size_t w = n >> 1, f = 1; w > 0; pw += w, f <<= 1, w >>= 1) {
        for (size_t i = 0; i < f; ++i) {
            uint32_t *A = c + i * 2 * w, *B = A + w;
            for (size_t j = 0; j < w; ++j) {
                uint64_t a = A[j], b = B[j];
                A[j] = (uint32_t)((a + b) % p);
                B[j] = (uint32_t)((a + p - b) * pw[j] % p);
            }
        }
    }
}

/* NoSimd DIT_core (fft-integral.inl:284-299) with Butterfly_DIT (:475-483):
 * t = B*alpha, A <- A+t, B <- A-t; starts at the last subarray (n-2). */
void or_fft_dit(uint32_t p, unsigned L, const uint32_t *pow_w, uint32_t *c)
{
    size_t n = (size_t)1 << L;
    size_t off = n - 2;
    for (size_t w = 1, f = n >> 1; w < n;) {
        const uint32_t *pw = pow_w + off;
        for (size_t i = 0; i < f; ++i) {
            uint32_t *A = c + i * 2 * w, *B = A + w;
            for (size_t j = 0; j < w; ++j) {
                uint64_t a = A[j], t = (uint64_t)B[j] * pw[j] % p;
                A[j] = (uint32_t)((a + t) % p);
                B[j] = (uint32_t)((a + p - t) % p);
            }
        }
        w <<= 1;
        f >>= 1;
        if (w < n) off -= w;
    }
}

/* fft.h:119-130 invroot: -1 for n=2, else -(w^(n/2-1)). */
uint32_t or_fft_invroot(uint32_t p, unsigned L, uint32_t w)
{
    if (L == 1) return p - 1;
    uint32_t x = or_powmod(w, ((uint64_t)1 << (L - 1)) - 1, p);
    return x ? p - x : 0;
}

/* fft.h:133-144 check_args. */
static int fft_check_args(uint32_t p, unsigned L, uint32_t root, uint32_t *w)
{
    if (!L) return OR_ERR_K_NOT_POSITIVE;
    if (root == 0) return or_compute_primitive_root(p, L, w);
    if (!or_is_primitive_root(p, root, L)) return OR_ERR_NOT_PRIMITIVE_ROOT;
    *w = root;
    return OR_OK;
}

static int fft_run(uint32_t p, unsigned L, uint32_t root, uint32_t *coeffs, uint32_t *root_used, int inverse)
{
    uint32_t w;
    int st = fft_check_args(p, L, root, &w);
    if (st) return st;
    size_t n = (size_t)1 << L;
    uint32_t *pw = (uint32_t *)malloc(sizeof(uint32_t) * (n - 1));
    if (!pw) return OR_ERR_INVALID_ARG;
    or_init_powers(p, L, w, pw, NULL);
    if (inverse)
        or_fft_dit(p, L, pw, coeffs);
    else
        or_fft_dif(p, L, pw, coeffs);
    free(pw);
    if (root_used) *root_used = w;
    return OR_OK;
}

int or_fft_direct(uint32_t p, unsigned L, uint32_t root, uint32_t *coeffs, uint32_t *root_used)
{
    return fft_run(p, L, root, coeffs, root_used, 0);
}

int or_fft_inverse(uint32_t p, unsigned L, uint32_t root, uint32_t *coeffs, uint32_t *root_used)
{
    return fft_run(p, L, root, coeffs, root_used, 1);
}

/* ------------------------------------------------------------------------ */
/* Single-prime FFT product (matpoly-mult-fft-wordsize-fast.inl:62-190).     */
/* ------------------------------------------------------------------------ */
void or_stage_forward(uint32_t p, unsigned L, const uint32_t *pow_w, uint32_t *polys, size_t count)
{
    size_t n = (size_t)1 << L;
    for (size_t e = 0; e < count; ++e) or_fft_dif(p, L, pow_w, polys + e * n);
}

void or_stage_inverse(uint32_t p, unsigned L, const uint32_t *pow_winv, uint32_t *polys, size_t count)
{
    size_t n = (size_t)1 << L;
    for (size_t e = 0; e < count; ++e) or_fft_dit(p, L, pow_winv, polys + e * n);
}

/* The pointwise loop of mul_fft (:160-164): at every point, C = A.B mod p
 * (BlasMatrixDomain::mul -> fgemm; here an exact u128 dot product). */
void or_stage_pointwise(uint32_t p, size_t m, size_t k, size_t n, size_t pts,
                        const uint32_t *ahat, const uint32_t *bhat, uint32_t *chat,
                        size_t pt0, size_t npts)
{
    uint64_t *brow = (uint64_t *)malloc(sizeof(uint64_t) * k * n);
    for (size_t t = pt0; t < pt0 + npts; ++t) {
        for (size_t q = 0; q < k * n; ++q) brow[q] = bhat[q * pts + t];
        for (size_t i = 0; i < m; ++i) {
            for (size_t j = 0; j < n; ++j) {
                u128 acc = 0;
                for (size_t q = 0; q < k; ++q) acc += (u128)((uint64_t)ahat[(i * k + q) * pts + t] * brow[q * n + j]);
                chat[(i * n + j) * pts + t] = (uint32_t)(acc % p);
            }
        }
    }
    free(brow);
}

/* The same three stages over `threads` OpenMP threads (polynomials, points). */
void or_stage_forward_mt(uint32_t p, unsigned L, const uint32_t *pow_w, uint32_t *polys, size_t count, int threads)
{
    size_t n = (size_t)1 << L;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (size_t e = 0; e < count; ++e) or_fft_dif(p, L, pow_w, polys + e * n);
}

void or_stage_inverse_mt(uint32_t p, unsigned L, const uint32_t *pow_winv, uint32_t *polys, size_t count, int threads)
{
    size_t n = (size_t)1 << L;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (size_t e = 0; e < count; ++e) or_fft_dit(p, L, pow_winv, polys + e * n);
}

void or_stage_pointwise_mt(uint32_t p, size_t m, size_t k, size_t n, size_t pts, const uint32_t *ahat,
                           const uint32_t *bhat, uint32_t *chat, size_t pt0, size_t npts, int threads)
{
#pragma omp parallel for num_threads(threads) schedule(static)
    for (size_t t = pt0; t < pt0 + npts; ++t) or_stage_pointwise(p, m, k, n, pts, ahat, bhat, chat, t, 1);
}

int or_polymat_mul_fft_prime(uint32_t p, size_t m, size_t k, size_t n,
                             const uint32_t *a, size_t sa, const uint32_t *b, size_t sb,
                             uint32_t *c, size_t max_rowdeg)
{
    if (!sa || !sb || !m || !k || !n) return OR_ERR_INVALID_ARG;
    size_t deg = max_rowdeg ? max_rowdeg : sa + sb - 2;
    unsigned lpts = 0;
    size_t pts = 1;
    while (pts <= deg) { pts <<= 1; ++lpts; }
    if (sa > pts || sb > pts) return OR_ERR_SIZE_MISMATCH;
    /* mul_fft :107-112 then FFT ctor checks (fft.h:133-144) */
    if ((p - 1) % pts != 0) return OR_ERR_BAD_FFT_PRIME;
    if (lpts == 0) return OR_ERR_K_NOT_POSITIVE;
    uint32_t w;
    int st = or_compute_primitive_root(p, lpts, &w);
    if (st) return st;
    uint32_t winv = or_fft_invroot(p, lpts, w);

    uint32_t *pw = (uint32_t *)malloc(sizeof(uint32_t) * pts);
    uint32_t *pwi = (uint32_t *)malloc(sizeof(uint32_t) * pts);
    uint32_t *a2 = (uint32_t *)calloc(m * k * pts, sizeof(uint32_t));
    uint32_t *b2 = (uint32_t *)calloc(k * n * pts, sizeof(uint32_t));
    uint32_t *c2 = (uint32_t *)calloc(m * n * pts, sizeof(uint32_t));
    if (!pw || !pwi || !a2 || !b2 || !c2) {
        free(pw); free(pwi); free(a2); free(b2); free(c2);
        return OR_ERR_INVALID_ARG;
    }
    or_init_powers(p, lpts, w, pw, NULL);
    or_init_powers(p, lpts, winv, pwi, NULL);
    for (size_t e = 0; e < m * k; ++e) memcpy(a2 + e * pts, a + e * sa, sizeof(uint32_t) * sa);
    for (size_t e = 0; e < k * n; ++e) memcpy(b2 + e * pts, b + e * sb, sizeof(uint32_t) * sb);

    or_stage_forward(p, lpts, pw, a2, m * k);
    or_stage_forward(p, lpts, pw, b2, k * n);
    or_stage_pointwise(p, m, k, n, pts, a2, b2, c2, 0, pts);
    or_stage_inverse(p, lpts, pwi, c2, m * n);
    /* :183-189 fscalin by pts^-1 */
    uint64_t inv_pts = or_invmod((uint32_t)(pts % p), p);
    for (size_t e = 0; e < m * n; ++e)
        for (size_t t = 0; t <= deg; ++t) c[e * (deg + 1) + t] = (uint32_t)(c2[e * pts + t] * inv_pts % p);
    free(pw); free(pwi); free(a2); free(b2); free(c2);
    return OR_OK;
}

/* matpoly-mult-naive.h:56-82. */
int or_polymat_mul_naive(uint32_t p, size_t m, size_t k, size_t n,
                         const uint32_t *a, size_t sa, const uint32_t *b, size_t sb, uint32_t *c)
{
    if (!sa || !sb) return OR_ERR_INVALID_ARG;
    size_t sc = sa + sb - 1;
    for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < n; ++j)
            for (size_t t = 0; t < sc; ++t) {
                size_t idx_min = (t + 1 < sb) ? 0 : t + 1 - sb;
                size_t idx_max = t < sa - 1 ? t : sa - 1;
                u128 acc = 0;
                for (size_t q = 0; q < k; ++q) {
                    const uint32_t *pa = a + (i * k + q) * sa;
                    const uint32_t *pb = b + (q * n + j) * sb;
                    for (size_t u = idx_min; u <= idx_max; ++u) acc += (uint64_t)pa[u] * pb[t - u];
                }
                c[(i * n + j) * sc + t] = (uint32_t)(acc % p);
            }
    return OR_OK;
}

/* matpoly-mult-fft-wordsize-fast.inl:209-316. */
int or_polymat_midproduct_fft_prime(uint32_t p, size_t m, size_t k, size_t n,
                                    const uint32_t *a, size_t sa, const uint32_t *b, size_t sb,
                                    uint32_t *c, size_t sc, int small_left, size_t n0, size_t n1)
{
    if (!sa || !sb || !sc || !m || !k || !n) return OR_ERR_INVALID_ARG;
    size_t hdeg = n0 ? n0 : sc;
    size_t deg = n1 ? n1 : 2 * hdeg - 1;
    if (deg < hdeg || sc < deg - hdeg) return OR_ERR_SIZE_MISMATCH;                   /* :214 */
    if ((small_left ? sb : sa) >= hdeg + deg) return OR_ERR_SIZE_MISMATCH;           /* :215-218 */
    unsigned lpts = 0;
    size_t pts = 1;
    while (pts < deg) { pts <<= 1; ++lpts; }                                          /* :221 */
    if (sa > pts || sb > pts || sc > pts || hdeg > pts) return OR_ERR_SIZE_MISMATCH;  /* copies into pts words */
    if ((p - 1) % pts != 0) return OR_ERR_BAD_FFT_PRIME;                              /* :265-272 */
    if (lpts == 0) return OR_ERR_K_NOT_POSITIVE;                                      /* fft.h:133-144 */
    uint32_t w;
    int st = or_compute_primitive_root(p, lpts, &w);
    if (st) return st;
    uint32_t winv = or_fft_invroot(p, lpts, w);                                       /* :274 FFTinv */

    uint32_t *pw = (uint32_t *)malloc(sizeof(uint32_t) * pts);
    uint32_t *pwi = (uint32_t *)malloc(sizeof(uint32_t) * pts);
    uint32_t *a2 = (uint32_t *)calloc(m * k * pts, sizeof(uint32_t));
    uint32_t *b2 = (uint32_t *)calloc(k * n * pts, sizeof(uint32_t));
    uint32_t *c2 = (uint32_t *)calloc(m * n * pts, sizeof(uint32_t));
    if (!pw || !pwi || !a2 || !b2 || !c2) {
        free(pw); free(pwi); free(a2); free(b2); free(c2);
        return OR_ERR_INVALID_ARG;
    }
    or_init_powers(p, lpts, w, pw, NULL);
    or_init_powers(p, lpts, winv, pwi, NULL);
    for (size_t e = 0; e < m * k; ++e) memcpy(a2 + e * pts, a + e * sa, sizeof(uint32_t) * sa);
    for (size_t e = 0; e < k * n; ++e) memcpy(b2 + e * pts, b + e * sb, sizeof(uint32_t) * sb);
    /* :228-235 reverse the first hdeg coefficients of the small operand */
    uint32_t *sm = small_left ? a2 : b2;
    size_t se = small_left ? m * k : k * n;
    for (size_t e = 0; e < se; ++e)
        for (size_t i = 0; i < hdeg / 2; ++i) {
            uint32_t t = sm[e * pts + i];
            sm[e * pts + i] = sm[e * pts + hdeg - 1 - i];
            sm[e * pts + hdeg - 1 - i] = t;
        }
    /* :279-290 forward transforms */
    or_stage_forward(p, lpts, small_left ? pw : pwi, a2, m * k);
    or_stage_forward(p, lpts, small_left ? pwi : pw, b2, k * n);
    /* :300-305 pointwise a.b */
    or_stage_pointwise(p, m, k, n, pts, a2, b2, c2, 0, pts);
    /* :311-313 FFTer.FFT_inverse: DIT with the root w */
    or_stage_inverse(p, lpts, pw, c2, m * n);
    /* :316-324 fscalin by pts^-1, then c.copy(c2, 0, c.size()-1) (:242) */
    uint64_t inv_pts = or_invmod((uint32_t)(pts % p), p);
    for (size_t e = 0; e < m * n; ++e)
        for (size_t t = 0; t < sc; ++t) c[e * sc + t] = (uint32_t)(c2[e * pts + t] * inv_pts % p);
    free(pw); free(pwi); free(a2); free(b2); free(c2);
    return OR_OK;
}

/* matpoly-mult-naive.h:85-125. */
int or_polymat_midproduct_naive(uint32_t p, size_t m, size_t k, size_t n,
                                const uint32_t *a, size_t sa, const uint32_t *b, size_t sb,
                                uint32_t *c, size_t sc, int small_left, size_t n0, size_t n1)
{
    if (!sa || !sb || !sc) return OR_ERR_INVALID_ARG;
    size_t hdeg = n0 ? n0 : sc;
    size_t deg = n1 ? n1 : 2 * hdeg - 1;
    if (hdeg == 0) return OR_ERR_INVALID_ARG;
    size_t kend = sa + sb - 1 < deg ? sa + sb - 1 : deg;
    for (size_t kk = hdeg - 1; kk < kend; ++kk) {
        size_t t = kk - hdeg + 1;
        if (t >= sc) return OR_ERR_SIZE_MISMATCH;
        /* smallLeft walks a from idx_a = kk - min(kk, sb-1), the other branch walks b
         * from kk - min(kk, sa-1); both sum a[u] b[kk-u] over every valid u */
        size_t u_lo = kk + 1 > sb ? kk + 1 - sb : 0;
        size_t u_hi = kk < sa - 1 ? kk : sa - 1;
        (void)small_left;
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) {
                u128 acc = 0;
                for (size_t q = 0; q < k; ++q) {
                    const uint32_t *pa = a + (i * k + q) * sa;
                    const uint32_t *pb = b + (q * n + j) * sb;
                    for (size_t u = u_lo; u <= u_hi; ++u) acc += (uint64_t)pa[u] * pb[kk - u];
                }
                c[(i * n + j) * sc + t] = (uint32_t)(acc % p);
            }
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Small fixed-width multiprecision (little-endian u32 limbs).               */
/* ------------------------------------------------------------------------ */
#define MPL 40 /* enough for 1280-bit bounds */

static int mp_cmp(const uint32_t *x, const uint32_t *y, int n)
{
    for (int i = n - 1; i >= 0; --i)
        if (x[i] != y[i]) return x[i] < y[i] ? -1 : 1;
    return 0;
}

static void mp_mul_small(uint32_t *x, int n, uint32_t v)
{
    uint64_t carry = 0;
    for (int i = 0; i < n; ++i) {
        uint64_t t = (uint64_t)x[i] * v + carry;
        x[i] = (uint32_t)t;
        carry = t >> 32;
    }
}

static void mp_mul(const uint32_t *x, int nx, const uint32_t *y, int ny, uint32_t *z, int nz)
{
    memset(z, 0, sizeof(uint32_t) * nz);
    for (int i = 0; i < nx; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < ny && i + j < nz; ++j) {
            uint64_t t = (uint64_t)x[i] * y[j] + z[i + j] + carry;
            z[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        for (int q = i + ny; q < nz && carry; ++q) {
            uint64_t t = (uint64_t)z[q] + carry;
            z[q] = (uint32_t)t;
            carry = t >> 32;
        }
    }
}

static int mp_bitlen(const uint32_t *x, int n)
{
    for (int i = n - 1; i >= 0; --i)
        if (x[i]) return i * 32 + 32 - __builtin_clz(x[i]);
    return 0;
}

/* x (n limbs) <- x mod P (np limbs) by shift-and-subtract. */
static void mp_mod(uint32_t *x, int n, const uint32_t *P, int np)
{
    int bx = mp_bitlen(x, n), bp = mp_bitlen(P, np);
    uint32_t sh[MPL + 1];
    for (int s = bx - bp; s >= 0; --s) {
        /* sh = P << s, n limbs */
        memset(sh, 0, sizeof(uint32_t) * n);
        int ls = s / 32, bs = s % 32;
        for (int i = 0; i < np; ++i) {
            if (i + ls < n) sh[i + ls] |= P[i] << bs;
            if (bs && i + ls + 1 < n) sh[i + ls + 1] |= P[i] >> (32 - bs);
        }
        if (mp_cmp(x, sh, n) >= 0) {
            uint64_t borrow = 0;
            for (int i = 0; i < n; ++i) {
                uint64_t t = (uint64_t)x[i] - sh[i] - borrow;
                x[i] = (uint32_t)t;
                borrow = (t >> 32) & 1;
            }
        }
    }
}

/* random-fftprime.h:107-133 generatePrimes + :161-166 compute_max_m. */
int or_generate_fft_primes(uint64_t pbound, const uint32_t *prodbound, int nlimbs,
                           unsigned kmin, uint32_t *primes, int max_primes)
{
    uint32_t prod[MPL];
    int np = 0;
    if (nlimbs > MPL - 2 || pbound < 4 || pbound > ((uint64_t)1 << 32)) return -1;
    memset(prod, 0, sizeof(prod));
    prod[0] = 1;
    if (!kmin) kmin++;
    int bits = 64 - __builtin_clzll(pbound - 2);
    for (int k = bits - 1; k > (int)kmin; --k) {
        uint64_t B = pbound - 1;
        int64_t m = (B & ((1ull << k) - 1)) ? (int64_t)(B >> k) : (int64_t)(B >> k) - 1;
        if (!(m & 1)) m--;
        for (; m > 0; m -= 2) {
            uint64_t p = ((uint64_t)m << k) + 1;
            if (or_is_prime_u64(p)) {
                if (np >= max_primes) return -1;
                primes[np++] = (uint32_t)p;
                mp_mul_small(prod, MPL, (uint32_t)p);
                if (mp_cmp(prod, prodbound, nlimbs) > 0 || mp_bitlen(prod, MPL) > nlimbs * 32) return np;
            }
        }
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Multiprecision product via CRT (matpoly-mult-fft-multiprecision.inl).     */
/* ------------------------------------------------------------------------ */
int or_polymat_mul_crt(const uint32_t *P, int nlimbs, size_t m, size_t k, size_t n,
                       const uint32_t *a, size_t sa, const uint32_t *b, size_t sb, uint32_t *c,
                       uint32_t *primes_used, int *nprimes_out)
{
    if (nlimbs < 1 || nlimbs > 16 || !sa || !sb) return OR_ERR_INVALID_ARG;
    size_t s = sa + sb - 1;
    if (s == 1) { /* constant matrices: PolynomialMatrixMulDomain::mul routes d<=2 to naive */
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) {
                uint32_t acc[MPL];
                memset(acc, 0, sizeof(acc));
                for (size_t q = 0; q < k; ++q) {
                    uint32_t prod[MPL];
                    mp_mul(a + (i * k + q) * nlimbs, nlimbs, b + (q * n + j) * nlimbs, nlimbs, prod, 2 * nlimbs + 2);
                    uint64_t carry = 0;
                    for (int z = 0; z < 2 * nlimbs + 2; ++z) {
                        uint64_t t = (uint64_t)acc[z] + prod[z] + carry;
                        acc[z] = (uint32_t)t;
                        carry = t >> 32;
                    }
                    mp_mod(acc, 2 * nlimbs + 2, P, nlimbs);
                }
                memcpy(c + (i * n + j) * nlimbs, acc, sizeof(uint32_t) * nlimbs);
            }
        if (nprimes_out) *nprimes_out = 0;
        return OR_OK;
    }
    unsigned lpts = 0;
    size_t pts = 1;
    while (pts < s) { pts <<= 1; ++lpts; } /* :151-154 */
    /* :616 bound = 2 P^2 k min(sa,sb) */
    uint32_t bound[MPL];
    int nb = 2 * nlimbs + 3;
    mp_mul(P, nlimbs, P, nlimbs, bound, nb);
    mp_mul_small(bound, nb, 2);
    mp_mul_small(bound, nb, (uint32_t)k);
    mp_mul_small(bound, nb, (uint32_t)(sa < sb ? sa : sb));
    uint32_t primes[64];
    int r = or_generate_fft_primes((uint64_t)1 << 31, bound, nb, lpts, primes, 64);
    if (r < 0) return OR_ERR_NOT_ENOUGH_PRIMES;
    if (primes_used) memcpy(primes_used, primes, sizeof(uint32_t) * r);
    if (nprimes_out) *nprimes_out = r;

    size_t na = m * k * sa, nbb = k * n * sb, nc = m * n * s;
    uint32_t *ar = (uint32_t *)malloc(sizeof(uint32_t) * na);
    uint32_t *br = (uint32_t *)malloc(sizeof(uint32_t) * nbb);
    uint32_t *cr = (uint32_t *)malloc(sizeof(uint32_t) * nc * r);
    if (!ar || !br || !cr) { free(ar); free(br); free(cr); return OR_ERR_INVALID_ARG; }
    for (int l = 0; l < r; ++l) {
        uint64_t pl = primes[l];
        /* RNS reduce-in (:192-193): Horner over limbs */
        for (size_t i = 0; i < na; ++i) {
            uint64_t v = 0;
            for (int q = nlimbs - 1; q >= 0; --q) v = ((v << 32) | a[i * nlimbs + q]) % pl;
            ar[i] = (uint32_t)v;
        }
        for (size_t i = 0; i < nbb; ++i) {
            uint64_t v = 0;
            for (int q = nlimbs - 1; q >= 0; --q) v = ((v << 32) | b[i * nlimbs + q]) % pl;
            br[i] = (uint32_t)v;
        }
        int st = or_polymat_mul_fft_prime((uint32_t)pl, m, k, n, ar, sa, br, sb, cr + (size_t)l * nc, 0);
        if (st) { free(ar); free(br); free(cr); return st; }
    }
    /* CRT out (RNS.convert :264) then %P (:626-628): Garner mixed radix,
     * then X mod P = sum v_i (prod_{j<i} p_j mod P). */
    uint32_t Mi[64][MPL];
    int w = nlimbs + 2;
    {
        uint32_t M[MPL];
        memset(M, 0, sizeof(M));
        M[0] = 1;
        for (int i = 0; i < r; ++i) {
            uint32_t t[MPL];
            memcpy(t, M, sizeof(t));
            mp_mod(t, MPL, P, nlimbs);
            memcpy(Mi[i], t, sizeof(t));
            mp_mul_small(M, MPL, primes[i]);
            mp_mod(M, MPL, P, nlimbs); /* only M mod P is needed */
        }
    }
    uint32_t inv[64][64];
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < i; ++j) inv[i][j] = or_invmod(primes[j] % primes[i], primes[i]);
    for (size_t e = 0; e < nc; ++e) {
        uint32_t v[64];
        for (int i = 0; i < r; ++i) {
            uint64_t pi = primes[i], t = cr[(size_t)i * nc + e];
            for (int j = 0; j < i; ++j) t = (t + pi - v[j] % pi) % pi * inv[i][j] % pi;
            v[i] = (uint32_t)t;
        }
        uint32_t acc[MPL];
        memset(acc, 0, sizeof(acc));
        for (int i = 0; i < r; ++i) {
            uint64_t carry = 0;
            for (int q = 0; q < w + 2; ++q) {
                uint64_t t = (uint64_t)Mi[i][q] * v[i] + acc[q] + carry;
                acc[q] = (uint32_t)t;
                carry = t >> 32;
            }
        }
        mp_mod(acc, w + 2, P, nlimbs);
        memcpy(c + e * nlimbs, acc, sizeof(uint32_t) * nlimbs);
    }
    free(ar); free(br); free(cr);
    return OR_OK;
}

/* PolynomialMatrixThreePrimesFFTMulDomain (three-primes.inl:57-191). */
int or_polymat_mul_three_primes(uint32_t p, size_t m, size_t k, size_t n,
                                const uint32_t *a, size_t sa, const uint32_t *b, size_t sb, uint32_t *c,
                                uint32_t *primes_used, int *nprimes_out)
{
    if (p >= (1u << 29)) return OR_ERR_PRIME_TOO_LARGE; /* :59-63 */
    if (p < 2 || !sa || !sb) return OR_ERR_INVALID_ARG;
    if (nprimes_out) *nprimes_out = 0;
    size_t s = sa + sb - 1;
    if (s == 1) return or_polymat_mul_naive(p, m, k, n, a, sa, b, sb, c);
    unsigned lpts = 0;
    size_t pts = 1;
    while (pts < s) { pts <<= 1; ++lpts; } /* :71-72, pts > deg */
    if ((p - 1) % pts == 0) return or_polymat_mul_fft_prime(p, m, k, n, a, sa, b, sb, c, 0); /* :107-111 */
    uint32_t bound[4] = {p - 1, 0, 0, 0}; /* (p-1)^2 k min(sa,sb) < 2^122 */
    mp_mul_small(bound, 4, p - 1);
    mp_mul_small(bound, 4, (uint32_t)k);
    mp_mul_small(bound, 4, (uint32_t)(sa < sb ? sa : sb));
    uint32_t primes[64];
    int r = or_generate_fft_primes((uint64_t)1 << 31, bound, 4, lpts, primes, 64);
    if (r < 0) return OR_ERR_NOT_ENOUGH_PRIMES; /* :118-123 */
    if (primes_used) memcpy(primes_used, primes, sizeof(uint32_t) * r);
    if (nprimes_out) *nprimes_out = r;
    size_t na = m * k * sa, nbb = k * n * sb, nc = m * n * s;
    uint32_t *ar = malloc(sizeof(uint32_t) * na), *br = malloc(sizeof(uint32_t) * nbb);
    uint32_t *cr = malloc(sizeof(uint32_t) * nc * r);
    if (!ar || !br || !cr) { free(ar); free(br); free(cr); return OR_ERR_INVALID_ARG; }
    for (int l = 0; l < r; ++l) {
        for (size_t i = 0; i < na; ++i) ar[i] = a[i] % primes[l]; /* copy when p_l > p, finit otherwise */
        for (size_t i = 0; i < nbb; ++i) br[i] = b[i] % primes[l];
        int st = or_polymat_mul_fft_prime(primes[l], m, k, n, ar, sa, br, sb, cr + (size_t)l * nc, 0);
        if (st) { free(ar); free(br); free(cr); return st; }
    }
    /* mixed-radix reconstruction (:164-181), on whole arrays as the reference does */
    for (size_t e = 0; e < nc; ++e) c[e] = cr[e] % p; /* freduce */
    uint64_t beta = 1 % p;
    for (int i = 1; i < r; ++i) {
        uint64_t pi = primes[i];
        uint32_t *ci = cr + (size_t)i * nc;
        for (int j = 0; j < i; ++j) {
            uint64_t alpha = or_invmod(primes[j] % primes[i], primes[i]);
            const uint32_t *cj = cr + (size_t)j * nc;
            for (size_t e = 0; e < nc; ++e) ci[e] = (uint32_t)((ci[e] + pi - cj[e] % pi) % pi * alpha % pi);
        }
        beta = beta * (primes[i - 1] % p) % p;
        for (size_t e = 0; e < nc; ++e) c[e] = (uint32_t)((c[e] + beta * (ci[e] % p)) % p);
    }
    free(ar); free(br); free(cr);
    return OR_OK;
}

/* ------------------------------------------------------------------------
 * Order bases: OrderBasis::M_Basis (order-basis.h:346-600) and
 * OrderBasis::PM_Basis (order-basis.h:219-326).
 * ---------------------------------------------------------------------- */

/* Row structure of a constant m x n matrix d (row-major) as a rank-profile
 * revealing PLUQ exposes it (the FFPACK::PLUQ call of order-basis.h:470-473):
 * the pivot rows are the row rank profile (row i is a pivot when it is not in
 * the span of rows 0..i-1), and the row permutation puts the pivot rows first
 * and the other rows after them, each group in increasing order.  rows[] gets
 * that order, *rank the rank, and Y[(s-rank)*m + t] the coefficients of
 * non-pivot row rows[s] in the pivot rows rows[0..rank):
 *   d[rows[s]] = sum_t Y[(s-rank)*m+t] d[rows[t]],
 * which is the L2 L1^-1 of the reference's PLUQ factors (:485-488 solve
 * X L1 = -L2 for X = -L2 L1^-1).  Computed here from a reduced echelon form
 * of the pivot rows and the inverse of their pivot-column submatrix. */
static int pluq_rows(uint32_t p, size_t m, size_t n, const uint32_t *d, size_t *rows, size_t *rank_out, uint32_t *Y)
{
    size_t rmax = m < n ? m : n, rank = 0, npiv = 0, nnon = 0;
    uint32_t *E = calloc(rmax * n + 1, 4), *v = malloc(n * 4 + 4);
    uint32_t *M = malloc(rmax * rmax * 4 + 4), *Mi = malloc(rmax * rmax * 4 + 4);
    size_t *pc = malloc(rmax * sizeof(size_t) + 8), *piv = malloc(m * sizeof(size_t)), *non = malloc(m * sizeof(size_t));
    if (!E || !v || !M || !Mi || !pc || !piv || !non) return OR_ERR_INVALID_ARG;
    for (size_t i = 0; i < m; ++i) {
        memcpy(v, d + i * n, n * 4);
        for (size_t t = 0; t < rank; ++t) {
            uint64_t f = v[pc[t]];
            if (!f) continue;
            for (size_t c = 0; c < n; ++c) v[c] = (uint32_t)((v[c] + (uint64_t)(p - E[t * n + c]) * f) % p);
        }
        size_t c0 = 0;
        while (c0 < n && !v[c0]) ++c0;
        if (c0 == n) {
            non[nnon++] = i;
            continue;
        }
        uint64_t inv = or_invmod(v[c0], p);
        for (size_t c = 0; c < n; ++c) E[rank * n + c] = (uint32_t)(v[c] * inv % p);
        for (size_t t = 0; t < rank; ++t) { /* keep the form reduced */
            uint64_t f = E[t * n + c0];
            if (!f) continue;
            for (size_t c = 0; c < n; ++c)
                E[t * n + c] = (uint32_t)((E[t * n + c] + (uint64_t)(p - E[rank * n + c]) * f) % p);
        }
        pc[rank++] = c0;
        piv[npiv++] = i;
    }
    /* M = d[pivot rows][pivot columns], inverted by Gauss-Jordan */
    for (size_t a = 0; a < rank; ++a)
        for (size_t b = 0; b < rank; ++b) {
            M[a * rank + b] = d[piv[a] * n + pc[b]];
            Mi[a * rank + b] = a == b;
        }
    for (size_t col = 0; col < rank; ++col) {
        size_t r = col;
        while (r < rank && !M[r * rank + col]) ++r;
        if (r == rank) return OR_ERR_INVALID_ARG; /* cannot happen: pivot rows are independent */
        for (size_t b = 0; b < rank; ++b) {
            uint32_t t = M[col * rank + b]; M[col * rank + b] = M[r * rank + b]; M[r * rank + b] = t;
            t = Mi[col * rank + b]; Mi[col * rank + b] = Mi[r * rank + b]; Mi[r * rank + b] = t;
        }
        uint64_t inv = or_invmod(M[col * rank + col], p);
        for (size_t b = 0; b < rank; ++b) {
            M[col * rank + b] = (uint32_t)(M[col * rank + b] * inv % p);
            Mi[col * rank + b] = (uint32_t)(Mi[col * rank + b] * inv % p);
        }
        for (size_t a = 0; a < rank; ++a) {
            uint64_t f = M[a * rank + col];
            if (a == col || !f) continue;
            for (size_t b = 0; b < rank; ++b) {
                M[a * rank + b] = (uint32_t)((M[a * rank + b] + (uint64_t)(p - M[col * rank + b]) * f) % p);
                Mi[a * rank + b] = (uint32_t)((Mi[a * rank + b] + (uint64_t)(p - Mi[col * rank + b]) * f) % p);
            }
        }
    }
    for (size_t s = 0; s < nnon; ++s)
        for (size_t t = 0; t < rank; ++t) {
            uint64_t acc = 0;
            for (size_t b = 0; b < rank; ++b) acc = (acc + (uint64_t)d[non[s] * n + pc[b]] * Mi[b * rank + t]) % p;
            Y[s * m + t] = (uint32_t)acc;
        }
    for (size_t t = 0; t < npiv; ++t) rows[t] = piv[t];
    for (size_t s = 0; s < nnon; ++s) rows[npiv + s] = non[s];
    *rank_out = rank;
    free(E); free(v); free(M); free(Mi); free(pc); free(piv); free(non);
    return OR_OK;
}

int or_m_basis(uint32_t p, size_t m, size_t n, const uint32_t *serie, size_t ssz, size_t order, uint64_t *shift,
               uint32_t *sigma, size_t *deg_out)
{
    if (p < 2 || !m || !n || !order || ssz < order || !serie || !shift || !sigma) return OR_ERR_INVALID_ARG;
    const size_t mm = m * m;
    uint32_t *sig = calloc((order + 1) * mm, 4), *tmp = malloc(mm * 4), *delta = malloc(m * n * 4);
    uint32_t *Y = malloc(m * m * 4 + 4);
    uint64_t *degree = calloc(m, 8), *t64 = malloc(m * 8);
    size_t *perm = malloc(m * sizeof(size_t)), *rows = malloc(m * sizeof(size_t));
    if (!sig || !tmp || !delta || !Y || !degree || !t64 || !perm || !rows) return OR_ERR_INVALID_ARG;
    for (size_t i = 0; i < m; ++i) sig[i * m + i] = 1; /* sigma = identity (:374-377) */
    size_t maxdeg = 0;
    for (size_t k = 0; k < order; ++k) {
        /* selection sort of the shift, same swaps on degree (:388-399) */
        for (size_t i = 0; i < m; ++i) {
            size_t im = i;
            for (size_t j = i + 1; j < m; ++j)
                if (shift[j] < shift[im]) im = j;
            uint64_t t = shift[i]; shift[i] = shift[im]; shift[im] = t;
            t = degree[i]; degree[i] = degree[im]; degree[im] = t;
            perm[i] = im;
        }
        const size_t lim = k < maxdeg ? k : maxdeg;
        /* Bperm on the rows of sigma[0..lim] (:418-444) */
        for (size_t l = 0; l <= lim; ++l)
            for (size_t i = 0; i < m; ++i)
                if (perm[i] != i)
                    for (size_t j = 0; j < m; ++j) {
                        uint32_t t = sig[l * mm + i * m + j];
                        sig[l * mm + i * m + j] = sig[l * mm + perm[i] * m + j];
                        sig[l * mm + perm[i] * m + j] = t;
                    }
        /* discrepancy: coefficient k of sigma.serie (:420-443) */
        for (size_t i = 0; i < m; ++i)
            for (size_t c = 0; c < n; ++c) {
                uint64_t acc = 0;
                for (size_t l = 0; l <= lim; ++l)
                    for (size_t j = 0; j < m; ++j)
                        acc = (acc + (uint64_t)sig[l * mm + i * m + j] * serie[(j * n + c) * ssz + (k - l)]) % p;
                delta[i * n + c] = (uint32_t)acc;
            }
        size_t rank;
        int st = pluq_rows(p, m, n, delta, rows, &rank, Y);
        if (st) return st;
        /* Qt on sigma[0..lim], then non-pivot rows -= (L2 L1^-1) pivot rows (:509-517) */
        for (size_t l = 0; l <= lim; ++l) {
            uint32_t *S = sig + l * mm;
            for (size_t s = 0; s < m; ++s) memcpy(tmp + s * m, S + rows[s] * m, m * 4);
            for (size_t s = rank; s < m; ++s)
                for (size_t j = 0; j < m; ++j) {
                    uint64_t acc = tmp[s * m + j];
                    for (size_t t = 0; t < rank; ++t)
                        acc = (acc + (uint64_t)(p - Y[(s - rank) * m + t]) % p * tmp[t * m + j]) % p;
                    tmp[s * m + j] = (uint32_t)acc;
                }
            memcpy(S, tmp, mm * 4);
        }
        /* Qt on degree and shift, then the degree bookkeeping (:546-566) */
        for (size_t s = 0; s < m; ++s) t64[s] = degree[rows[s]];
        memcpy(degree, t64, m * 8);
        for (size_t s = 0; s < m; ++s) t64[s] = shift[rows[s]];
        memcpy(shift, t64, m * 8);
        uint64_t dmax = 0;
        for (size_t i = 0; i < rank; ++i) {
            if (degree[i] > dmax) dmax = degree[i];
            degree[i]++;
            shift[i]++;
        }
        for (size_t i = rank; i < m; ++i)
            if (degree[i] < dmax) degree[i] = dmax;
        size_t md = maxdeg > dmax + 1 ? maxdeg : (size_t)dmax + 1;
        maxdeg = md < order ? md : order;
        /* pivot rows times x (:573-580) */
        for (size_t l = maxdeg; l-- > 0;)
            for (size_t i = 0; i < rank; ++i) memcpy(sig + (l + 1) * mm + i * m, sig + l * mm + i * m, m * 4);
        for (size_t i = 0; i < rank; ++i) memset(sig + i * m, 0, m * 4);
    }
    const size_t so = order + 1;
    for (size_t e = 0; e < mm; ++e)
        for (size_t l = 0; l < so; ++l) sigma[e * so + l] = l <= maxdeg ? sig[l * mm + e] : 0;
    *deg_out = maxdeg;
    free(sig); free(tmp); free(delta); free(Y); free(degree); free(t64); free(perm); free(rows);
    return OR_OK;
}

/* product of polfirst polynomial matrices (a m x k of size sa, b k x n of size sb):
 * the FFT-prime route when 2^ceil(log2(sa+sb-1)) | p-1, the naive one otherwise
 * (both give the unique product) */
static int pm_mul(uint32_t p, size_t m, size_t k, size_t n, const uint32_t *a, size_t sa, const uint32_t *b,
                  size_t sb, uint32_t *c)
{
    size_t pts = 1;
    while (pts < sa + sb - 1) pts <<= 1;
    if (sa + sb > 2 && p < (1u << 31) && (p - 1) % pts == 0)
        return or_polymat_mul_fft_prime(p, m, k, n, a, sa, b, sb, c, 0);
    return or_polymat_mul_naive(p, m, k, n, a, sa, b, sb, c);
}

int or_pm_basis(uint32_t p, size_t m, size_t n, const uint32_t *serie, size_t ssz, size_t order, uint64_t *shift,
                uint32_t *sigma, size_t *deg_out)
{
    if (p < 2 || !m || !n || !order || ssz < order || !serie || !shift || !sigma) return OR_ERR_INVALID_ARG;
    if (order <= OR_MBASIS_THRESHOLD) return or_m_basis(p, m, n, serie, ssz, order, shift, sigma, deg_out);
    const size_t ord1 = order >> 1, ord2 = order - ord1, mm = m * m;
    uint32_t *s1 = malloc(mm * (ord1 + 1) * 4), *s2 = malloc(mm * (ord2 + 1) * 4);
    uint32_t *ser = malloc(m * n * order * 4), *ser2 = calloc(m * n * ord2, 4);
    if (!s1 || !s2 || !ser || !ser2) return OR_ERR_INVALID_ARG;
    size_t d1, d2;
    int st = or_pm_basis(p, m, n, serie, ssz, ord1, shift, s1, &d1); /* first half (:249-261) */
    if (st) return st;
    /* sigma1 as a polfirst matrix of size d1+1, serie truncated to order */
    uint32_t *s1c = malloc(mm * (d1 + 1) * 4);
    for (size_t e = 0; e < mm; ++e) memcpy(s1c + e * (d1 + 1), s1 + e * (ord1 + 1), (d1 + 1) * 4);
    for (size_t e = 0; e < m * n; ++e) memcpy(ser + e * order, serie + e * ssz, order * 4);
    /* serie update: coefficients ord1..order-1 of sigma1.serie (:274, midproductgen) */
    st = or_polymat_midproduct_naive(p, m, m, n, s1c, d1 + 1, ser, order, ser2, ord2, 1, ord1 + 1, order);
    if (st) return st;
    st = or_pm_basis(p, m, n, ser2, ord2, ord2, shift, s2, &d2); /* second half (:290) */
    if (st) return st;
    uint32_t *s2c = malloc(mm * (d2 + 1) * 4), *prod = malloc(mm * (d1 + d2 + 1) * 4);
    for (size_t e = 0; e < mm; ++e) memcpy(s2c + e * (d2 + 1), s2 + e * (ord2 + 1), (d2 + 1) * 4);
    st = pm_mul(p, m, m, m, s2c, d2 + 1, s1c, d1 + 1, prod); /* sigma = sigma2.sigma1 (:297) */
    if (st) return st;
    const size_t so = order + 1, d = d1 + d2;
    for (size_t e = 0; e < mm; ++e)
        for (size_t l = 0; l < so; ++l) sigma[e * so + l] = l <= d ? prod[e * (d + 1) + l] : 0;
    *deg_out = d;
    free(s1); free(s2); free(ser); free(ser2); free(s1c); free(s2c); free(prod);
    return OR_OK;
}
