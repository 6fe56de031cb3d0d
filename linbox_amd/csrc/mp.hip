// Multiprecision modulus route: PolynomialMatrixFFTMulDomain<Modular<integer>>::mul
// (matpoly-mult-fft-multiprecision.inl:611-630) over ZRing<integer>::mul_crtla
// (:141-277), and the few-primes route of PolynomialMatrixThreePrimesFFTMulDomain
// (three-primes.inl:67-191) for one-word moduli that are not FFT primes for the
// length.  The reference takes the few-primes domain for p < 2^29 and the
// multiprecision one above (wordsize.inl:62-83); the product is the same either
// way, and here every one-word modulus takes the few-primes route, which needs
// no multiword arithmetic (five times faster at 64x64, degree 4095).
//
// The integer product of the [0,P) representatives has coefficients below
// bound = 2 P^2 k min(sa,sb) (:616).  It is computed modulo FFT primes
// p_i < 2^31 with 2^(lpts+1) | p_i - 1, chosen in the deterministic order of
// RandomFFTPrime::generatePrimes (random-fftprime.h:107-133) until their
// product exceeds the bound, then recombined on the device:
//   1. residue kernel: every nlimbs-word coefficient -> its residue mod each p_i
//      (Horner over the 32-bit limbs with Shoup products);
//   2. one single-prime NTT product per p_i (the same kernels as lbx_polymat_mul_u32);
//   3. CRT kernel: Garner's mixed-radix digits v_i of each coefficient, then
//      X mod P = sum_i v_i (prod_{j<i} p_j mod P) in 32-bit limbs, reduced by
//      shift-and-subtract, written as nlimbs words.
// The reference picks ~23-bit primes for fgemm on doubles (:162-164); the
// primes here are 31-bit because every step is 32-bit integer arithmetic.
// Products are unique, so the output is identical.
#include <cstring>
#include <string>
#include <vector>

#include "lbx_ctx.h"
#include "mp_div.h"

namespace lbx {

constexpr int kMaxPrimes = 160;
constexpr int kMaxLimbs = 64;

struct MpPlan {
    std::vector<uint32_t> primes;
    unsigned lpts = 0;
    int nlimbs = 0;
    bool reduce = false;          // few-primes plans: some prime is below the modulus
    uint32_t *d_const = nullptr;  // device constant block, layout described at MpLayout / FpLayout
};

namespace {

typedef std::vector<uint32_t> Limbs;

void limbs_trim(Limbs &x)
{
    while (x.size() > 1 && x.back() == 0) x.pop_back();
}

void limbs_mul_small(Limbs &x, uint32_t v)
{
    uint64_t carry = 0;
    for (auto &w : x) {
        const uint64_t t = (uint64_t)w * v + carry;
        w = (uint32_t)t;
        carry = t >> 32;
    }
    if (carry) x.push_back((uint32_t)carry);
}

Limbs limbs_mul(const Limbs &a, const Limbs &b)
{
    Limbs z(a.size() + b.size(), 0);
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t carry = 0;
        for (size_t j = 0; j < b.size(); ++j) {
            const uint64_t t = (uint64_t)a[i] * b[j] + z[i + j] + carry;
            z[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        z[i + b.size()] = (uint32_t)carry;
    }
    limbs_trim(z);
    return z;
}

int limbs_cmp(const Limbs &a, const Limbs &b)
{
    const size_t n = a.size() > b.size() ? a.size() : b.size();
    for (size_t i = n; i-- > 0;) {
        const uint32_t x = i < a.size() ? a[i] : 0, y = i < b.size() ? b[i] : 0;
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

int limbs_bits(const Limbs &a)
{
    for (size_t i = a.size(); i-- > 0;)
        if (a[i]) return (int)(i * 32 + 32 - __builtin_clz(a[i]));
    return 0;
}

Limbs limbs_mod(Limbs x, const Limbs &P)
{
    const int bp = limbs_bits(P);
    for (int sh = limbs_bits(x) - bp; sh >= 0; --sh) {
        Limbs t(x.size() + 1, 0);
        for (size_t i = 0; i < P.size(); ++i) {
            const size_t w = i + sh / 32;
            if (w < t.size()) t[w] |= P[i] << (sh % 32);
            if (sh % 32 && w + 1 < t.size()) t[w + 1] |= P[i] >> (32 - sh % 32);
        }
        if (limbs_cmp(x, t) >= 0) {
            uint64_t borrow = 0;
            for (size_t i = 0; i < x.size(); ++i) {
                const uint64_t d = (uint64_t)x[i] - (i < t.size() ? t[i] : 0) - borrow;
                x[i] = (uint32_t)d;
                borrow = (d >> 32) & 1;
            }
        }
    }
    limbs_trim(x);
    return x;
}

typedef unsigned __int128 u128;

uint64_t mulmod_u64(uint64_t a, uint64_t b, uint64_t n) { return (uint64_t)((u128)a * b % n); }

uint64_t powmod_u64(uint64_t a, uint64_t e, uint64_t n)
{
    uint64_t r = 1 % n;
    a %= n;
    while (e) {
        if (e & 1) r = mulmod_u64(r, a, n);
        a = mulmod_u64(a, a, n);
        e >>= 1;
    }
    return r;
}

// Miller-Rabin with the first twelve prime bases: deterministic below 2^64.
bool is_prime(uint64_t n)
{
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t q : bases) {
        if (n == q) return true;
        if (n % q == 0) return false;
    }
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++s;
    }
    for (uint64_t a : bases) {
        uint64_t x = powmod_u64(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int r = 1; r < s && composite; ++r) {
            x = mulmod_u64(x, x, n);
            if (x == n - 1) composite = false;
        }
        if (composite) return false;
    }
    return true;
}

// random-fftprime.h:107-133 generatePrimes with compute_max_m (:161-166).
bool generate_primes(uint64_t pbound, const Limbs &prodbound, unsigned kmin, std::vector<uint32_t> &out)
{
    out.clear();
    Limbs prod{1};
    if (!kmin) ++kmin;
    const int bits = 64 - __builtin_clzll(pbound - 2);
    for (int k = bits - 1; k > (int)kmin; --k) {
        const uint64_t B = pbound - 1;
        int64_t m = (B & ((1ull << k) - 1)) ? (int64_t)(B >> k) : (int64_t)(B >> k) - 1;
        if (!(m & 1)) --m;
        for (; m > 0; m -= 2) {
            const uint64_t p = ((uint64_t)m << k) + 1;
            if (is_prime(p)) {
                if (out.size() >= (size_t)kMaxPrimes) return false;
                out.push_back((uint32_t)p);
                limbs_mul_small(prod, (uint32_t)p);
                if (limbs_cmp(prod, prodbound) > 0) return true;
            }
        }
    }
    return false;
}

// Device constant block (32-bit words):
//   [0..r)         primes p_i
//   [r..2r)        floor(2^32/p_i)            (Shoup constant of 1)
//   [2r..3r)       2^32 mod p_i
//   [3r..4r)       Shoup constant of 2^32 mod p_i
//   [4r..4r+2r^2)  Garner inverses (p_j mod p_i)^-1 and their Shoup constants, i*r+j
//   then r*nlimbs  prod_{j<i} p_j mod P, then nlimbs words of P;
//   for the register kernels (nq = their word count NLM, else 0): off and s,
//   Q = P 2^(32 off + s) in nq words, and r*nq words of
//   (prod_{j<i} p_j mod P) 2^(32 off + s) (mp_div.h).
struct MpLayout {
    int r, nl, nq = 0;
    __host__ __device__ int primes() const { return 0; }
    __host__ __device__ int onep() const { return r; }
    __host__ __device__ int r32() const { return 2 * r; }
    __host__ __device__ int r32p() const { return 3 * r; }
    __host__ __device__ int inv() const { return 4 * r; }
    __host__ __device__ int mi() const { return 4 * r + 2 * r * r; }
    __host__ __device__ int P() const { return mi() + r * nl; }
    __host__ __device__ int sh() const { return P() + nl; }
    __host__ __device__ int Q() const { return sh() + 2; }
    __host__ __device__ int mq() const { return Q() + nq; }
    __host__ __device__ int words() const { return mq() + r * nq; }
};

// Register widths (RM primes, NLM words) of the CRT kernel mp_product
// launches for r primes and an nl-word modulus, {0, 0} for the general
// kernel; a plan's constants are laid out for that width (nq = NLM).  Narrow
// moduli get narrow kernels: a word-size P (3 primes) takes 2-word
// arithmetic and remainders instead of 8-word ones.
struct CrtWidth {
    int rm, nlm;
};
inline CrtWidth crt_width(int r, int nl)
{
    if (r <= 4 && nl <= 2) return {4, 2};
    if (r <= 8 && nl <= 4) return {8, 4};
    if (r <= 16 && nl <= 8) return {16, 8};
    if (r <= 32 && nl <= 16) return {32, 16};
    return {0, 0};
}

// Word index of element idx (entry idx / size, coefficient idx % size) of a
// view.  A polfirst view without window (stride = size, no shift) maps idx to
// itself, which spares the per-element 64-bit division.
__device__ __forceinline__ size_t elem_index(const DevView &v, size_t idx, size_t size)
{
    if (v.layout == 0 && v.stor == size && v.shift == 0) return idx;
    const size_t e = idx / size, t = idx - e * size;
    return view_index(v, e, t);
}

// Residues of nlimbs-word coefficients modulo every prime:
// out[i][e*size + t] for e < entries, t < size.
// The limb count is bounded at compile time (NLM >= nl), so the limbs stay in
// registers and the Horner steps unroll.
template <int NLM>
__global__ void __launch_bounds__(256) mp_residue_kernel(DevView in, int nl, size_t entries, size_t size,
                                                         const uint32_t *__restrict__ K, int r,
                                                         uint32_t *__restrict__ out)
{
    const size_t total = entries * size;
    const MpLayout lay{r, nl};
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        const uint32_t *src = in.base + elem_index(in, idx, size) * nl;
        uint32_t limb[NLM];
#pragma unroll
        for (int l = 0; l < NLM; ++l) limb[l] = l < nl ? src[l] : 0u;
        for (int i = 0; i < r; ++i) {
            const uint32_t p = K[lay.primes() + i], onep = K[lay.onep() + i];
            const uint32_t c = K[lay.r32() + i], cp = K[lay.r32p() + i];
            uint32_t v = 0;
#pragma unroll
            for (int l = NLM - 1; l >= 0; --l)
                if (l < nl) v = add_mod(shoup_mul(v, c, cp, p), shoup_mul(limb[l], 1u, onep, p), p);
            out[(size_t)i * total + idx] = v;
        }
    }
}

// Garner recombination of r residues and reduction mod P.
__global__ void __launch_bounds__(128) mp_crt_kernel(const uint32_t *__restrict__ res, size_t entries, size_t size,
                                                     const uint32_t *__restrict__ K, int r, int nl, DevView out)
{
    const size_t total = entries * size;
    const MpLayout lay{r, nl};
    for (size_t idx = (size_t)blockIdx.x * 128 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 128) {
        uint32_t v[kMaxPrimes];
        uint32_t acc[kMaxLimbs + 2];
        for (int l = 0; l < nl + 2; ++l) acc[l] = 0;
        for (int i = 0; i < r; ++i) {
            const uint32_t p = K[lay.primes() + i], onep = K[lay.onep() + i];
            uint32_t t = res[(size_t)i * total + idx];
            for (int j = 0; j < i; ++j) {
                const uint32_t vj = shoup_mul(v[j], 1u, onep, p);
                t = shoup_mul(sub_mod(t, vj, p), K[lay.inv() + 2 * (i * r + j)], K[lay.inv() + 2 * (i * r + j) + 1], p);
            }
            v[i] = t;
            // acc += v_i * (prod_{j<i} p_j mod P)
            uint64_t carry = 0;
            const uint32_t *mi = K + lay.mi() + i * nl;
            for (int l = 0; l < nl; ++l) {
                const uint64_t s = (uint64_t)mi[l] * t + acc[l] + carry;
                acc[l] = (uint32_t)s;
                carry = s >> 32;
            }
            for (int l = nl; l < nl + 2; ++l) {
                const uint64_t s = (uint64_t)acc[l] + carry;
                acc[l] = (uint32_t)s;
                carry = s >> 32;
            }
        }
        // acc < r 2^31 P <= 2^37 P: binary long division by P << sh, sh = 40..0
        const uint32_t *P = K + lay.P();
        for (int sh = 40; sh >= 0; --sh) {
            const int ws = sh >> 5, bs = sh & 31;
            uint32_t shifted[kMaxLimbs + 2];
            for (int l = 0; l < nl + 2; ++l) {
                const int src = l - ws;
                uint32_t w = 0;
                if (src >= 0 && src < nl) w = P[src] << bs;
                if (bs && src - 1 >= 0 && src - 1 < nl) w |= P[src - 1] >> (32 - bs);
                shifted[l] = w;
            }
            // P < 2^(32 nl) so P << sh (sh <= 40) fits the nl+2 words of acc
            bool ge = true;
            for (int l = nl + 1; l >= 0; --l) {
                if (acc[l] != shifted[l]) {
                    ge = acc[l] > shifted[l];
                    break;
                }
            }
            if (ge) {
                uint64_t borrow = 0;
                for (int l = 0; l < nl + 2; ++l) {
                    const uint64_t d = (uint64_t)acc[l] - shifted[l] - borrow;
                    acc[l] = (uint32_t)d;
                    borrow = (d >> 32) & 1;
                }
            }
        }
        uint32_t *dst = out.base + elem_index(out, idx, size) * nl;
        for (int l = 0; l < nl; ++l) dst[l] = acc[l];
    }
}

// The same recombination with the prime count and the limb count bounded at
// compile time (RM, NLM): every array index is then a constant, so v and the
// accumulator stay in registers (the general kernel above indexes them at run
// time and keeps them in scratch memory).  The accumulator is formed against
// P moved to the top of NLM words (Q = P 2^(32 off + s), normalized), reduced
// by two quotient digits of Knuth's division (mp_div.h) instead of 41 steps of
// shift-and-subtract, and shifted back into nl words.
template <int RM, int NLM>
__global__ void __launch_bounds__(128) mp_crt_small_kernel(const uint32_t *__restrict__ res, size_t entries,
                                                           size_t size, const uint32_t *__restrict__ K, int r, int nl,
                                                           DevView out)
{
    const size_t total = entries * size;
    const MpLayout lay{r, nl, NLM};
    const int off = (int)K[lay.sh()], s = (int)K[lay.sh() + 1];
    uint32_t q[NLM];
#pragma unroll
    for (int l = 0; l < NLM; ++l) q[l] = K[lay.Q() + l];
    const uint32_t qinv = div21_inverse(q[NLM - 1]);  // Q is normalized: its top word has its top bit set
    for (size_t idx = (size_t)blockIdx.x * 128 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 128) {
        uint32_t v[RM], rv[RM];
        uint32_t acc[NLM + 2];
#pragma unroll
        for (int l = 0; l < NLM + 2; ++l) acc[l] = 0;
#pragma unroll
        for (int i = 0; i < RM; ++i) rv[i] = i < r ? res[(size_t)i * total + idx] : 0u;  // loads issued together
#pragma unroll
        for (int i = 0; i < RM; ++i) {
            if (i < r) {
                const uint32_t p = K[lay.primes() + i], onep = K[lay.onep() + i];
                uint32_t t = rv[i];
#pragma unroll
                for (int j = 0; j < i; ++j) {
                    const uint32_t vj = shoup_mul(v[j], 1u, onep, p);
                    t = shoup_mul(sub_mod(t, vj, p), K[lay.inv() + 2 * (i * r + j)], K[lay.inv() + 2 * (i * r + j) + 1],
                                  p);
                }
                v[i] = t;
                // acc += v_i * (prod_{j<i} p_j mod P) 2^(32 off + s)
                const uint32_t *mi = K + lay.mq() + i * NLM;
                uint64_t carry = 0;
#pragma unroll
                for (int l = 0; l < NLM; ++l) {
                    const uint64_t sum = (uint64_t)mi[l] * t + acc[l] + carry;
                    acc[l] = (uint32_t)sum;
                    carry = sum >> 32;
                }
                const uint64_t top = (uint64_t)acc[NLM] + carry;
                acc[NLM] = (uint32_t)top;
                acc[NLM + 1] += (uint32_t)(top >> 32);
            }
        }
        // acc < r 2^31 Q <= 2^36 Q: acc mod Q = (X mod P) 2^(32 off + s) in acc[0..NLM)
        knuth_rem2<NLM>(acc, q, qinv);
        uint32_t *dst = out.base + elem_index(out, idx, size) * nl;
#pragma unroll
        for (int l = 0; l < NLM; ++l)
            if (l >= off) dst[l - off] = s ? (acc[l] >> s) | (acc[l + 1] << (32 - s)) : acc[l];
        for (int l = NLM - off; l < nl; ++l) dst[l] = 0;  // words of P's leading zero limbs
    }
}

// ---- few-primes route (three-primes.inl:67-191) ----------------------------
// Constants of a few-primes plan, r primes, modulus p < 2^32.
struct FpLayout {
    int r;
    __host__ __device__ int primes() const { return 0; }
    __host__ __device__ int onep() const { return r; }            // Shoup constant of 1 mod p_i
    __host__ __device__ int inv() const { return 2 * r; }         // p_j^-1 mod p_i and its Shoup constant
    __host__ __device__ int beta() const { return 2 * r + 2 * r * r; }  // prod_{j<i} p_j mod p, Shoup constant
    __host__ __device__ int words() const { return beta() + 2 * r; }
};

// Mixed-radix recombination (three-primes.inl:164-181): the Garner digits v_i
// of each coefficient from its residues res[i][idx] (c_i <- (c_i - c_j) p_j^-1
// mod p_i for j < i), then c = sum_i v_i beta_i mod p with beta_i = prod_{j<i}
// p_j mod p, written to the output view.  Every value is a 32-bit word: v_i <
// p_i < 2^31 and Shoup products take any 32-bit multiplicand.  WIDE (p >= 2^31):
// the Shoup remainder x*w - q*p < 2p and the sum are held in 64 bits.
template <int RM, bool WIDE>
__global__ void __launch_bounds__(256) few_primes_crt_kernel(const uint32_t *__restrict__ res, size_t entries,
                                                             size_t size, const uint32_t *__restrict__ K, int r,
                                                             uint32_t p, DevView out)
{
    const size_t total = entries * size;
    const FpLayout lay{r};
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        uint32_t v[RM], rv[RM];
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < RM; ++i) rv[i] = i < r ? res[(size_t)i * total + idx] : 0u;  // loads issued together
#pragma unroll
        for (int i = 0; i < RM; ++i) {
            if (i < r) {
                const uint32_t pi = K[lay.primes() + i], onep = K[lay.onep() + i];
                uint32_t t = rv[i];
#pragma unroll
                for (int j = 0; j < i; ++j) {
                    const uint32_t vj = shoup_mul(v[j], 1u, onep, pi);
                    t = shoup_mul(sub_mod(t, vj, pi), K[lay.inv() + 2 * (i * r + j)],
                                  K[lay.inv() + 2 * (i * r + j) + 1], pi);
                }
                v[i] = t;
                const uint32_t w = K[lay.beta() + 2 * i], wp = K[lay.beta() + 2 * i + 1];
                if constexpr (!WIDE) {
                    acc = add_mod(acc, shoup_mul(t, w, wp, p), p);
                } else {
                    uint64_t y = (uint64_t)t * w - (uint64_t)mulhi32(t, wp) * p;
                    if (y >= p) y -= p;
                    y += acc;
                    if (y >= p) y -= p;
                    acc = (uint32_t)y;
                }
            }
        }
        out.base[elem_index(out, idx, size)] = acc;
    }
}

// Residues of one-word coefficients modulo every prime of a few-primes plan,
// out[i][e*size + t] (the finit of three-primes.inl:143-146), for plans with a
// prime below the modulus.
__global__ void __launch_bounds__(256) fp_residue_kernel(DevView in, size_t entries, size_t size,
                                                         const uint32_t *__restrict__ K, int r,
                                                         uint32_t *__restrict__ out)
{
    const size_t total = entries * size;
    const FpLayout lay{r};
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        const uint32_t x = in.base[elem_index(in, idx, size)];
        for (int i = 0; i < r; ++i)
            out[(size_t)i * total + idx] = shoup_mul(x, 1u, K[lay.onep() + i], K[lay.primes() + i]);
    }
}

constexpr int kFewMaxPrimes = 8;

unsigned grid_of(size_t total, unsigned block)
{
    const size_t g = (total + block - 1) / block;
    return (unsigned)(g < 8192 ? (g ? g : 1) : 8192);
}

// The r per-prime products c_i = a_i b_i (polfirst, contiguous; operand i at
// a + i*astride) as one call of the batch kernels, when the pointwise form
// this shape takes is their plain one: a few launches in place of r separate
// products.  LBX_ERR_UNSUPPORTED (nothing launched) otherwise.
lbx_status primes_as_batch(lbx_ctx *ctx, const std::vector<uint32_t> &primes, const uint32_t *a, size_t astride,
                           const uint32_t *b, size_t bstride, uint32_t *c, size_t m, size_t k, size_t n, size_t sa,
                           size_t sb, unsigned lpts, int method, hipStream_t s)
{
    if (sa + sb <= 2 || choose_method(method, m, k, n, lpts) != LBX_METHOD_PLAIN_U64) return LBX_ERR_UNSUPPORTED;
    std::vector<lbx_job> jobs(primes.size());
    for (size_t i = 0; i < primes.size(); ++i)
        jobs[i] = lbx_job{primes[i], m, k, n, sa, sb, a + i * astride, b + i * bstride, c + i * m * n * (sa + sb - 1)};
    return batch_products(ctx, jobs.data(), jobs.size(), s);
}

constexpr int kBatchMinPrimes = 6;

lbx_status get_plan(lbx_ctx *ctx, const uint32_t *Pw, int nl, size_t k, size_t minsz, unsigned lpts, MpPlan **out)
{
    std::string key((const char *)Pw, (size_t)nl * 4);
    key += "|" + std::to_string(k) + "|" + std::to_string(minsz) + "|" + std::to_string(lpts);
    auto it = ctx->mp_plans.find(key);
    if (it != ctx->mp_plans.end()) {
        *out = it->second;
        return LBX_OK;
    }
    Limbs P(Pw, Pw + nl);
    limbs_trim(P);
    if (P.size() == 1 && P[0] < 2) return LBX_ERR_INVALID_ARG;
    Limbs bound = limbs_mul(P, P);  // :616 bound = 2 P^2 k min(sa,sb)
    limbs_mul_small(bound, 2);
    limbs_mul_small(bound, (uint32_t)k);
    limbs_mul_small(bound, (uint32_t)minsz);
    std::vector<uint32_t> primes;
    if (!generate_primes((uint64_t)1 << 31, bound, lpts, primes)) return LBX_ERR_NOT_ENOUGH_PRIMES;
    const int r = (int)primes.size();
    // word count of the register kernel mp_product launches for (r, nl), 0 for the general one
    const int nq = crt_width(r, nl).nlm;
    const MpLayout lay{r, nl, nq};
    std::vector<uint32_t> K((size_t)lay.words(), 0);
    // P moved to the top of nq words with its top bit set: 2^(32 off + s)
    const int np = (int)P.size(), off = nq - np, sbits = __builtin_clz(P.back());
    auto place = [&](const Limbs &x, uint32_t *dst) {  // x < P: x 2^(32 off + s) in nq words
        uint32_t prev = 0;
        for (int l = 0; l < np; ++l) {
            const uint32_t w = l < (int)x.size() ? x[l] : 0u;
            dst[off + l] = sbits ? (w << sbits) | prev : w;
            prev = sbits ? w >> (32 - sbits) : 0u;
        }
    };
    if (nq) {
        K[lay.sh()] = (uint32_t)off;
        K[lay.sh() + 1] = (uint32_t)sbits;
        place(P, &K[lay.Q()]);
    }
    Limbs M{1};
    for (int i = 0; i < r; ++i) {
        const uint32_t p = primes[i];
        K[lay.primes() + i] = p;
        K[lay.onep() + i] = shoup_pre(1, p);
        const uint32_t r32 = (uint32_t)(((uint64_t)1 << 32) % p);
        K[lay.r32() + i] = r32;
        K[lay.r32p() + i] = shoup_pre(r32, p);
        for (int j = 0; j < i; ++j) {
            const uint32_t inv = invmod_h(primes[j] % p, p);
            K[lay.inv() + 2 * (i * r + j)] = inv;
            K[lay.inv() + 2 * (i * r + j) + 1] = shoup_pre(inv, p);
        }
        const Limbs mi = limbs_mod(M, P);
        for (int l = 0; l < nl; ++l) K[lay.mi() + i * nl + l] = l < (int)mi.size() ? mi[l] : 0;
        if (nq) place(mi, &K[lay.mq() + i * nq]);
        limbs_mul_small(M, p);
        M = limbs_mod(M, P);
    }
    for (int l = 0; l < nl; ++l) K[lay.P() + l] = Pw[l];
    MpPlan *plan = new MpPlan();
    plan->primes = primes;
    plan->lpts = lpts;
    plan->nlimbs = nl;
    hipError_t e = hipMalloc(&plan->d_const, K.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(plan->d_const, K.data(), K.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (plan->d_const) (void)hipFree(plan->d_const);
        delete plan;
        return from_hip(e);
    }
    ctx->mp_plans[key] = plan;
    *out = plan;
    return LBX_OK;
}

// Plan of the few-primes route: bound = (p-1)^2 k min(sa,sb) (three-primes.inl:75-76,
// 180-181 of mul, :225-226 of midproduct), FFT primes below 2^31 with
// 2^lpts | p_i - 1 in the order of generatePrimes until their product exceeds
// the bound (:116-121).  r = 0 (the caller then takes the multiprecision
// route) when more than kFewMaxPrimes would be needed; `reduce` when some
// p_i <= p, whose operands need the reduction of :143-146.
lbx_status get_few_plan(lbx_ctx *ctx, uint32_t p, size_t k, size_t minsz, unsigned lpts, MpPlan **out)
{
    std::string key = "fp|" + std::to_string(p) + "|" + std::to_string(k) + "|" + std::to_string(minsz) + "|" +
                      std::to_string(lpts);
    auto it = ctx->mp_plans.find(key);
    if (it != ctx->mp_plans.end()) {
        *out = it->second;
        return LBX_OK;
    }
    Limbs bound{p - 1};
    limbs_mul_small(bound, p - 1);
    limbs_mul_small(bound, (uint32_t)k);
    limbs_mul_small(bound, (uint32_t)minsz);
    std::vector<uint32_t> primes;
    if (!generate_primes((uint64_t)1 << 31, bound, lpts, primes)) return LBX_ERR_NOT_ENOUGH_PRIMES;
    MpPlan *plan = new MpPlan();
    plan->lpts = lpts;
    plan->nlimbs = 1;
    for (uint32_t q : primes) plan->reduce = plan->reduce || q <= p;
    if ((int)primes.size() <= kFewMaxPrimes) {
        const int r = (int)primes.size();
        const FpLayout lay{r};
        std::vector<uint32_t> K((size_t)lay.words(), 0);
        uint32_t beta = 1 % p;
        for (int i = 0; i < r; ++i) {
            const uint32_t pi = primes[i];
            K[lay.primes() + i] = pi;
            K[lay.onep() + i] = shoup_pre(1, pi);
            for (int j = 0; j < i; ++j) {
                const uint32_t inv = invmod_h(primes[j] % pi, pi);
                K[lay.inv() + 2 * (i * r + j)] = inv;
                K[lay.inv() + 2 * (i * r + j) + 1] = shoup_pre(inv, pi);
            }
            K[lay.beta() + 2 * i] = beta;
            K[lay.beta() + 2 * i + 1] = shoup_pre(beta, p);
            beta = (uint32_t)((uint64_t)beta * (pi % p) % p);  // beta *= p_i (:179-180)
        }
        hipError_t e = hipMalloc(&plan->d_const, K.size() * 4);
        if (e == hipSuccess) e = hipMemcpy(plan->d_const, K.data(), K.size() * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (plan->d_const) (void)hipFree(plan->d_const);
            delete plan;
            return from_hip(e);
        }
        plan->primes = primes;
    }
    ctx->mp_plans[key] = plan;
    *out = plan;
    return LBX_OK;
}

}  // namespace

void mp_release(lbx_ctx *ctx)
{
    for (auto &kv : ctx->mp_plans) {
        if (kv.second->d_const) (void)hipFree(kv.second->d_const);
        delete kv.second;
    }
    ctx->mp_plans.clear();
}

lbx_status mp_product(lbx_ctx *ctx, const uint32_t *Pw, int nl, const DevView &a, const DevView &b, const DevView &c,
                      size_t m, size_t k, size_t n, int method, hipStream_t s, const MidSpec *mid, unsigned mid_L,
                      size_t count)
{
    // product: s = max_rowdeg+1 coefficients when given, else sa+sb-1 (mul_crtla :150-152;
    // three-primes.inl:69-73); with fewer than sa+sb-1 the cyclic product mod x^pts - 1
    const size_t sa = a.size, sb = b.size, sc = mid ? c.size : (count ? count : sa + sb - 1);
    const bool full = !mid && sc == sa + sb - 1;
    unsigned lpts = 0;
    size_t pts = 1;
    while (pts < sc) {  // mul_crtla :151-154
        pts <<= 1;
        ++lpts;
    }
    if (mid) {  // midproduct_crtla :446-447: points from deg, not from the output size
        lpts = mid_L;
        pts = (size_t)1 << lpts;
    }
    const size_t minsz = sa < sb ? sa : sb;
    if (nl == 1 && Pw[0] >= 2) {
        // few-primes route (wordsize.inl:66-68): with every p_i > p the operands
        // are already residues mod p_i (three-primes.inl:138-142), so the
        // per-prime products read them where they are; else one pass writes
        // their residues first (:143-146)
        MpPlan *fp;
        LBX_TRY(get_few_plan(ctx, Pw[0], k, minsz, lpts, &fp));
        const int r = (int)fp->primes.size();
        if (r > 0) {
            void *ra = nullptr, *rb = nullptr, *rc;
            if (fp->reduce) {
                LBX_TRY(ensure(ctx, "mp_ra", (size_t)r * m * k * sa * 4, &ra));
                LBX_TRY(ensure(ctx, "mp_rb", (size_t)r * k * n * sb * 4, &rb));
                hipLaunchKernelGGL(fp_residue_kernel, dim3(grid_of(m * k * sa, 256)), dim3(256), 0, s, a, m * k, sa,
                                   fp->d_const, r, (uint32_t *)ra);
                LBX_TRY_HIP(hipGetLastError());
                hipLaunchKernelGGL(fp_residue_kernel, dim3(grid_of(k * n * sb, 256)), dim3(256), 0, s, b, k * n, sb,
                                   fp->d_const, r, (uint32_t *)rb);
                LBX_TRY_HIP(hipGetLastError());
            }
            LBX_TRY(ensure(ctx, "mp_rc", (size_t)r * m * n * sc * 4, &rc));
            for (int i = 0; i < r; ++i) {
                DevView va = a, vb = b;
                if (fp->reduce) {
                    va = DevView{(uint32_t *)ra + (size_t)i * m * k * sa, LBX_POLFIRST, m, k, sa, 0, sa};
                    vb = DevView{(uint32_t *)rb + (size_t)i * k * n * sb, LBX_POLFIRST, k, n, sb, 0, sb};
                }
                DevView vc{(uint32_t *)rc + (size_t)i * m * n * sc, LBX_POLFIRST, m, n, sc, 0, sc};
                if (mid)
                    LBX_TRY(fft_prime_product(ctx, fp->primes[i], lpts, va, vb, vc, m, k, n, sc, method, s, mid));
                else if (sa + sb <= 2 && sc == 1)
                    LBX_TRY(constant_product(ctx, fp->primes[i], va, vb, vc, m, k, n, s));
                else
                    LBX_TRY(fft_prime_product(ctx, fp->primes[i], lpts, va, vb, vc, m, k, n, sc, method, s));
            }
            const dim3 grid(grid_of(m * n * sc, 256));
            if (Pw[0] >= (1u << 31))
                hipLaunchKernelGGL((few_primes_crt_kernel<kFewMaxPrimes, true>), grid, dim3(256), 0, s,
                                   (const uint32_t *)rc, m * n, sc, fp->d_const, r, Pw[0], c);
            else
                hipLaunchKernelGGL((few_primes_crt_kernel<kFewMaxPrimes, false>), grid, dim3(256), 0, s,
                                   (const uint32_t *)rc, m * n, sc, fp->d_const, r, Pw[0], c);
            LBX_TRY_HIP(hipGetLastError());
            return LBX_OK;
        }
    }
    MpPlan *plan;
    LBX_TRY(get_plan(ctx, Pw, nl, k, minsz, lpts, &plan));
    const int r = (int)plan->primes.size();
    void *ra, *rb, *rc;
    LBX_TRY(ensure(ctx, "mp_ra", (size_t)r * m * k * sa * 4, &ra));
    LBX_TRY(ensure(ctx, "mp_rb", (size_t)r * k * n * sb * 4, &rb));
    LBX_TRY(ensure(ctx, "mp_rc", (size_t)r * m * n * sc * 4, &rc));
    auto residues = nl <= 8 ? mp_residue_kernel<8> : mp_residue_kernel<kMaxLimbs>;
    hipLaunchKernelGGL(residues, dim3(grid_of(m * k * sa, 256)), dim3(256), 0, s, a, nl, m * k, sa, plan->d_const, r,
                       (uint32_t *)ra);
    LBX_TRY_HIP(hipGetLastError());
    hipLaunchKernelGGL(residues, dim3(grid_of(k * n * sb, 256)), dim3(256), 0, s, b, nl, k * n, sb, plan->d_const, r,
                       (uint32_t *)rb);
    LBX_TRY_HIP(hipGetLastError());
    // many primes (multiword moduli): the products as one batch; a new shape
    // builds and caches a batch plan (graph capture), which pays off here
    lbx_status bst = LBX_ERR_UNSUPPORTED;
    if (full && r >= kBatchMinPrimes)
        bst = primes_as_batch(ctx, plan->primes, (const uint32_t *)ra, m * k * sa, (const uint32_t *)rb, k * n * sb,
                              (uint32_t *)rc, m, k, n, sa, sb, lpts, method, s);
    if (bst != LBX_ERR_UNSUPPORTED) LBX_TRY(bst);
    for (int i = 0; i < r && bst == LBX_ERR_UNSUPPORTED; ++i) {
        DevView va{(uint32_t *)ra + (size_t)i * m * k * sa, LBX_POLFIRST, m, k, sa, 0, sa};
        DevView vb{(uint32_t *)rb + (size_t)i * k * n * sb, LBX_POLFIRST, k, n, sb, 0, sb};
        DevView vc{(uint32_t *)rc + (size_t)i * m * n * sc, LBX_POLFIRST, m, n, sc, 0, sc};
        if (mid)
            LBX_TRY(fft_prime_product(ctx, plan->primes[i], lpts, va, vb, vc, m, k, n, sc, method, s, mid));
        else if (sa + sb <= 2 && sc == 1)
            LBX_TRY(constant_product(ctx, plan->primes[i], va, vb, vc, m, k, n, s));
        else
            LBX_TRY(fft_prime_product(ctx, plan->primes[i], lpts, va, vb, vc, m, k, n, sc, method, s));
    }
    const dim3 grid(grid_of(m * n * sc, 128));
    const int nlm = crt_width(r, nl).nlm;
    if (nlm == 2)
        hipLaunchKernelGGL((mp_crt_small_kernel<4, 2>), grid, dim3(128), 0, s, (const uint32_t *)rc, m * n, sc,
                           plan->d_const, r, nl, c);
    else if (nlm == 4)
        hipLaunchKernelGGL((mp_crt_small_kernel<8, 4>), grid, dim3(128), 0, s, (const uint32_t *)rc, m * n, sc,
                           plan->d_const, r, nl, c);
    else if (nlm == 8)
        hipLaunchKernelGGL((mp_crt_small_kernel<16, 8>), grid, dim3(128), 0, s, (const uint32_t *)rc, m * n, sc,
                           plan->d_const, r, nl, c);
    else if (nlm == 16)
        hipLaunchKernelGGL((mp_crt_small_kernel<32, 16>), grid, dim3(128), 0, s, (const uint32_t *)rc, m * n, sc,
                           plan->d_const, r, nl, c);
    else
        hipLaunchKernelGGL(mp_crt_kernel, grid, dim3(128), 0, s, (const uint32_t *)rc, m * n, sc, plan->d_const, r,
                           nl, c);
    LBX_TRY_HIP(hipGetLastError());
    return LBX_OK;
}

// ---- products over the integers: PolynomialMatrixFFTMulDomain<ZRing<integer>> ----
// (matpoly-mult-fft-multiprecision.inl:43-89 with mul_crtla, :141-277).
// Coefficients are nl-word (inputs) and cl-word (output) two's complement
// integers.  As in the reference, maxA = 2^(largest bit size of a's
// coefficients), likewise maxB, and bound = 2 maxA maxB k min(sa,sb) (:78-86);
// the product is taken modulo FFT primes whose product M exceeds the bound and
// reconstructed in the symmetric range.  The primes are odd, so the mixed-radix
// digits of (M-1)/2 are all (p_i-1)/2: a value is negative (X > (M-1)/2) when
// its Garner digits, read from the most significant, first exceed them.  The
// output X - [X > (M-1)/2] M is formed modulo 2^(32 cl), which holds it.
constexpr int kZzMaxIn = 32;      // words of an input coefficient
constexpr int kZzMaxOut = 80;     // words of an output coefficient
constexpr int kZzMaxPrimes = 96;

namespace {

// Constant block: primes, Shoup constants of 1, 2^32 mod p_i and its Shoup
// constant, 2^(32 nl) mod p_i, Garner inverses (p_j mod p_i)^-1 with their
// Shoup constants at 2(i r + j), then W_i = prod_{j<i} p_j mod 2^(32 cl) in cl
// words each, then M mod 2^(32 cl).
struct ZzLayout {
    int r, cl;
    __host__ __device__ int primes() const { return 0; }
    __host__ __device__ int onep() const { return r; }
    __host__ __device__ int r32() const { return 2 * r; }
    __host__ __device__ int r32p() const { return 3 * r; }
    __host__ __device__ int top() const { return 4 * r; }
    __host__ __device__ int inv() const { return 5 * r; }
    __host__ __device__ int W() const { return 5 * r + 2 * r * r; }
    __host__ __device__ int M() const { return W() + r * cl; }
    __host__ __device__ int words() const { return M() + cl; }
};

// Largest bit size of the absolute values of nl-word two's complement
// coefficients (Givaro's bitsize, 0 for 0), max-accumulated into *out.
__global__ void __launch_bounds__(256) zz_logmax_kernel(DevView in, int nl, size_t entries, size_t size,
                                                        unsigned *__restrict__ out)
{
    const size_t total = entries * size;
    unsigned best = 0;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        const uint32_t *src = in.base + elem_index(in, idx, size) * nl;
        const bool neg = src[nl - 1] >> 31;
        // |x|: the words themselves, or their two's complement negation
        uint32_t carry = 1;
        unsigned bits = 0;
        for (int l = 0; l < nl; ++l) {
            uint32_t w = src[l];
            if (neg) {
                const uint32_t t = ~w + carry;
                carry = carry && t == 0;
                w = t;
            }
            if (w) bits = 32 * l + 32 - __builtin_clz(w);
        }
        best = bits > best ? bits : best;
    }
    for (int off = 32; off; off >>= 1) {
        const unsigned o = __shfl_xor(best, off);
        best = o > best ? o : best;
    }
    if ((threadIdx.x & 63) == 0 && best) atomicMax(out, best);
}

// Residues mod every prime of nl-word two's complement coefficients: Horner over
// the words as unsigned, minus 2^(32 nl) when the sign bit is set.
template <int NLM>
__global__ void __launch_bounds__(256) zz_residue_kernel(DevView in, int nl, size_t entries, size_t size,
                                                         const uint32_t *__restrict__ K, int r,
                                                         uint32_t *__restrict__ out)
{
    const size_t total = entries * size;
    const ZzLayout lay{r, 0};
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        const uint32_t *src = in.base + elem_index(in, idx, size) * nl;
        uint32_t limb[NLM];
#pragma unroll
        for (int l = 0; l < NLM; ++l) limb[l] = l < nl ? src[l] : 0u;
        const bool neg = src[nl - 1] >> 31;
        for (int i = 0; i < r; ++i) {
            const uint32_t p = K[lay.primes() + i], onep = K[lay.onep() + i];
            const uint32_t c = K[lay.r32() + i], cp = K[lay.r32p() + i];
            uint32_t v = 0;
#pragma unroll
            for (int l = NLM - 1; l >= 0; --l)
                if (l < nl) v = add_mod(shoup_mul(v, c, cp, p), shoup_mul(limb[l], 1u, onep, p), p);
            if (neg) v = sub_mod(v, K[lay.top() + i], p);
            out[(size_t)i * total + idx] = v;
        }
    }
}

// Garner digits of each coefficient's r residues, the sign test against the
// digits (p_i-1)/2, and X - [negative] M modulo 2^(32 cl) in cl words.
__global__ void __launch_bounds__(128) zz_crt_kernel(const uint32_t *__restrict__ res, size_t entries, size_t size,
                                                     const uint32_t *__restrict__ K, int r, int cl, DevView out)
{
    const size_t total = entries * size;
    const ZzLayout lay{r, cl};
    for (size_t idx = (size_t)blockIdx.x * 128 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 128) {
        uint32_t v[kZzMaxPrimes];
        uint32_t acc[kZzMaxOut];
        for (int l = 0; l < cl; ++l) acc[l] = 0;
        for (int i = 0; i < r; ++i) {
            const uint32_t p = K[lay.primes() + i], onep = K[lay.onep() + i];
            uint32_t t = res[(size_t)i * total + idx];
            for (int j = 0; j < i; ++j) {
                const uint32_t vj = shoup_mul(v[j], 1u, onep, p);
                t = shoup_mul(sub_mod(t, vj, p), K[lay.inv() + 2 * (i * r + j)], K[lay.inv() + 2 * (i * r + j) + 1], p);
            }
            v[i] = t;
            // acc += v_i W_i (mod 2^(32 cl))
            const uint32_t *w = K + lay.W() + i * cl;
            uint64_t carry = 0;
            for (int l = 0; l < cl; ++l) {
                const uint64_t s = (uint64_t)w[l] * t + acc[l] + carry;
                acc[l] = (uint32_t)s;
                carry = s >> 32;
            }
        }
        bool neg = false;
        for (int i = r - 1; i >= 0; --i) {
            const uint32_t h = (K[lay.primes() + i] - 1) / 2;
            if (v[i] != h) {
                neg = v[i] > h;
                break;
            }
        }
        if (neg) {
            const uint32_t *Mm = K + lay.M();
            uint64_t borrow = 0;
            for (int l = 0; l < cl; ++l) {
                const uint64_t d = (uint64_t)acc[l] - Mm[l] - borrow;
                acc[l] = (uint32_t)d;
                borrow = (d >> 32) & 1;
            }
        }
        uint32_t *dst = out.base + elem_index(out, idx, size) * cl;
        for (int l = 0; l < cl; ++l) dst[l] = acc[l];
    }
}

// Largest coefficient bit size of a view (one pass, one read-back).
lbx_status zz_logmax(lbx_ctx *ctx, const DevView &v, int nl, unsigned *bits, hipStream_t s)
{
    void *w = nullptr;
    LBX_TRY(ensure(ctx, "zz_logmax", 256, &w));
    LBX_TRY_HIP(hipMemsetAsync(w, 0, 4, s));
    const size_t entries = v.rows * v.cols;
    hipLaunchKernelGGL(zz_logmax_kernel, dim3(grid_of(entries * v.size, 256)), dim3(256), 0, s, v, nl, entries, v.size,
                       (unsigned *)w);
    LBX_TRY_HIP(hipGetLastError());
    LBX_TRY_HIP(hipMemcpyAsync(bits, w, 4, hipMemcpyDeviceToHost, s));
    LBX_TRY_HIP(hipStreamSynchronize(s));
    return LBX_OK;
}

}  // namespace

lbx_status zz_product(lbx_ctx *ctx, int nl, int cl, const DevView &a, const DevView &b, const DevView &c, size_t m,
                      size_t k, size_t n, size_t sc, int method, hipStream_t s)
{
    const size_t sa = a.size, sb = b.size;
    const bool full = sc == sa + sb - 1;
    unsigned lpts = 0;
    size_t pts = 1;
    while (pts < sc) {  // mul_crtla :151-154
        pts <<= 1;
        ++lpts;
    }
    unsigned la = 0, lb = 0;
    LBX_TRY(zz_logmax(ctx, a, nl, &la, s));
    LBX_TRY(zz_logmax(ctx, b, nl, &lb, s));
    // bound = 2 maxA maxB k min(sa,sb), maxA = 2^la, maxB = 2^lb (:78-86);
    // |coefficient| < bound/2 must fit cl words with its sign
    Limbs bound{1};
    for (unsigned i = 0; i < la + lb + 1; ++i) limbs_mul_small(bound, 2);
    limbs_mul_small(bound, (uint32_t)k);
    limbs_mul_small(bound, (uint32_t)(sa < sb ? sa : sb));
    limbs_trim(bound);
    size_t bits = 32 * (bound.size() - 1) + 32 - __builtin_clz(bound.back());  // bound < 2^bits
    if (bits > (size_t)32 * cl) return LBX_ERR_SIZE_MISMATCH;  // bound/2 < 2^(bits-1) needs bits-1 bits and a sign
    std::vector<uint32_t> primes;
    if (!generate_primes((uint64_t)1 << 31, bound, lpts, primes) || (int)primes.size() > kZzMaxPrimes)
        return LBX_ERR_NOT_ENOUGH_PRIMES;
    const int r = (int)primes.size();
    const ZzLayout lay{r, cl};
    std::vector<uint32_t> K((size_t)lay.words(), 0);
    Limbs W{1};  // prod_{j<i} p_j, kept below 2^(32 cl)
    for (int i = 0; i < r; ++i) {
        const uint32_t p = primes[i];
        K[lay.primes() + i] = p;
        K[lay.onep() + i] = shoup_pre(1, p);
        const uint32_t r32 = (uint32_t)(((uint64_t)1 << 32) % p);
        K[lay.r32() + i] = r32;
        K[lay.r32p() + i] = shoup_pre(r32, p);
        uint64_t top = 1;
        for (int l = 0; l < nl; ++l) top = top * r32 % p;
        K[lay.top() + i] = (uint32_t)top;
        for (int j = 0; j < i; ++j) {
            const uint32_t inv = invmod_h(primes[j] % p, p);
            K[lay.inv() + 2 * (i * r + j)] = inv;
            K[lay.inv() + 2 * (i * r + j) + 1] = shoup_pre(inv, p);
        }
        for (int l = 0; l < cl; ++l) K[lay.W() + i * cl + l] = l < (int)W.size() ? W[l] : 0;
        limbs_mul_small(W, p);
        if ((int)W.size() > cl) W.resize(cl);
    }
    for (int l = 0; l < cl; ++l) K[lay.M() + l] = l < (int)W.size() ? W[l] : 0;
    void *dk, *ra, *rb, *rc;
    LBX_TRY(ensure(ctx, "zz_const", K.size() * 4, &dk));
    // the stream is idle here (zz_logmax waited for it): a blocking copy from K
    LBX_TRY_HIP(hipMemcpy(dk, K.data(), K.size() * 4, hipMemcpyHostToDevice));
    LBX_TRY(ensure(ctx, "mp_ra", (size_t)r * m * k * sa * 4, &ra));
    LBX_TRY(ensure(ctx, "mp_rb", (size_t)r * k * n * sb * 4, &rb));
    LBX_TRY(ensure(ctx, "mp_rc", (size_t)r * m * n * sc * 4, &rc));
    auto residues = nl <= 8 ? zz_residue_kernel<8> : zz_residue_kernel<kZzMaxIn>;
    hipLaunchKernelGGL(residues, dim3(grid_of(m * k * sa, 256)), dim3(256), 0, s, a, nl, m * k, sa,
                       (const uint32_t *)dk, r, (uint32_t *)ra);
    LBX_TRY_HIP(hipGetLastError());
    hipLaunchKernelGGL(residues, dim3(grid_of(k * n * sb, 256)), dim3(256), 0, s, b, nl, k * n, sb,
                       (const uint32_t *)dk, r, (uint32_t *)rb);
    LBX_TRY_HIP(hipGetLastError());
    lbx_status bst = LBX_ERR_UNSUPPORTED;
    if (full && r >= kBatchMinPrimes)
        bst = primes_as_batch(ctx, primes, (const uint32_t *)ra, m * k * sa, (const uint32_t *)rb, k * n * sb,
                              (uint32_t *)rc, m, k, n, sa, sb, lpts, method, s);
    if (bst != LBX_ERR_UNSUPPORTED) LBX_TRY(bst);
    for (int i = 0; i < r && bst == LBX_ERR_UNSUPPORTED; ++i) {
        DevView va{(uint32_t *)ra + (size_t)i * m * k * sa, LBX_POLFIRST, m, k, sa, 0, sa};
        DevView vb{(uint32_t *)rb + (size_t)i * k * n * sb, LBX_POLFIRST, k, n, sb, 0, sb};
        DevView vc{(uint32_t *)rc + (size_t)i * m * n * sc, LBX_POLFIRST, m, n, sc, 0, sc};
        if (sa + sb <= 2 && sc == 1)
            LBX_TRY(constant_product(ctx, primes[i], va, vb, vc, m, k, n, s));
        else
            LBX_TRY(fft_prime_product(ctx, primes[i], lpts, va, vb, vc, m, k, n, sc, method, s));
    }
    hipLaunchKernelGGL(zz_crt_kernel, dim3(grid_of(m * n * sc, 128)), dim3(128), 0, s, (const uint32_t *)rc, m * n,
                       sc, (const uint32_t *)dk, r, cl, c);
    LBX_TRY_HIP(hipGetLastError());
    return LBX_OK;
}

}  // namespace lbx

using namespace lbx;

extern "C" lbx_status lbx_polymat_mul_mp(lbx_ctx *ctx, const uint32_t *P, int nlimbs, lbx_polmat *c,
                                         const lbx_polmat *a, const lbx_polmat *b, size_t max_rowdeg, int method,
                                         void *stream)
{
    return guarded([&]() -> lbx_status {
        if (!ctx || !P || !a || !b || !c) return LBX_ERR_INVALID_ARG;
        if (nlimbs < 1 || nlimbs > kMaxLimbs) return LBX_ERR_PRIME_TOO_LARGE;
        if (method < 0 || method > 2) return LBX_ERR_INVALID_ARG;
        LBX_TRY(check_view(a));
        LBX_TRY(check_view(b));
        LBX_TRY(check_view(c));
        bool nonzero = false, big = false;
        for (int l = 0; l < nlimbs; ++l) {
            nonzero |= P[l] != 0;
            if (l > 0) big |= P[l] != 0;
        }
        if (!nonzero || (!big && P[0] < 2)) return LBX_ERR_INVALID_ARG;
        if (a->cols != b->rows || c->rows != a->rows || c->cols != b->cols) return LBX_ERR_SIZE_MISMATCH;
        if (!a->size || !b->size) return LBX_ERR_INVALID_ARG;
        const size_t m = a->rows, k = a->cols, n = b->cols;
        const size_t sc = max_rowdeg ? max_rowdeg + 1 : a->size + b->size - 1;
        if (c->size < sc) return LBX_ERR_SIZE_MISMATCH;
        size_t pts = 1;
        while (pts < sc) pts <<= 1;
        if (a->size > pts || b->size > pts) return LBX_ERR_SIZE_MISMATCH;  // the padded copies hold pts coefficients
        if (k > 0xffffffffu || a->size > 0xffffffffu || sc > 0xffffffffu) return LBX_ERR_UNSUPPORTED;

        std::lock_guard<std::mutex> lock(ctx->mu);
        DeviceGuard guard(ctx->device);
        hipStream_t s = (hipStream_t)stream;
        StreamOrder stream_order(ctx, s);
        const size_t wb = 4 * (size_t)nlimbs;
        Staged sa_, sb_, sc_;
        lbx_status st = stage_in(ctx, "stage_a", a, wb, true, &sa_, s);
        if (st == LBX_OK) st = stage_in(ctx, "stage_b", b, wb, true, &sb_, s);
        if (st == LBX_OK) st = stage_in(ctx, "stage_c", c, wb, !(c->shift == 0 && sc >= c->storage_size), &sc_, s);
        const bool host_io = sa_.owned || sb_.owned || sc_.owned;
        if (st == LBX_OK)
            st = mp_product(ctx, P, nlimbs, devview(a, sa_.dev), devview(b, sb_.dev), devview(c, sc_.dev), m, k, n,
                            method, s, nullptr, 0, sc);
        if (st == LBX_OK && sc_.owned) {
            st = device_to_host(ctx, c->data, sc_.dev, storage_words(c) * wb, s);
        }
        if (host_io) (void)hipStreamSynchronize(s);
        release(sa_);
        release(sb_);
        release(sc_);
        if (st == LBX_OK) c->size = sc;
        return st;
    });
}

extern "C" lbx_status lbx_polymat_mul_zz(lbx_ctx *ctx, lbx_polmat *c, int c_limbs, const lbx_polmat *a,
                                         const lbx_polmat *b, int ab_limbs, size_t max_rowdeg, int method,
                                         void *stream)
{
    return guarded([&]() -> lbx_status {
        if (!ctx || !a || !b || !c) return LBX_ERR_INVALID_ARG;
        if (ab_limbs < 1 || ab_limbs > kZzMaxIn || c_limbs < 1 || c_limbs > kZzMaxOut) return LBX_ERR_UNSUPPORTED;
        if (method < 0 || method > 2) return LBX_ERR_INVALID_ARG;
        LBX_TRY(check_view(a));
        LBX_TRY(check_view(b));
        LBX_TRY(check_view(c));
        if (a->cols != b->rows || c->rows != a->rows || c->cols != b->cols) return LBX_ERR_SIZE_MISMATCH;
        if (!a->size || !b->size) return LBX_ERR_INVALID_ARG;
        const size_t m = a->rows, k = a->cols, n = b->cols;
        const size_t sc = max_rowdeg ? max_rowdeg + 1 : a->size + b->size - 1;
        if (c->size < sc) return LBX_ERR_SIZE_MISMATCH;
        size_t pts = 1;
        while (pts < sc) pts <<= 1;
        if (a->size > pts || b->size > pts) return LBX_ERR_SIZE_MISMATCH;
        if (k > 0xffffffffu || a->size > 0xffffffffu || sc > 0xffffffffu) return LBX_ERR_UNSUPPORTED;

        std::lock_guard<std::mutex> lock(ctx->mu);
        DeviceGuard guard(ctx->device);
        hipStream_t s = (hipStream_t)stream;
        StreamOrder stream_order(ctx, s);
        const size_t wi = 4 * (size_t)ab_limbs, wo = 4 * (size_t)c_limbs;
        Staged sa_, sb_, sc_;
        lbx_status st = stage_in(ctx, "stage_a", a, wi, true, &sa_, s);
        if (st == LBX_OK) st = stage_in(ctx, "stage_b", b, wi, true, &sb_, s);
        if (st == LBX_OK) st = stage_in(ctx, "stage_c", c, wo, !(c->shift == 0 && sc >= c->storage_size), &sc_, s);
        const bool host_io = sa_.owned || sb_.owned || sc_.owned;
        if (st == LBX_OK)
            st = zz_product(ctx, ab_limbs, c_limbs, devview(a, sa_.dev), devview(b, sb_.dev), devview(c, sc_.dev), m, k,
                            n, sc, method, s);
        if (st == LBX_OK && sc_.owned) {
            st = device_to_host(ctx, c->data, sc_.dev, storage_words(c) * wo, s);
        }
        if (host_io) (void)hipStreamSynchronize(s);
        release(sa_);
        release(sb_);
        release(sc_);
        if (st == LBX_OK) c->size = sc;
        return st;
    });
}
