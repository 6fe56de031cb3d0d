// Order bases of matrix power series on the device.
//
//  * lbx_order_basis_m_u32 follows OrderBasis::M_Basis (order-basis.h:346-600):
//    one workgroup runs all `order` steps.  Each step sorts the shift with the
//    reference's selection-sort swaps (:388-399), forms the discrepancy
//    (coefficient k of sigma.serie, :416-444), splits its rows into the row
//    rank profile and the rest (the rank-profile PLUQ of :470-473), subtracts
//    from every other row its combination of the pivot rows (:509-517),
//    updates degrees and shifts (:546-566) and multiplies the pivot rows by x
//    (:573-580).  Rows of sigma never move in memory: a row map carries the
//    permutations, so a swap costs one word instead of a row of every
//    coefficient.
//  * lbx_order_basis_pm_u32 follows OrderBasis::PM_Basis (order-basis.h:219-326):
//    the host recursion over device buffers, M_Basis below order 32 (:47-48,232),
//    the serie update with the library's middle product (:274) and
//    sigma = sigma2.sigma1 with its polynomial-matrix product (:297).
//
// The discrepancy's rows are eliminated in increasing row order without
// divisions: a row r' below pivot row r becomes a.r' - b.r (a the pivot entry,
// b the entry of r' in the pivot column), and each row keeps its scale and its
// coefficients over the original pivot rows, so that a row that ends at zero
// gives its combination of pivot rows after one inverse of its scale.
#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "lbx_ctx.h"

// wave-wide 64-bit minimum and maximum of the device library (DPP reductions)
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_min_u64(unsigned long long);
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(unsigned long long);

namespace lbx {
namespace {

constexpr int kMbThreads = 1024;
constexpr int kMbWaves = kMbThreads / 64;
constexpr int kMbMaxM = 128;  // rows of the serie (= dimension of sigma)
constexpr int kMbMaxN = 64;   // columns of the serie
constexpr size_t kMbThreshold = 32;  // MBASIS_THRESHOLD (order-basis.h:47-48)
constexpr uint64_t kShiftLimit = (uint64_t)1 << 48;  // shift packed with a 7-bit row index

// x mod p for any 64-bit x and 2 <= p < 2^32, mu = floor((2^64-1)/p):
// q = hi64(x mu) undercounts x/p by less than 3, so r < 3p.
__device__ __forceinline__ uint32_t red64(uint64_t x, uint32_t p, uint64_t mu)
{
    const uint64_t q = __umul64hi(x, mu);
    uint64_t r = x - q * p;
    if (r >= p) r -= p;
    if (r >= p) r -= p;
    return (uint32_t)r;
}

__device__ __forceinline__ uint32_t mulm(uint32_t a, uint32_t b, uint32_t p, uint64_t mu)
{
    return red64((uint64_t)a * b, p, mu);
}

__device__ uint32_t invm(uint32_t a, uint32_t p, uint64_t mu)  // a^(p-2), a != 0
{
    uint32_t r = 1, b = a;
    for (uint32_t e = p - 2; e; e >>= 1) {
        if (e & 1) r = mulm(r, b, p, mu);
        b = mulm(b, b, p, mu);
    }
    return r;
}

// Montgomery product x y 2^-32 mod p for odd p < 2^31 (pinv = -p^-1 mod 2^32):
// x y + q p < 2^63, so the quotient by 2^32 is below 2p.
__device__ __forceinline__ uint32_t mont(uint32_t x, uint32_t y, uint32_t p, uint32_t pinv)
{
    const uint64_t t = (uint64_t)x * y;
    const uint32_t q = (uint32_t)t * pinv;
    return csub((uint32_t)((t + (uint64_t)q * p) >> 32), p);
}

// Sum of products kept as a 64-bit word plus a count of wraps of 2^64.
struct Acc {
    uint64_t lo = 0;
    uint32_t wraps = 0;
    __device__ __forceinline__ void add(uint32_t a, uint32_t b)
    {
        const uint64_t t = (uint64_t)a * b, s = lo + t;
        wraps += s < t ? 1u : 0u;
        lo = s;
    }
    // r64 = 2^64 mod p
    __device__ __forceinline__ uint32_t get(uint32_t p, uint64_t mu, uint32_t r64) const
    {
        return red64((uint64_t)red64(lo, p, mu) + (uint64_t)wraps * r64, p, mu);
    }
};

// Also clears the grid barrier's arrival count for the M_Basis launch that follows.
__global__ void __launch_bounds__(256) gather_coeff_major(DevView v, size_t count, uint32_t *out, unsigned *bar)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) bar[0] = 0;
    const size_t entries = v.rows * v.cols, total = entries * count;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t t = i / entries, e = i - t * entries;
        out[i] = v.base[view_index(v, e, t)];
    }
}

// Sum of products for p < 2^31: products are below 2^62, and the word is
// folded at 2^62 (2^62 = f62 mod p) after every second product, so it stays
// below 2^62 + 2^33 + 2^63.
struct AccSmall {
    uint64_t v = 0;
    __device__ __forceinline__ void add2(uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1, uint32_t f62)
    {
        v += (uint64_t)a0 * b0;
        v += (uint64_t)a1 * b1;
        v = (v & (((uint64_t)1 << 62) - 1)) + (uint64_t)(uint32_t)(v >> 62) * f62;
    }
    __device__ __forceinline__ void add1(uint32_t a, uint32_t b, uint32_t f62)
    {
        v += (uint64_t)a * b;
        v = (v & (((uint64_t)1 << 62) - 1)) + (uint64_t)(uint32_t)(v >> 62) * f62;
    }
    __device__ __forceinline__ uint32_t get(uint32_t p, uint64_t mu, uint32_t) const { return red64(v, p, mu); }
};

// Any p < 2^32: a 64-bit word and a count of its wraps.
struct AccWide {
    Acc a;
    __device__ __forceinline__ void add2(uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1, uint32_t)
    {
        a.add(a0, b0);
        a.add(a1, b1);
    }
    __device__ __forceinline__ void add1(uint32_t x, uint32_t y, uint32_t) { a.add(x, y); }
    __device__ __forceinline__ uint32_t get(uint32_t p, uint64_t mu, uint32_t r64) const { return a.get(p, mu, r64); }
};

// sum_j x[j] y[j*ystride], len terms
template <typename A>
__device__ __forceinline__ void dot(A &acc, const uint32_t *x, const uint32_t *y, int ystride, int len, uint32_t f62)
{
    int j = 0;
    for (; j + 7 < len; j += 8) {  // eight terms' loads issued together
        uint32_t xv[8], yv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            xv[u] = x[j + u];
            yv[u] = y[(j + u) * ystride];
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) acc.add2(xv[u], yv[u], xv[u + 1], yv[u + 1], f62);
    }
    for (; j + 1 < len; j += 2) acc.add2(x[j], y[j * ystride], x[j + 1], y[(j + 1) * ystride], f62);
    if (j < len) acc.add1(x[j], y[j * ystride], f62);
}

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) const void global_cvoid;

// Copies rows of `len` words from global memory into LDS with
// global_load_lds_dword (lane i of an instruction fills LDS word base + i):
// one instruction per 64 words of a row, one row per wave in turn, and no
// register staging, so every load of the copy can be in flight at once.  The
// caller waits (s_waitcnt 0) and synchronises before reading the copy.
template <typename Src, typename Dst>
__device__ __forceinline__ void stage_rows(int nrows, int len, Src src, Dst dst, int wave, int lane)
{
    for (int r = wave; r < nrows; r += kMbWaves)
        for (int j0 = 0; j0 < len; j0 += 64)
            if (j0 + lane < len)
                __builtin_amdgcn_global_load_lds((global_cvoid *)(src(r) + j0 + lane), (lds_void *)(dst(r) + j0), 4,
                                                 0, 0);
}

constexpr int kMbStageWords = 24576;  // LDS words for staged coefficient matrices
constexpr int kMbMaxOut = 4;          // m*n <= kMbMaxOut*kMbThreads
constexpr unsigned kNone = 0xffffffffu;

// One workgroup, kMbThreads threads.  S: serie, coefficient-major
// S[(t*m + j)*n + c]; W: order+1 coefficient matrices of m x m words,
// W[(l*m + phys)*m + j], phys a physical row.  Dynamic LDS: R, Z and D
// (m x n words each) and `stage` words for staged coefficient matrices.
//
// Per step k:
//  1. the shift sort (wave 0, keys in registers, one wave-wide minimum per position);
//  2. the discrepancy: rows that were pivots at step k-1 reuse that step's
//     discrepancy row (x.sigma_t has coefficient k+1 equal to coefficient k of
//     sigma_t, order-basis.h:446-456), the others are summed over staged blocks
//     of W[l] rows and S[k-l];
//  3. elimination, one barrier per pivot: the pivot column is never written, so
//     every thread reads its row's multiplier from it, and the next pivot is
//     searched for in the same pass;
//  4. non-pivot rows -= combination . pivot rows and pivot rows times x, on
//     staged blocks of W taken from the highest coefficient down, so the
//     shifted pivot rows never overwrite a coefficient still to be read.
// MONT (odd p < 2^31): the elimination runs on Montgomery forms x 2^32 mod p,
// whose zero pattern is the same and whose ratios Z/scale are the same.
template <typename A, bool MONT>
__global__ void __launch_bounds__(kMbThreads) m_basis_kernel(const uint32_t *__restrict__ S, uint32_t *__restrict__ W,
                                                             DevView out, uint64_t *__restrict__ shift_io,
                                                             uint64_t *__restrict__ deg_out, int order, int m, int n,
                                                             int stage, int out_count, uint32_t p, uint64_t mu,
                                                             uint32_t r64, uint32_t f62, uint32_t pinv, uint32_t r2m,
                                                             uint32_t one)
{
    // one = 1 in the elimination's representation (2^32 mod p with MONT)
    auto mul = [&](uint32_t x, uint32_t y) -> uint32_t {
        if constexpr (MONT) return mont(x, y, p, pinv);
        else return mulm(x, y, p, mu);
    };
    auto amb = [&](uint32_t a, uint32_t x, uint32_t b, uint32_t y) -> uint32_t {
        const uint32_t u = mul(a, x), v = mul(b, y);
        return u >= v ? u - v : u + (p - v);
    };
    extern __shared__ uint32_t dyn[];
    uint32_t *R = dyn;              // discrepancy rows under elimination, logical order
    uint32_t *Z = dyn + m * n;      // coefficients over the pivot rows, then the combination
    uint32_t *D = dyn + 2 * m * n;  // discrepancy before elimination, by physical row
    uint32_t *ST = dyn + 3 * m * n; // staged coefficient matrices
    __shared__ uint64_t shift[kMbMaxM], degree[kMbMaxM], tmp64[kMbMaxM], tmp64b[kMbMaxM];
    __shared__ int rowmap[kMbMaxM], tmpmap[kMbMaxM], pivot_of[kMbMaxM], oldrow[kMbMaxM], comp[kMbMaxM];
    __shared__ int was_piv[kMbMaxM];  // by physical row: pivot at the previous step
    __shared__ uint32_t scale[kMbMaxM];
    __shared__ unsigned smin[3];
    __shared__ int s_maxdeg, s_nc;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t mm = (size_t)m * m, nw = (size_t)(order + 1) * mm;
    const int mn = m * n;
    for (size_t i = tid; i < nw; i += kMbThreads) W[i] = 0;
    __syncthreads();
    for (int i = tid; i < m; i += kMbThreads) {
        W[(size_t)i * m + i] = 1;
        rowmap[i] = i;
        shift[i] = shift_io[i];
        degree[i] = 0;
        was_piv[i] = 0;
    }
    if (tid == 0) s_maxdeg = 0;
    __syncthreads();

    for (int k = 0; k < order; ++k) {
        const int maxdeg = s_maxdeg, lim = k < maxdeg ? k : maxdeg;
        // ---- 1. selection sort of the shift (order-basis.h:388-399) ----
        if (wave == 0) {
            uint64_t sh0 = lane < m ? shift[lane] : 0, sh1 = lane + 64 < m ? shift[lane + 64] : 0;
            uint64_t dg0 = lane < m ? degree[lane] : 0, dg1 = lane + 64 < m ? degree[lane + 64] : 0;
            int rm0 = lane < m ? rowmap[lane] : 0, rm1 = lane + 64 < m ? rowmap[lane + 64] : 0;
            // shifts spread over less than 2^24: 32-bit keys (shift - min) << 7 | position,
            // and only the positions' origins move; the arrays are gathered once at the end
            uint64_t lo = ~(uint64_t)0, hi = 0;
            if (lane < m) { lo = sh0; hi = sh0; }
            if (lane + 64 < m) { lo = sh1 < lo ? sh1 : lo; hi = sh1 > hi ? sh1 : hi; }
            for (int off = 32; off; off >>= 1) {
                const uint64_t a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
                lo = a < lo ? a : lo;
                hi = b > hi ? b : hi;
            }
            const bool narrow = hi - lo < ((uint64_t)1 << 24);
            if (narrow) {
                uint32_t v0 = (uint32_t)(sh0 - lo), v1 = (uint32_t)(sh1 - lo);
                int o0 = lane, o1 = lane + 64;
                for (int i = 0; i < m - 1; ++i) {
                    uint32_t best = 0xffffffffu;
                    if (lane < m && lane >= i) best = v0 << 7 | (uint32_t)lane;
                    if (lane + 64 < m && lane + 64 >= i) {
                        const uint32_t c1 = v1 << 7 | (uint32_t)(lane + 64);
                        best = c1 < best ? c1 : best;
                    }
                    best = __ockl_wfred_min_u32(best);  // wave-wide minimum (DPP), uniform
                    // every remaining key holds the largest shift: each later step would
                    // pick the element already in place, so the order is final
                    if ((best >> 7) == (uint32_t)(hi - lo)) break;
                    const int im = (int)(best & 127);
                    if (im != i) {
                        const int li = i & 63, lm = im & 63;
                        const bool hi_i = i >= 64, hi_m = im >= 64;
                        // uniform lanes: read them with v_readlane
                        const uint32_t vi = (uint32_t)__builtin_amdgcn_readlane((int)(hi_i ? v1 : v0), li);
                        const uint32_t vm = best >> 7;
                        const int oi = __builtin_amdgcn_readlane(hi_i ? o1 : o0, li);
                        const int om = __builtin_amdgcn_readlane(hi_m ? o1 : o0, lm);
                        if (lane == li) {
                            if (hi_i) { v1 = vm; o1 = om; } else { v0 = vm; o0 = om; }
                        }
                        if (lane == lm) {
                            if (hi_m) { v1 = vi; o1 = oi; } else { v0 = vi; o0 = oi; }
                        }
                    }
                }
                // gather by origin (all lanes read before any lane writes)
                if (lane < m) { sh0 = shift[o0]; dg0 = degree[o0]; rm0 = rowmap[o0]; }
                if (lane + 64 < m) { sh1 = shift[o1]; dg1 = degree[o1]; rm1 = rowmap[o1]; }
            }
            for (int i = 0; !narrow && i < m - 1; ++i) {
                uint64_t best = ~(uint64_t)0;  // key = shift << 7 | position: first minimum
                if (lane < m && lane >= i) best = sh0 << 7 | (uint64_t)lane;
                if (lane + 64 < m && lane + 64 >= i) {
                    const uint64_t c1 = sh1 << 7 | (uint64_t)(lane + 64);
                    best = c1 < best ? c1 : best;
                }
                for (int off = 32; off; off >>= 1) {
                    const uint64_t o = __shfl_xor(best, off);
                    best = o < best ? o : best;
                }
                if ((best >> 7) == hi) break;  // as above: the rest is in place
                const int im = (int)(best & 127);
                if (im != i) {
                    const int li = i & 63, lm = im & 63;
                    const bool hi_i = i >= 64, hi_m = im >= 64;
                    const uint64_t si = __shfl(hi_i ? sh1 : sh0, li), sm = __shfl(hi_m ? sh1 : sh0, lm);
                    const uint64_t di = __shfl(hi_i ? dg1 : dg0, li), dm = __shfl(hi_m ? dg1 : dg0, lm);
                    const int ri = __shfl(hi_i ? rm1 : rm0, li), rmm = __shfl(hi_m ? rm1 : rm0, lm);
                    if (lane == li) {
                        if (hi_i) { sh1 = sm; dg1 = dm; rm1 = rmm; } else { sh0 = sm; dg0 = dm; rm0 = rmm; }
                    }
                    if (lane == lm) {
                        if (hi_m) { sh1 = si; dg1 = di; rm1 = ri; } else { sh0 = si; dg0 = di; rm0 = ri; }
                    }
                }
            }
            if (lane < m) { shift[lane] = sh0; degree[lane] = dg0; rowmap[lane] = rm0; }
            if (lane + 64 < m) { shift[lane + 64] = sh1; degree[lane + 64] = dg1; rowmap[lane + 64] = rm1; }
            // rows to sum: those that were not pivots at the previous step
            int base = 0;
            for (int h = 0; h < m; h += 64) {
                const int r = h + lane;
                const bool need = r < m && !was_piv[h == 0 ? rm0 : rm1];
                const uint64_t bal = __ballot(need);
                if (need) comp[base + __popcll(bal & ((1ull << lane) - 1))] = r;
                base += __popcll(bal);
            }
            if (lane == 0) s_nc = base;
        }
        __syncthreads();
        // ---- 2. discrepancy: coefficient k of sigma.serie (order-basis.h:416-456) ----
        const int nc = s_nc;
        {
            A acc[kMbMaxOut];
            const int blk = nc * m + mn;  // staged words per coefficient
            const int lb = stage / blk > 0 ? stage / blk : 1;
            for (int l0 = 0; nc && l0 <= lim; l0 += lb) {
                const int l1 = l0 + lb < lim + 1 ? l0 + lb : lim + 1;
                stage_rows(
                    (l1 - l0) * nc, m,
                    [&](int pr) {
                        const int lq = pr / nc, i = pr - lq * nc;
                        return W + ((size_t)(l0 + lq) * m + rowmap[comp[i]]) * m;
                    },
                    [&](int pr) {
                        const int lq = pr / nc, i = pr - lq * nc;
                        return ST + lq * blk + i * m;
                    },
                    wave, lane);
                // serie blocks S[k-l], rows of n words
                stage_rows(
                    (l1 - l0) * m, n,
                    [&](int pr) {
                        const int lq = pr / m, j = pr - lq * m;
                        return S + (size_t)(k - l0 - lq) * mn + j * n;
                    },
                    [&](int pr) {
                        const int lq = pr / m, j = pr - lq * m;
                        return ST + lq * blk + nc * m + j * n;
                    },
                    wave, lane);
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
#pragma unroll
                for (int t = 0; t < kMbMaxOut; ++t) {
                    const int o = tid + t * kMbThreads;
                    if (o < nc * n) {
                        const int i = o / n, c = o - i * n;
                        for (int lq = 0; lq < l1 - l0; ++lq)
                            dot(acc[t], ST + lq * blk + i * m, ST + lq * blk + nc * m + c, n, m, f62);
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int t = 0; t < kMbMaxOut; ++t) {
                const int o = tid + t * kMbThreads;
                if (o < nc * n) {
                    const int i = o / n, c = o - i * n;
                    R[comp[i] * n + c] = acc[t].get(p, mu, r64);
                }
            }
        }
        for (int o = tid; o < mn; o += kMbThreads) {
            const int r = o / n, c = o - r * n, ph = rowmap[r];
            if (was_piv[ph]) R[o] = D[ph * n + c];
            Z[o] = 0;
        }
        for (int i = tid; i < m; i += kMbThreads) {
            scale[i] = one;
            pivot_of[i] = -1;
        }
        if (tid == 0) {
            smin[0] = kNone;
            smin[1] = kNone;
            smin[2] = kNone;
        }
        __syncthreads();
        {
            bool found = false;
            for (int r = wave; r < m; r += kMbWaves) {  // one row per wave, one lane per column
                const uint32_t v = lane < n ? R[r * n + lane] : 0u;
                if (lane < n) {
                    D[rowmap[r] * n + lane] = v;
                    if constexpr (MONT) R[r * n + lane] = mont(v, r2m, p, pinv);
                }
                const uint64_t bal = __ballot(v != 0);
                if (bal && !found) {
                    found = true;
                    if (lane == 0) atomicMin(&smin[0], (unsigned)(r << 8 | __builtin_ctzll(bal)));
                }
            }
        }
        __syncthreads();
        // ---- 3. row rank profile and combinations (the PLUQ of :470-488) ----
        int rank = 0;
        for (;;) {
            // a pivot position is encoded row << 8 | column (n <= 64): the same row-major
            // order as row * n + column, decoded without a division on the pivot's path
            const unsigned hit = smin[rank % 3];
            if (hit == kNone) break;
            const int s = (int)(hit >> 8), c0 = (int)(hit & 255u);
            const uint32_t a = R[s * n + c0], sc_s = scale[s];
            // Each row below s is the word vector [R_r | Z_r], updated as a.x - b.y
            // against [R_s | Z_s] with Z_s[rank] read as -scale_s (so that
            // Z_r[rank] = b scale_s).  Columns of earlier pivots are zero in both
            // rows and column c0 becomes a b - b a = 0; b is read by the whole wave
            // before its lanes store, and no other wave touches row r.
            const uint32_t neg_sc = p - sc_s;
            const int wdt = n + rank + 1;
            unsigned *nxt = &smin[(rank + 1) % 3];
            // Every row is updated, b = 0 included (the row and its scale are then both
            // multiplied by a, which keeps Z/scale and the zero pattern): no branch per
            // row, and two rows per pass for independent chains.  Lane j holds word j
            // and word j+64 of the pivot vector (n <= 64, so word j+64 is a combination word).
            const int j1 = lane + 64;
            const bool has0 = lane < wdt, has1 = j1 < wdt, isr0 = lane < n;
            const int zo0 = lane - n, zo1 = j1 - n;
            const uint32_t y0 = !has0 ? 0u : (isr0 ? R[s * n + lane] : (zo0 == rank ? neg_sc : Z[s * n + zo0]));
            const uint32_t y1 = !has1 ? 0u : (zo1 == rank ? neg_sc : Z[s * n + zo1]);
            auto update = [&](int r, uint32_t b) -> bool {  // returns: lane's column nonzero
                bool nz = false;
                if (has0) {
                    uint32_t *xp = isr0 ? R + r * n + lane : Z + r * n + zo0;
                    const uint32_t v = amb(a, *xp, b, y0);
                    *xp = v;
                    nz = isr0 && v != 0;
                }
                if (has1) {
                    uint32_t *xp = Z + r * n + zo1;
                    *xp = amb(a, *xp, b, y1);
                }
                return nz;
            };
            bool found = false;  // this wave's rows come in increasing order: only its first hit counts
            for (int r = s + 1 + wave; r < m; r += 2 * kMbWaves) {
                const int r2 = r + kMbWaves;
                const bool two = r2 < m;
                const uint32_t b = R[r * n + c0], b2 = two ? R[r2 * n + c0] : 0u;
                const bool nz = update(r, b);
                const bool nz2 = two && update(r2, b2);
                if (lane == 0) {
                    scale[r] = mul(a, scale[r]);
                    if (two) scale[r2] = mul(a, scale[r2]);
                }
                const uint64_t bal = __ballot(nz), bal2 = __ballot(nz2);
                if (!found && (bal || bal2)) {
                    found = true;
                    const unsigned hit = bal ? (unsigned)(r << 8 | __builtin_ctzll(bal))
                                             : (unsigned)(r2 << 8 | __builtin_ctzll(bal2));
                    if (lane == 0) atomicMin(nxt, hit);
                }
            }
            if (tid == 0) {
                pivot_of[s] = rank;
                smin[(rank + 2) % 3] = kNone;
            }
            ++rank;
            __syncthreads();
        }
        // non-pivot rows: combination = Z / scale (one inverse per row, one lane each)
        for (int r = tid; r < m; r += kMbThreads)
            if (pivot_of[r] < 0 && scale[r] != 1) scale[r] = invm(scale[r], p, mu);
        __syncthreads();
        for (int r = wave; r < m; r += kMbWaves) {
            if (pivot_of[r] >= 0 || scale[r] == 1) continue;
            const uint32_t inv = scale[r];
            if (lane < rank) Z[r * n + lane] = mulm(Z[r * n + lane], inv, p, mu);  // plain product: Z~/scale~ = Z/scale
        }
        // Qt: pivot rows first, then the others, each in increasing order; a
        // non-pivot row's place counts the non-pivot rows above it (m <= 128:
        // two waves, one ballot per 64 rows)
        static_assert(kMbMaxM <= 2 * 64 && kMbThreads >= 2 * 64, "two ballots cover every row");
        if (tid < 2 * 64) {
            const uint64_t b0 = __ballot(lane < m && pivot_of[lane] < 0);
            const uint64_t b1 = __ballot(lane + 64 < m && pivot_of[lane + 64] < 0);
            const int r = tid;
            if (r < m) {
                int pos = pivot_of[r];
                if (pos < 0) {
                    const uint64_t below = (1ull << lane) - 1;
                    pos = rank + (wave == 0 ? __popcll(b0 & below) : __popcll(b0) + __popcll(b1 & below));
                }
                oldrow[pos] = r;
                tmpmap[pos] = rowmap[r];
                tmp64[pos] = shift[r];
                tmp64b[pos] = degree[r];
                was_piv[rowmap[r]] = pivot_of[r] >= 0;
            }
        }
        __syncthreads();
        for (int i = tid; i < m; i += kMbThreads) {
            rowmap[i] = tmpmap[i];
            shift[i] = tmp64[i];
            degree[i] = tmp64b[i];
        }
        __syncthreads();
        // ---- 4. non-pivot rows -= combination . pivot rows (:509-517), pivot rows times x (:573-580) ----
        {
            const int lb = stage / (int)mm > 0 ? stage / (int)mm : 1;
            for (int l1 = lim + 1; l1 > 0;) {
                const int l0 = l1 - lb > 0 ? l1 - lb : 0;
                stage_rows(
                    (l1 - l0) * m, m,
                    [&](int pr) {
                        const int lq = pr / m, i = pr - lq * m;
                        return W + ((size_t)(l0 + lq) * m + rowmap[i]) * m;
                    },
                    [&](int pr) { return ST + pr * m; }, wave, lane);
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                const int nrows = m - rank;
                for (int pr = wave; pr < (l1 - l0) * nrows; pr += kMbWaves) {
                    const int lq = pr / nrows, si = rank + pr - lq * nrows;
                    const uint32_t *st = ST + lq * (int)mm, *y = Z + oldrow[si] * n;
                    uint32_t *dst = W + ((size_t)(l0 + lq) * m + rowmap[si]) * m;
                    for (int j = lane; j < m; j += 64) {
                        A acc;
                        dot(acc, y, st + j, m, rank, f62);
                        const uint32_t sub = acc.get(p, mu, r64), v = st[si * m + j];
                        dst[j] = v >= sub ? v - sub : v + (p - sub);
                    }
                }
                for (int pr = wave; pr < (l1 - l0) * rank; pr += kMbWaves) {
                    const int lq = pr / rank, t = pr - lq * rank;
                    const uint32_t *src = ST + lq * (int)mm + t * m;
                    uint32_t *dst = W + ((size_t)(l0 + lq + 1) * m + rowmap[t]) * m;
                    for (int j = lane; j < m; j += 64) dst[j] = src[j];
                }
                __syncthreads();
                l1 = l0;
            }
            for (int o = tid; o < rank * m; o += kMbThreads) W[(size_t)rowmap[o / m] * m + o % m] = 0;
        }
        // ---- degrees and shifts (:546-566) ----
        if (tid == 0) {
            uint64_t dmax = 0;
            for (int i = 0; i < rank; ++i) dmax = degree[i] > dmax ? degree[i] : dmax;
            for (int i = 0; i < rank; ++i) {
                degree[i]++;
                shift[i]++;
            }
            for (int i = rank; i < m; ++i) degree[i] = degree[i] < dmax ? dmax : degree[i];
            const uint64_t md = (uint64_t)maxdeg > dmax + 1 ? (uint64_t)maxdeg : dmax + 1;
            s_maxdeg = (int)(md < (uint64_t)order ? md : (uint64_t)order);
        }
        __syncthreads();
    }
    // sigma.resize(max_degree+1), shift2 = shift (:592-597); out_count coefficients
    // written when nonzero, zero above the degree
    const int md = s_maxdeg;
    const size_t cnt = out_count ? (size_t)out_count : (size_t)md + 1;
    for (size_t o = tid; o < cnt * mm; o += kMbThreads) {
        const size_t l = o / mm, e = o - l * mm, i = e / m, j = e - i * m;
        out.base[view_index(out, e, l)] = l <= (size_t)md ? W[(l * m + rowmap[i]) * m + j] : 0u;
    }
    for (int i = tid; i < m; i += kMbThreads) shift_io[i] = shift[i];
    if (tid == 0) *deg_out = (uint64_t)md;
}

// ---------------------------------------------------------------------------
// M_Basis over several workgroups (orders up to kMcMaxOrder: the PM_Basis
// base case).  Row operations on sigma act on every column alike, so sigma's
// columns are split over G workgroups: group g keeps columns [g cw, g cw + cw)
// of every coefficient of sigma, and rows g cw .. of the serie that those
// columns meet, resident in LDS for the whole call.  Per step each group
// forms its share of the discrepancy (the sum over its columns), the shares
// are exchanged through global memory with one grid barrier, and every group
// then runs the same sort, elimination and bookkeeping on the summed
// discrepancy (so every group takes the same decisions) and applies the row
// operations to its own columns.  The steps follow m_basis_kernel and the
// reference's order-basis.h:346-600; the differences are in how they are
// computed:
//  * the shift sort handles one value at a time: the reference's selection
//    sort places the elements of the smallest remaining value in order of
//    position, and each element it displaces follows a chain of those
//    positions, which one pass of the wave resolves;
//  * each wave keeps its rows of [R | Z] in registers during elimination and
//    stores a row only when it is that wave's candidate for the next pivot;
//  * a x - b y takes one Montgomery reduction (a x + b (p - y) < 2^64 for
//    p < 2^31), and rows carry no scale: every row below pivot j is
//    multiplied by its entry a_j, so a row's scale is the product of the a_j
//    of the pivots above it, read from a prefix product at the end.
constexpr int kMcMaxOrder = 32;
constexpr int kMcMaxGroups = 64;
constexpr int kMcLdsWords = 36864;  // dynamic LDS words (the static arrays take the rest of the 160 KiB)
constexpr unsigned kMcSpinLimit = 1u << 24;  // polls of the grid barrier before it reports a failure

// a x - b y mod p on Montgomery forms, ny = p - y (odd p < 2^31)
__device__ __forceinline__ uint32_t amb1(uint32_t a, uint32_t x, uint32_t b, uint32_t ny, uint32_t p, uint32_t pinv)
{
    const uint64_t t = (uint64_t)a * x + (uint64_t)b * ny;
    const uint32_t q = (uint32_t)t * pinv;
    const uint32_t r = (uint32_t)((t + (uint64_t)q * p) >> 32);
    return r >= p ? r - p : r;
}

// All groups of the grid meet here for the `count`-th time.  bar[0] counts
// arrivals, bar[1] is set when a poll limit runs out (the host reports the
// call as failed, and later barriers no longer wait).  The shares exchanged
// through the barrier are written and read with agent-scope atomics.
__device__ __forceinline__ void grid_barrier(unsigned *bar, unsigned count)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned target = count * gridDim.x;
        unsigned polls = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            // the failure flag only matters once a limit ran out: read it now and then
            if ((polls & 255) == 255 && __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (++polls == kMcSpinLimit) {
                __hip_atomic_store(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
    }
    __syncthreads();
}

// sum of x[j] y[j*ystride], len terms, on two independent accumulators
__device__ __forceinline__ void dot2(AccSmall &a0, AccSmall &a1, const uint32_t *x, const uint32_t *y, int ystride,
                                     int len, uint32_t f62)
{
    int j = 0;
    for (; j + 3 < len; j += 4) {
        const uint32_t x0 = x[j], x1 = x[j + 1], x2 = x[j + 2], x3 = x[j + 3];
        const uint32_t y0 = y[j * ystride], y1 = y[(j + 1) * ystride], y2 = y[(j + 2) * ystride],
                       y3 = y[(j + 3) * ystride];
        a0.add2(x0, y0, x1, y1, f62);
        a1.add2(x2, y2, x3, y3, f62);
    }
    for (; j < len; ++j) a0.add1(x[j], y[j * ystride], f62);
}

__device__ __forceinline__ uint32_t sum2(const AccSmall &a0, const AccSmall &a1, uint32_t p, uint64_t mu)
{
    return csub(a0.get(p, mu, 0) + a1.get(p, mu, 0), p);
}

// G = gridDim.x groups of kMbThreads threads; ROWS >= m / kMbWaves rows of the
// elimination per wave.  S: serie, coefficient-major S[(t*m + j)*n + c]; part:
// 2 x G x m x n words of exchanged discrepancy shares; bar: the grid barrier's
// two words, zero at launch.  Dynamic LDS, in words: W (order+1) x m x cw
// (W[(l*m + phys)*cw + jj], coefficient l of sigma's column g*cw+jj in
// physical row phys), Sg order x cw x n (serie rows of the group's columns),
// V m x (2n+1) (row r of the discrepancy under elimination, [R_r | Z_r],
// logical order), D m x n (discrepancy before elimination, by physical row).
// out_count: coefficients of sigma written (zero above the degree); 0 writes
// degree+1 of them.
template <int ROWS>
__global__ void __launch_bounds__(kMbThreads) m_basis_cols_kernel(
    const uint32_t *__restrict__ S, DevView out, uint64_t *__restrict__ shift_io, uint64_t *__restrict__ deg_out,
    uint32_t *__restrict__ part, unsigned *__restrict__ bar, int order, int m, int n, int cw, int out_count,
    uint32_t p, uint64_t mu, uint32_t f62, uint32_t pinv, uint32_t r2m, uint32_t one)
{
    const int G = gridDim.x, g = blockIdx.x, j0 = g * cw;
    const int cols = m - j0 < cw ? m - j0 : cw;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int mn = m * n, vs = 2 * n + 1, cst = m * cw;  // cst: words per coefficient of the group's W
    extern __shared__ uint32_t dyn[];
    uint32_t *W = dyn;
    uint32_t *Sg = W + (order + 1) * cst;
    uint32_t *V = Sg + order * cw * n;
    uint32_t *D = V + m * vs;
    __shared__ uint64_t shift[kMbMaxM], degree[kMbMaxM], tmp64[kMbMaxM], tmp64b[kMbMaxM];
    __shared__ int rowmap[kMbMaxM], tmpmap[kMbMaxM], oldrow[kMbMaxM], pivot_of[kMbMaxM], comp[kMbMaxM],
        cidx[kMbMaxM], was_piv[kMbMaxM], qtab[kMbMaxM], piv_s[kMbMaxN], piv_c[kMbMaxN];
    __shared__ uint32_t pref[kMbMaxN + 1], rinv[kMbMaxM], piv_a[kMbMaxN];  // pivots: row, column, entry
    __shared__ int s_maxdeg, s_nc, s_npub, s_turn;

    for (int i = tid; i < (order + 1) * cst; i += kMbThreads) {
        const int q = i / cw, jj = i - q * cw, ph = q % m, l = q / m;
        W[i] = (l == 0 && ph == j0 + jj) ? 1u : 0u;
    }
    for (int i = tid; i < order * cw * n; i += kMbThreads) {
        const int q = i / n, c = i - q * n, jj = q % cw, t = q / cw;
        Sg[i] = jj < cols ? S[((size_t)t * m + j0 + jj) * n + c] : 0u;
    }
    for (int i = tid; i < m; i += kMbThreads) {
        rowmap[i] = i;
        shift[i] = shift_io[i];
        degree[i] = 0;
        was_piv[i] = 0;
    }
    if (tid == 0) {
        s_maxdeg = 0;
        pref[0] = one;
    }
    __syncthreads();
    unsigned nbar = 0;  // grid barriers passed (the same count in every group)
    const uint64_t below = (1ull << lane) - 1;

    for (int k = 0; k < order; ++k) {
        const int maxdeg = s_maxdeg, lim = k < maxdeg ? k : maxdeg;
        // ---- 1. selection sort of the shift (order-basis.h:388-399), one value at a time ----
        if (wave == 0) {
            const bool v0 = lane < m, v1 = lane + 64 < m;
            uint64_t k0 = v0 ? shift[lane] : 0, k1 = v1 ? shift[lane + 64] : 0;
            int o0 = lane, o1 = lane + 64;
            for (int start = 0; start < m - 1;) {
                const bool a0 = v0 && lane >= start, a1 = v1 && lane + 64 >= start;
                uint64_t lo = ~(uint64_t)0, hi = 0;
                if (a0) { lo = k0; hi = k0; }
                if (a1) { lo = k1 < lo ? k1 : lo; hi = k1 > hi ? k1 : hi; }
                lo = __ockl_wfred_min_u64(lo);
                hi = __ockl_wfred_max_u64(hi);
                if (lo == hi) break;  // what remains holds one value: every later swap is the identity
                // the smallest value's positions q_0 < q_1 < ... (c of them): step start+j
                // swaps positions start+j and q_j, since the value's elements not yet placed
                // never move and q_j >= start+j
                const bool m0 = a0 && k0 == lo, m1 = a1 && k1 == lo;
                const uint64_t b0 = __ballot(m0), b1 = __ballot(m1);
                const int n0 = __popcll(b0), end = start + n0 + __popcll(b1);
                const int r0 = __popcll(b0 & below), r1 = n0 + __popcll(b1 & below);
                if (m0) qtab[r0] = lane;
                if (m1) qtab[r1] = lane + 64;
                __builtin_amdgcn_s_waitcnt(0);
                __builtin_amdgcn_wave_barrier();
                // destination: start+rank for the value's elements; an element displaced
                // from x in [start, end) goes to q_{x-start}, and on while that is below end
                auto dest = [&](int x, bool isv, int rk) -> int {
                    if (isv) return start + rk;
                    if (x < start || x >= end) return x;
                    int y = qtab[x - start];
                    while (y < end) y = qtab[y - start];
                    return y;
                };
                const int d0 = v0 ? dest(lane, m0, r0) : lane;
                const int d1 = v1 ? dest(lane + 64, m1, r1) : lane + 64;
                __builtin_amdgcn_s_waitcnt(0);
                __builtin_amdgcn_wave_barrier();
                if (v0) { tmp64[d0] = k0; tmpmap[d0] = o0; }
                if (v1) { tmp64[d1] = k1; tmpmap[d1] = o1; }
                __builtin_amdgcn_s_waitcnt(0);
                __builtin_amdgcn_wave_barrier();
                if (v0) { k0 = tmp64[lane]; o0 = tmpmap[lane]; }
                if (v1) { k1 = tmp64[lane + 64]; o1 = tmpmap[lane + 64]; }
                __builtin_amdgcn_s_waitcnt(0);
                __builtin_amdgcn_wave_barrier();
                start = end;
            }
            // gather by origin (all lanes read before any lane writes)
            uint64_t dg0 = 0, dg1 = 0;
            int rm0 = 0, rm1 = 0;
            if (v0) { dg0 = degree[o0]; rm0 = rowmap[o0]; }
            if (v1) { dg1 = degree[o1]; rm1 = rowmap[o1]; }
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
            if (v0) { shift[lane] = k0; degree[lane] = dg0; rowmap[lane] = rm0; }
            if (v1) { shift[lane + 64] = k1; degree[lane + 64] = dg1; rowmap[lane + 64] = rm1; }
            // rows to sum: those that were not pivots at the previous step
            int base = 0;
            for (int h = 0; h < m; h += 64) {
                const int r = h + lane;
                const bool need = r < m && !was_piv[h == 0 ? rm0 : rm1];
                const uint64_t bal = __ballot(need);
                const int idx = base + __popcll(bal & below);
                if (r < m) cidx[r] = need ? idx : -1;
                if (need) comp[idx] = r;
                base += __popcll(bal);
            }
            if (lane == 0) s_nc = base;
        }
        __syncthreads();
        // ---- 2. this group's share of the discrepancy (order-basis.h:416-456) ----
        const int nc = s_nc, slot = (int)(nbar & 1);
        for (int o = tid; o < nc * n; o += kMbThreads) {
            const int i = o / n, c = o - i * n, ph = rowmap[comp[i]];
            AccSmall acc0, acc1;
            for (int l = 0; l <= lim; ++l)
                dot2(acc0, acc1, W + l * cst + ph * cw, Sg + (k - l) * cw * n + c, n, cols, f62);
            const uint32_t v = sum2(acc0, acc1, p, mu);
            if (G == 1) V[comp[i] * vs + c] = v;
            else __hip_atomic_store(part + (slot * G + g) * mn + o, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (G > 1 && nc) grid_barrier(bar, ++nbar);
        else __syncthreads();
        // the summed discrepancy, its Montgomery form in V, an empty Z part; rows
        // that were pivots at the previous step keep theirs: x.sigma_t has the same
        // discrepancy (:446-456).  One row per wave, one column per lane.
        for (int r = wave; r < m; r += kMbWaves) {
            const int ph = rowmap[r];
            const bool wp = was_piv[ph] != 0;
            if (lane < n) {
                uint32_t v;
                if (wp) {
                    v = D[ph * n + lane];
                } else if (G == 1) {
                    v = V[r * vs + lane];
                } else {
                    // plain loads, eight in flight: the grid barrier's acquire fence has
                    // invalidated this CU's caches
                    const uint32_t *q = part + slot * G * mn + cidx[r] * n + lane;
                    uint64_t sum = 0;
                    int h = 0;
                    for (; h + 16 <= G; h += 16) {
                        uint32_t t[16];
#pragma unroll
                        for (int j = 0; j < 16; ++j) t[j] = q[(h + j) * mn];
#pragma unroll
                        for (int j = 0; j < 16; ++j) sum += t[j];
                    }
                    for (; h + 8 <= G; h += 8) {
                        uint32_t t[8];
#pragma unroll
                        for (int j = 0; j < 8; ++j) t[j] = q[(h + j) * mn];
#pragma unroll
                        for (int j = 0; j < 8; ++j) sum += t[j];
                    }
                    for (; h < G; ++h) sum += q[h * mn];
                    v = red64(sum, p, mu);
                }
                D[ph * n + lane] = v;
                V[r * vs + lane] = mont(v, r2m, p, pinv);
                V[r * vs + n + lane] = 0;
            }
        }
        for (int i = tid; i < m; i += kMbThreads) pivot_of[i] = -1;
        if (tid == 0) {
            s_npub = 0;
            s_turn = 0;
        }
        __syncthreads();
        // ---- 3. row rank profile and combinations (the PLUQ of :470-488) ----
        // Wave w holds rows r0 = w*ROWS .. r0+ROWS-1 of [R | Z] in registers (lane j: words
        // j and j+64).  Waves take turns in order: on its turn wave w runs through its own
        // rows, each against the pivots before it, publishes each pivot's vector as it finds
        // it (s_npub) and then hands the turn on (s_turn).  Until its turn a wave applies to
        // its rows every pivot already published, so the waves below apply a block's pivots
        // while it is still being eliminated.  Every row below a pivot meets it once, in pivot
        // order, as in a pivot-by-pivot elimination.  Pivot vector j: [R_s | Z_s (j words) |
        // -scale_s | 0 ...]; a row below it becomes a [R_r | Z_r] - b [pivot vector],
        // a = R_s[c_j], b = R_r[c_j].  The counters are workgroup-scope release/acquire
        // words in LDS; a wait that runs past its poll limit raises the failure flag, and
        // every later wait in the call then gives up at once.
        const int r0 = wave * ROWS;
        uint32_t x0[ROWS], x1[ROWS];
#pragma unroll
        for (int u = 0; u < ROWS; ++u) {
            const int r = r0 + u;
            x0[u] = r < m && lane < n ? V[r * vs + lane] : 0u;
            x1[u] = 0u;
        }
        // applies pivots [applied, upto) to this wave's rows
        int applied = 0;
        auto apply_to = [&](int upto) {
            for (; applied < upto; ++applied) {
                const int j = applied, s = piv_s[j], c0 = piv_c[j];
                const uint32_t a = piv_a[j];
                const int wdt = n + j + 1;
                const uint32_t y0 = lane < wdt ? V[s * vs + lane] : 0u, ny0 = p - y0;
                // the rows' updates are independent: straight-line code lets their
                // multiply chains overlap
                if (wdt > 64) {
                    const uint32_t y1 = lane + 64 < wdt ? V[s * vs + lane + 64] : 0u, ny1 = p - y1;
#pragma unroll
                    for (int u = 0; u < ROWS; ++u) {
                        const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)x0[u], c0);
                        x0[u] = amb1(a, x0[u], b, ny0, p, pinv);
                        x1[u] = amb1(a, x1[u], b, ny1, p, pinv);
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < ROWS; ++u) {
                        const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)x0[u], c0);
                        x0[u] = amb1(a, x0[u], b, ny0, p, pinv);
                    }
                }
            }
        };
        if (r0 < m) {
            unsigned polls = 0;
            bool mine = false;
            for (;;) {
                // the turn first: once it is this wave's, every pivot of the waves
                // before it is published and counted below
                const int turn = __hip_atomic_load(&s_turn, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                const int pub = __hip_atomic_load(&s_npub, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                apply_to(pub);
                if (pub == n) break;  // every pivot found: these rows hold no new one
                if (turn >= wave) {
                    mine = true;
                    break;
                }
                // the failure flag (global memory) only matters once a poll limit
                // ran out somewhere: read it now and then, not on every poll
                if ((polls & 255) == 255 && __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
                if (++polls == kMcSpinLimit) {
                    __hip_atomic_store(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (mine) {
                int j = applied;
                uint32_t aprod = pref[j];  // the scale of the next pivot row
#pragma unroll
                for (int u = 0; u < ROWS; ++u) {
                    const int r = r0 + u;
                    const uint64_t bal = __ballot(r < m && lane < n && x0[u] != 0);
                    if (bal) {
                        const int c0 = (int)__builtin_ctzll(bal), wdt = n + j + 1;
                        const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)x0[u], c0);
                        const uint32_t y0 = lane == wdt - 1 ? p - aprod : x0[u];
                        const uint32_t y1 = lane + 64 == wdt - 1 ? p - aprod : x1[u];
                        if (lane < wdt) V[r * vs + lane] = y0;
                        if (lane + 64 < wdt) V[r * vs + lane + 64] = y1;
                        const uint32_t next = mont(aprod, a, p, pinv);
                        if (lane == 0) {
                            piv_s[j] = r;
                            piv_c[j] = c0;
                            piv_a[j] = a;
                            pivot_of[r] = j;
                            pref[j + 1] = next;
                            __hip_atomic_store(&s_npub, j + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                        const uint32_t ny0 = p - y0, ny1 = p - y1;
#pragma unroll
                        for (int v = u + 1; v < ROWS; ++v) {
                            const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)x0[v], c0);
                            x0[v] = amb1(a, x0[v], b, ny0, p, pinv);
                            if (wdt > 64) x1[v] = amb1(a, x1[v], b, ny1, p, pinv);
                        }
                        aprod = next;
                        ++j;
                    }
                }
                applied = j;
                if (lane == 0) __hip_atomic_store(&s_turn, wave + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        __syncthreads();
        const int rank = s_npub;
        // a non-pivot row's scale is the product of the entries of the pivots above it;
        // its combination is Z / scale: plain inverse s^-1 = mont(montpow(scale, p-2), 1)
        if (tid < 2 * 64) {
            const uint64_t pm0 = __ballot(lane < m && pivot_of[lane] >= 0);
            const uint64_t pm1 = __ballot(lane + 64 < m && pivot_of[lane + 64] >= 0);
            const int r = tid;
            if (r < m && pivot_of[r] < 0) {
                const int cnt = wave == 0 ? __popcll(pm0 & below) : __popcll(pm0) + __popcll(pm1 & below);
                uint32_t x = one, bp = pref[cnt];
                for (uint32_t e = p - 2; e; e >>= 1) {
                    if (e & 1) x = mont(x, bp, p, pinv);
                    bp = mont(bp, bp, p, pinv);
                }
                rinv[r] = mont(x, 1u, p, pinv);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < ROWS; ++u) {
            const int r = r0 + u;
            if (r < m && pivot_of[r] < 0) {
                const uint32_t inv = rinv[r];
                if (lane >= n && lane < n + rank) V[r * vs + lane] = mont(x0[u], inv, p, pinv);
                if (lane + 64 < n + rank) V[r * vs + lane + 64] = mont(x1[u], inv, p, pinv);
            }
        }
        // Qt: pivot rows first, then the others, each in increasing order
        static_assert(kMbMaxM <= 2 * 64 && kMbThreads >= 2 * 64, "two ballots cover every row");
        if (tid < 2 * 64) {
            const uint64_t b0 = __ballot(lane < m && pivot_of[lane] < 0);
            const uint64_t b1 = __ballot(lane + 64 < m && pivot_of[lane + 64] < 0);
            const int r = tid;
            if (r < m) {
                int pos = pivot_of[r];
                if (pos < 0) pos = rank + (wave == 0 ? __popcll(b0 & below) : __popcll(b0) + __popcll(b1 & below));
                oldrow[pos] = r;
                tmpmap[pos] = rowmap[r];
                tmp64[pos] = shift[r];
                tmp64b[pos] = degree[r];
                was_piv[rowmap[r]] = pivot_of[r] >= 0;
            }
        }
        __syncthreads();
        for (int i = tid; i < m; i += kMbThreads) {
            rowmap[i] = tmpmap[i];
            shift[i] = tmp64[i];
            degree[i] = tmp64b[i];
        }
        __syncthreads();
        // ---- 4. this group's columns: non-pivot rows -= combination . pivot rows (:509-517) ----
        {
            const int nrows = m - rank, per_l = nrows * cw, total = (lim + 1) * per_l;
            for (int o = tid; o < total; o += kMbThreads) {
                const int l = o / per_l, rem = o - l * per_l, q = rem / cw, jj = rem - q * cw, si = rank + q;
                const uint32_t *y = V + oldrow[si] * vs + n;
                const uint32_t *w = W + l * cst + jj;
                AccSmall acc0, acc1;
                int t = 0;
                for (; t + 3 < rank; t += 4) {
                    const uint32_t w0 = w[rowmap[t] * cw], w1 = w[rowmap[t + 1] * cw], w2 = w[rowmap[t + 2] * cw],
                                   w3 = w[rowmap[t + 3] * cw];
                    acc0.add2(y[t], w0, y[t + 1], w1, f62);
                    acc1.add2(y[t + 2], w2, y[t + 3], w3, f62);
                }
                for (; t < rank; ++t) acc0.add1(y[t], w[rowmap[t] * cw], f62);
                const uint32_t sub = sum2(acc0, acc1, p, mu);
                uint32_t *dst = W + l * cst + rowmap[si] * cw + jj;
                const uint32_t v = *dst;
                *dst = v >= sub ? v - sub : v + (p - sub);
            }
        }
        __syncthreads();
        // ---- pivot rows times x (:573-580); degrees and shifts (:546-566) ----
        for (int o = tid; o < rank * cw; o += kMbThreads) {
            const int t = o / cw, jj = o - t * cw;
            uint32_t *col = W + rowmap[t] * cw + jj;
            // eight coefficients at a time from the top: a block's writes reach only
            // coefficients already read
            for (int top = lim; top >= 0; top -= 8) {
                uint32_t v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (top - u >= 0) v[u] = col[(top - u) * cst];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (top - u >= 0) col[(top - u + 1) * cst] = v[u];
            }
            col[0] = 0;
        }
        if (wave == kMbWaves - 1) {
            uint64_t dm = 0;
            for (int i = lane; i < rank; i += 64) dm = degree[i] > dm ? degree[i] : dm;
            dm = __ockl_wfred_max_u64(dm);
            for (int i = lane; i < m; i += 64) {
                if (i < rank) {
                    degree[i]++;
                    shift[i]++;
                } else if (degree[i] < dm) {
                    degree[i] = dm;
                }
            }
            if (lane == 0) {
                const uint64_t md = (uint64_t)maxdeg > dm + 1 ? (uint64_t)maxdeg : dm + 1;
                s_maxdeg = (int)(md < (uint64_t)order ? md : (uint64_t)order);
            }
        }
        __syncthreads();
    }
    // sigma.resize(max_degree+1), shift2 = shift (:592-597)
    const int md = s_maxdeg;
    const int cnt = out_count ? out_count : md + 1;
    for (int o = tid; o < cnt * m * cols; o += kMbThreads) {
        const int q = o / cols, l = q / m, jj = o - q * cols, i = q - l * m;
        const uint32_t v = l <= md ? W[l * cst + rowmap[i] * cw + jj] : 0u;
        out.base[view_index(out, (size_t)i * m + j0 + jj, l)] = v;
    }
    if (g == 0) {
        for (int i = tid; i < m; i += kMbThreads) shift_io[i] = shift[i];
        if (tid == 0) *deg_out = (uint64_t)md;
    }
}

unsigned grid_for(size_t total)
{
    const size_t g = (total + 255) / 256;
    return (unsigned)(g < 4096 ? (g ? g : 1) : 4096);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device bytes of the buffers each recursion level keeps live (sigma1, serie2,
// sigma2 and the two degree words), summed down the deepest path.
size_t pm_level_bytes(size_t m, size_t n, size_t order)
{
    if (order <= kMbThreshold) return 0;
    const size_t o1 = order >> 1, o2 = order - o1;
    const size_t own =
        align256(m * m * (o1 + 1) * 4) + align256(m * n * o2 * 4) + align256(m * m * (o2 + 1) * 4) + align256(16);
    return own + std::max(pm_level_bytes(m, n, o1), pm_level_bytes(m, n, o2));
}

struct MbScratch {
    uint32_t *S, *W, *part;
    uint64_t *shift, *deg;
    unsigned *bar;  // grid barrier: arrivals, then the failure flag
};

struct PmRun {
    lbx_ctx *ctx;
    uint32_t p;
    size_t m, n;
    int method;
    hipStream_t s;
    MbScratch mb;
    uint8_t *arena;
    size_t used, cap;

    uint32_t *take(size_t bytes)
    {
        uint32_t *q = (uint32_t *)(arena + used);
        used += align256(bytes);
        return q;
    }
};

DevView polfirst_view(uint32_t *base, size_t rows, size_t cols, size_t stor, size_t size)
{
    DevView v;
    v.base = base;
    v.layout = LBX_POLFIRST;
    v.rows = rows;
    v.cols = cols;
    v.stor = stor;
    v.shift = 0;
    v.size = size;
    return v;
}

size_t mc_lds_words(size_t m, size_t n, size_t order, size_t cw)
{
    return cw * ((order + 1) * m + order * n) + m * n + m * (n + 1) + m * n;
}

// Column split of m_basis_cols_kernel: *G groups of *cw columns, or *G = 0 for
// the one-workgroup kernel (orders above kMcMaxOrder, or LDS too small for one
// column).  By default about m^2 n / 8192 groups (at most 16): the work each
// group repeats (sort and elimination) against the work it shares out.
// LBX_MBASIS_GROUPS=N asks for N groups, 0 for the one-workgroup kernel.
// The grid of m_basis_cols_kernel must fit on the device at once: one group
// per CU (its LDS takes most of a CU's 160 KiB).
lbx_status mc_check_groups(int G)
{
    int dev = 0, cus = 0;
    LBX_TRY_HIP(hipGetDevice(&dev));
    LBX_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return G <= cus ? LBX_OK : LBX_ERR_UNSUPPORTED;
}

void cols_split(size_t m, size_t n, size_t order, int *G, int *cw)
{
    *G = 0;
    *cw = 0;
    if (order > (size_t)kMcMaxOrder) return;
    size_t target = std::min<size_t>(16, std::max<size_t>(1, m * m * n / 8192));
    if (const char *e = getenv("LBX_MBASIS_GROUPS")) {
        const long want = atol(e);
        if (want <= 0) return;
        target = (size_t)want;
    }
    size_t w = (m + target - 1) / target;
    while (w > 0 && mc_lds_words(m, n, order, w) > (size_t)kMcLdsWords) --w;
    if (!w) return;
    const size_t groups = (m + w - 1) / w;
    if (groups > (size_t)kMcMaxGroups) return;
    *G = (int)groups;
    *cw = (int)w;
}

// T = a_top . b_top (m x m coefficient matrices: coefficient ta of a, tb of b);
// c[.][0] -= T and c[.][P] = T.  A cyclic product mod x^P - 1 of two operands
// whose full product has P+1 coefficients holds coefficient P in coefficient
// 0; this puts it back.  With fix = false only the degree word is written.
// deg = d[0] + d[1] (the degrees of sigma2 and sigma1).
__global__ void __launch_bounds__(256) wrap_fix_kernel(DevView a, size_t ta, DevView b, size_t tb, DevView c,
                                                       size_t P, int m, uint32_t p, uint64_t mu, uint32_t r64,
                                                       bool fix, const uint64_t *d, uint64_t *deg)
{
    // one workgroup per row i of T, one thread per column j
    __shared__ uint32_t arow[kMbMaxM];
    const int i = blockIdx.x, j = threadIdx.x;
    if (i == 0 && j == 0) *deg = d[0] + d[1];
    if (!fix) return;
    if (j < m) arow[j] = a.base[view_index(a, (size_t)i * m + j, ta)];
    __syncthreads();
    if (j >= m) return;
    // the loads of column j are independent: eight in flight at a time
    Acc acc;
    int q = 0;
    for (; q + 8 <= m; q += 8) {
        uint32_t bv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) bv[u] = b.base[view_index(b, (size_t)(q + u) * m + j, tb)];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc.add(arow[q + u], bv[u]);
    }
    for (; q < m; ++q) acc.add(arow[q], b.base[view_index(b, (size_t)q * m + j, tb)]);
    const uint32_t t = acc.get(p, mu, r64);
    const size_t e = (size_t)i * m + j;
    uint32_t *c0 = c.base + view_index(c, e, 0);
    const uint32_t v = *c0;
    *c0 = v >= t ? v - t : v + (p - t);
    c.base[view_index(c, e, P)] = t;
}

// M_Basis on a device view of the serie (coefficients 0..order-1) into sigma.
// The degree goes to *deg on the device; out_count coefficients of sigma are
// written (zero above the degree), or degree+1 when out_count is 0.
lbx_status run_m_basis(PmRun &R, const DevView &serie, const DevView &sigma, size_t order, uint64_t *deg,
                       size_t out_count)
{
    const size_t m = R.m, n = R.n;
    const size_t g = serie.rows * serie.cols * order;
    hipLaunchKernelGGL(gather_coeff_major, dim3(grid_for(g)), dim3(256), 0, R.s, serie, order, R.mb.S, R.mb.bar);
    LBX_TRY_HIP(hipGetLastError());
    uint64_t mu = ~(uint64_t)0 / R.p;
    const uint32_t r32 = (uint32_t)(((uint64_t)1 << 32) % R.p), r64 = (uint32_t)((uint64_t)r32 * r32 % R.p);
    uint32_t f62 = (uint32_t)(((uint64_t)1 << 62) % R.p);
    const bool mont_ok = R.p < (1u << 31) && (R.p & 1);
    uint32_t pinv = 0, r2m = 0, one = 1;
    if (mont_ok) {
        uint32_t inv = R.p;  // Newton iteration for p^-1 mod 2^32
        for (int i = 0; i < 5; ++i) inv *= 2u - R.p * inv;
        pinv = 0u - inv;
        one = r32;
        r2m = (uint32_t)((uint64_t)r32 * r32 % R.p);
    }
    int G = 0, cw = 0;
    if (mont_ok) cols_split(m, n, order, &G, &cw);
    if (G > 0) {
        const size_t lds = mc_lds_words(m, n, order, (size_t)cw) * 4;
        // rows of the elimination per wave: m / 16, rounded up to 2, 4 or 8
        const void *kern = m <= 2 * (size_t)kMbWaves   ? (const void *)m_basis_cols_kernel<2>
                           : m <= 4 * (size_t)kMbWaves ? (const void *)m_basis_cols_kernel<4>
                                                       : (const void *)m_basis_cols_kernel<8>;
        LBX_TRY_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const uint32_t *S = R.mb.S;
        DevView out = sigma;
        uint64_t *shift_io = R.mb.shift;
        uint32_t *part = R.mb.part;
        unsigned *bar = R.mb.bar;
        int order_i = (int)order, m_i = (int)m, n_i = (int)n, oc = (int)out_count;
        uint32_t p = R.p;
        void *args[] = {&S,   &out, &shift_io, &deg, &part, &bar,  &order_i, &m_i, &n_i, &cw,
                        &oc,  &p,   &mu,       &f62, &pinv, &r2m, &one};
        // Every group waits for the others at each step, so all G must become
        // resident.  G <= 64 groups of one CU each fit on the device together,
        // and groups still queued behind other streams' work are dispatched as
        // that work drains (the barrier's poll limit bounds the wait), so an
        // ordinary launch on the caller's stream serves.  A cooperative launch
        // would go through a queue of its own, with a cross-queue hand-off
        // before and after every base case.
        LBX_TRY(mc_check_groups(G));
        LBX_TRY_HIP(hipLaunchKernel(kern, dim3(G), dim3(kMbThreads), args, lds, R.s));
        return LBX_OK;
    }
    // staged words: enough for every coefficient of a step, at most kMbStageWords
    const size_t need = std::max((order + 1) * (m * m + m * n), m * m);
    const int stage = (int)std::min(need, (size_t)kMbStageWords);
    const size_t lds = (3 * m * n + (size_t)stage) * 4;
    auto kern = mont_ok ? m_basis_kernel<AccSmall, true>
                        : (R.p < (1u << 31) ? m_basis_kernel<AccSmall, false> : m_basis_kernel<AccWide, false>);
    LBX_TRY_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(1), dim3(kMbThreads), lds, R.s, R.mb.S, R.mb.W, sigma, R.mb.shift, deg, (int)order,
                       (int)m, (int)n, stage, (int)out_count, R.p, mu, r64, f62, pinv, r2m, one);
    LBX_TRY_HIP(hipGetLastError());
    return LBX_OK;
}

unsigned log2_ceil(size_t x)
{
    unsigned L = 0;
    while (((size_t)1 << L) < x) ++L;
    return L;
}

// c = a.b on device views with the routes of lbx_polymat_mul_u32.
lbx_status product(PmRun &R, const DevView &a, const DevView &b, const DevView &c, size_t m, size_t k, size_t n)
{
    const size_t sa = a.size, sb = b.size, deg = sa + sb - 2;
    const unsigned L = log2_ceil(deg + 1);
    const bool small_p = R.p < (1u << 31);
    if (sa + sb <= 2 && small_p) return constant_product(R.ctx, R.p, a, b, c, m, k, n, R.s);
    if (small_p && (R.p - 1) % ((size_t)1 << L) == 0)
        return fft_prime_product(R.ctx, R.p, L, a, b, c, m, k, n, deg + 1, R.method, R.s);
    return mp_product(R.ctx, &R.p, 1, a, b, c, m, k, n, R.method, R.s);
}

// sigma = sigma2.sigma1 (:297) on all stored coefficients: sigma2 has o2+1 and
// sigma1 o1+1 (zero above their degrees), so sigma gets order+1 coefficients,
// zero above d1+d2, without the degrees being read back.  When order is a
// power of two and an FFT prime reaches it, the product is taken mod
// x^order - 1 (the length the true degrees would need, since d1+d2 <= order)
// and the one coefficient that wraps is put back.  *deg = d[0] + d[1].
lbx_status sigma_product(PmRun &R, uint32_t *s2, size_t o2, uint32_t *s1, size_t o1, const DevView &sigma,
                         size_t order, const uint64_t *d, uint64_t *deg)
{
    const size_t m = R.m;
    const DevView a = polfirst_view(s2, m, m, o2 + 1, o2 + 1), b = polfirst_view(s1, m, m, o1 + 1, o1 + 1);
    const unsigned Lp = log2_ceil(order);
    const size_t P = (size_t)1 << Lp;
    const uint64_t mu = ~(uint64_t)0 / R.p;
    const uint32_t r32 = (uint32_t)(((uint64_t)1 << 32) % R.p), r64 = (uint32_t)((uint64_t)r32 * r32 % R.p);
    DevView c = sigma;
    const bool wrap = P == order && R.p < (1u << 31) && (R.p - 1) % P == 0;
    if (wrap) {
        c.size = P;
        LBX_TRY(fft_prime_product(R.ctx, R.p, Lp, a, b, c, m, m, m, P, R.method, R.s));
    } else {
        c.size = order + 1;
        LBX_TRY(product(R, a, b, c, m, m, m));
    }
    hipLaunchKernelGGL(wrap_fix_kernel, dim3(wrap ? (unsigned)m : 1u), dim3(256), 0, R.s, a, o2, b, o1, sigma, P,
                       (int)m, R.p, mu, r64, wrap, d, deg);
    return from_hip(hipGetLastError());
}

// PM_Basis (:219-326) over device buffers.  sigma gets order+1 coefficients,
// zero above the degree (written to *deg on the device), except that a call
// that is one base case writes out_count of them (degree+1 when 0).
lbx_status pm_basis(PmRun &R, const DevView &serie, const DevView &sigma, size_t order, uint64_t *deg,
                    size_t out_count)
{
    if (order <= kMbThreshold) return run_m_basis(R, serie, sigma, order, deg, out_count);
    const size_t m = R.m, n = R.n, o1 = order >> 1, o2 = order - o1;
    const size_t mark = R.used;
    uint32_t *s1 = R.take(m * m * (o1 + 1) * 4), *ser2 = R.take(m * n * o2 * 4), *s2 = R.take(m * m * (o2 + 1) * 4);
    uint64_t *d12 = (uint64_t *)R.take(16);
    DevView serie1 = serie;
    serie1.size = o1;  // serie1.copy(serie, 0, ord1-1) (:258) as a view
    LBX_TRY(pm_basis(R, serie1, polfirst_view(s1, m, m, o1 + 1, o1 + 1), o1, d12, o1 + 1));
    // serie2 = coefficients ord1..order-1 of sigma1.serie: midproductgen(serie2, sigma1, serie, true,
    // ord1+1, ord1+ord2) (:274) with c.size() = ord2 > 4, the FFT middle product; sigma1 enters
    // with its o1+1 stored coefficients, zero above its degree
    const DevView a = polfirst_view(s1, m, m, o1 + 1, o1 + 1);
    DevView b = serie;
    b.size = order;
    const DevView c = polfirst_view(ser2, m, n, o2, o2);
    const MidSpec mid{true, o1 + 1};
    const unsigned L = log2_ceil(order);
    if (R.p < (1u << 31) && (R.p - 1) % ((size_t)1 << L) == 0)
        LBX_TRY(fft_prime_product(R.ctx, R.p, L, a, b, c, m, m, n, o2, R.method, R.s, &mid));
    else
        LBX_TRY(mp_product(R.ctx, &R.p, 1, a, b, c, m, m, n, R.method, R.s, &mid, L));
    LBX_TRY(pm_basis(R, polfirst_view(ser2, m, n, o2, o2), polfirst_view(s2, m, m, o2 + 1, o2 + 1), o2, d12 + 1,
                     o2 + 1));
    // sigma = sigma2.sigma1 (:297); the resize to d1+d2+1 is the caller's
    LBX_TRY(sigma_product(R, s2, o2, s1, o1, sigma, order, d12, deg));
    R.used = mark;
    return LBX_OK;
}

lbx_status order_basis(lbx_ctx *ctx, uint32_t p, lbx_polmat *sigma, const lbx_polmat *serie, size_t order,
                       size_t *shift, size_t *degree, int method, void *stream, bool divide)
{
    if (!ctx || !sigma || !serie || !shift || !degree) return LBX_ERR_INVALID_ARG;
    LBX_TRY(check_view(sigma));
    LBX_TRY(check_view(serie));
    if (p < 2 || !order || method < 0 || method > 2) return LBX_ERR_INVALID_ARG;
    const size_t m = serie->rows, n = serie->cols;
    if (sigma->rows != m || sigma->cols != m) return LBX_ERR_SIZE_MISMATCH;
    if (serie->size < order || sigma->size < order + 1) return LBX_ERR_SIZE_MISMATCH;
    if (m > (size_t)kMbMaxM || n > (size_t)kMbMaxN || m * n > (size_t)kMbMaxOut * kMbThreads)
        return LBX_ERR_UNSUPPORTED;
    // PM_Basis: the middle products and products of the recursion need
    // transforms of up to 2^ceil(log2(order+1)) points; M_Basis alone keeps
    // its sigma history (order+1 coefficient matrices) on the device
    const size_t max_order = divide ? ((size_t)1 << kMaxLog) - 1 : ((size_t)1 << kMaxLdsLog) - 1;
    if (order > max_order) return LBX_ERR_UNSUPPORTED;
    for (size_t i = 0; i < m; ++i)
        if (shift[i] >= kShiftLimit) return LBX_ERR_UNSUPPORTED;

    std::lock_guard<std::mutex> lock(ctx->mu);
    DeviceGuard guard(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    StreamOrder stream_order(ctx, s);
    const size_t base_order = divide ? std::min(order, kMbThreshold) : order;
    const size_t b_S = align256(m * n * base_order * 4), b_W = align256((base_order + 1) * m * m * 4);
    const size_t b_part = align256(2 * (size_t)kMcMaxGroups * m * n * 4);
    const size_t b_sh = align256(m * 8), b_small = 256;
    const size_t arena = divide ? pm_level_bytes(m, n, order) : 0;
    void *scratch = nullptr;
    LBX_TRY(ensure(ctx, "order_basis", b_S + b_W + b_part + b_sh + 2 * b_small + arena, &scratch));
    uint8_t *base = (uint8_t *)scratch;
    PmRun R{ctx, p, m, n, method, s, {}, base + b_S + b_W + b_part + b_sh + 2 * b_small, 0, arena};
    R.mb.S = (uint32_t *)base;
    R.mb.W = (uint32_t *)(base + b_S);
    R.mb.part = (uint32_t *)(base + b_S + b_W);
    R.mb.shift = (uint64_t *)(base + b_S + b_W + b_part);
    R.mb.deg = (uint64_t *)(base + b_S + b_W + b_part + b_sh);
    R.mb.bar = (unsigned *)(base + b_S + b_W + b_part + b_sh + b_small);

    Staged ss, sg;
    lbx_status st = stage_in(ctx, "stage_a", serie, 4, true, &ss, s);
    if (st == LBX_OK) st = stage_in(ctx, "stage_c", sigma, 4, true, &sg, s);
    std::vector<uint64_t> sh(shift, shift + m);
    uint64_t d = 0;
    unsigned failed = 0;
    if (st == LBX_OK) st = from_hip(hipMemcpyAsync(R.mb.shift, sh.data(), m * 8, hipMemcpyHostToDevice, s));
    if (st == LBX_OK) st = from_hip(hipMemsetAsync(R.mb.bar, 0, 8, s));
    if (st == LBX_OK) {
        const DevView dser = devview(serie, ss.dev), dsig = devview(sigma, sg.dev);
        st = divide ? pm_basis(R, dser, dsig, order, R.mb.deg, 0) : run_m_basis(R, dser, dsig, order, R.mb.deg, 0);
    }
    // the one read-back of the call: shift, degree and the grid barrier's failure flag
    if (st == LBX_OK) st = from_hip(hipMemcpyAsync(sh.data(), R.mb.shift, m * 8, hipMemcpyDeviceToHost, s));
    if (st == LBX_OK) st = from_hip(hipMemcpyAsync(&d, R.mb.deg, 8, hipMemcpyDeviceToHost, s));
    if (st == LBX_OK) st = from_hip(hipMemcpyAsync(&failed, R.mb.bar + 1, 4, hipMemcpyDeviceToHost, s));
    if (st == LBX_OK && sg.owned)
        st = device_to_host(ctx, sigma->data, sg.dev, storage_words(sigma) * 4, s);
    const lbx_status sync = from_hip(hipStreamSynchronize(s));
    if (st == LBX_OK) st = sync;
    if (st == LBX_OK && failed) st = LBX_ERR_HIP;  // a grid barrier gave up waiting
    release(ss);
    release(sg);
    if (st == LBX_OK) {
        for (size_t i = 0; i < m; ++i) shift[i] = (size_t)sh[i];
        *degree = (size_t)d;
        sigma->size = (size_t)d + 1;
    }
    return st;
}

}  // namespace
}  // namespace lbx

using namespace lbx;

extern "C" lbx_status lbx_order_basis_pm_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *sigma, const lbx_polmat *serie,
                                             size_t order, size_t *shift, size_t *degree, int method, void *stream)
{
    return guarded([&]() -> lbx_status {
        return order_basis(ctx, p, sigma, serie, order, shift, degree, method, stream, true);
    });
}

extern "C" lbx_status lbx_order_basis_m_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *sigma, const lbx_polmat *serie,
                                            size_t order, size_t *shift, size_t *degree, void *stream)
{
    return guarded([&]() -> lbx_status {
        return order_basis(ctx, p, sigma, serie, order, shift, degree, LBX_METHOD_AUTO, stream, false);
    });
}
