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
#include <vector>

#include "lbx_ctx.h"

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

// a*x - b*y mod p
__device__ __forceinline__ uint32_t axmby(uint32_t a, uint32_t x, uint32_t b, uint32_t y, uint32_t p, uint64_t mu)
{
    const uint32_t u = mulm(a, x, p, mu), v = mulm(b, y, p, mu);
    return u >= v ? u - v : u + (p - v);
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

__global__ void __launch_bounds__(256) gather_coeff_major(DevView v, size_t count, uint32_t *out)
{
    const size_t entries = v.rows * v.cols, total = entries * count;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t t = i / entries, e = i - t * entries;
        out[i] = v.base[view_index(v, e, t)];
    }
}

// One workgroup, kMbThreads threads.  S: serie, coefficient-major
// S[(t*m + j)*n + c]; W: (order+1) coefficient matrices of m x m words,
// W[(l*m + phys)*m + j], phys a physical row.  Dynamic LDS: R (m x n) and
// Z (m x n) words.
__global__ void __launch_bounds__(kMbThreads) m_basis_kernel(const uint32_t *__restrict__ S, uint32_t *__restrict__ W,
                                                             DevView out, uint64_t *__restrict__ shift_io,
                                                             uint64_t *__restrict__ deg_out, int order, int m, int n,
                                                             uint32_t p, uint64_t mu, uint32_t r64)
{
    extern __shared__ uint32_t dyn[];
    uint32_t *R = dyn;          // discrepancy rows under elimination, logical row order
    uint32_t *Z = dyn + m * n;  // coefficients over the pivot rows, then the combination
    __shared__ uint64_t shift[kMbMaxM], degree[kMbMaxM], tmp64[kMbMaxM], tmp64b[kMbMaxM];
    __shared__ int rowmap[kMbMaxM], tmpmap[kMbMaxM], pivot_of[kMbMaxM], newpos[kMbMaxM], oldrow[kMbMaxM];
    __shared__ uint32_t scale[kMbMaxM];
    __shared__ unsigned smin[2];
    __shared__ int s_maxdeg;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t mm = (size_t)m * m, nw = (size_t)(order + 1) * mm;
    for (size_t i = tid; i < nw; i += kMbThreads) W[i] = 0;
    __syncthreads();
    for (int i = tid; i < m; i += kMbThreads) {
        W[(size_t)i * m + i] = 1;
        rowmap[i] = i;
        shift[i] = shift_io[i];
        degree[i] = 0;
    }
    if (tid == 0) s_maxdeg = 0;
    __syncthreads();

    for (int k = 0; k < order; ++k) {
        const int maxdeg = s_maxdeg, lim = k < maxdeg ? k : maxdeg;
        // ---- selection sort of the shift (order-basis.h:388-399), wave 0 ----
        if (wave == 0) {
            // positions lane and lane+64; key = shift << 7 | position
            uint64_t sh0 = lane < m ? shift[lane] : 0, sh1 = lane + 64 < m ? shift[lane + 64] : 0;
            uint64_t dg0 = lane < m ? degree[lane] : 0, dg1 = lane + 64 < m ? degree[lane + 64] : 0;
            int rm0 = lane < m ? rowmap[lane] : 0, rm1 = lane + 64 < m ? rowmap[lane + 64] : 0;
            for (int i = 0; i < m - 1; ++i) {
                uint64_t best = ~(uint64_t)0;
                if (lane < m && lane >= i) best = sh0 << 7 | (uint64_t)lane;
                if (lane + 64 < m && lane + 64 >= i) {
                    const uint64_t c1 = sh1 << 7 | (uint64_t)(lane + 64);
                    best = c1 < best ? c1 : best;
                }
                for (int off = 32; off; off >>= 1) {
                    const uint64_t o = __shfl_xor(best, off);
                    best = o < best ? o : best;
                }
                const int im = (int)(best & 127);
                if (im != i) {  // swap positions i and im (uniform branch)
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
        }
        __syncthreads();
        // ---- discrepancy: coefficient k of sigma.serie (order-basis.h:416-444) ----
        for (int o = tid; o < m * n; o += kMbThreads) {
            const int r = o / n, c = o - r * n;
            const uint32_t *wrow = W + (size_t)rowmap[r] * m;
            Acc acc;
            for (int l = 0; l <= lim; ++l) {
                const uint32_t *wl = wrow + (size_t)l * mm;
                const uint32_t *sl = S + ((size_t)(k - l) * m) * n + c;
                for (int j = 0; j < m; ++j) acc.add(wl[j], sl[(size_t)j * n]);
            }
            R[o] = acc.get(p, mu, r64);
            Z[o] = 0;
        }
        for (int i = tid; i < m; i += kMbThreads) {
            scale[i] = 1;
            pivot_of[i] = -1;
        }
        if (tid == 0) {
            smin[0] = 0xffffffffu;
            smin[1] = 0xffffffffu;
        }
        __syncthreads();
        // ---- row rank profile and combinations (the PLUQ of :470-488) ----
        int next = 0, rank = 0;
        for (;;) {
            const int slot = rank & 1;
            // first nonzero entry (row-major) among rows >= next
            for (int o = next * n + tid; o < m * n; o += kMbThreads)
                if (R[o]) atomicMin(&smin[slot], (unsigned)o);
            __syncthreads();
            const unsigned hit = smin[slot];
            if (hit == 0xffffffffu) break;
            const int s = (int)(hit / n), c0 = (int)(hit - (unsigned)s * n);
            const uint32_t a = R[s * n + c0];
            // rows below s: row = a*row - b*pivot row; one wave per row, one lane per column
            for (int r = s + 1 + wave; r < m; r += kMbWaves) {
                const uint32_t b = R[r * n + c0];
                if (b == 0) continue;
                uint32_t rv = 0, zv = 0;
                if (lane < n) rv = axmby(a, R[r * n + lane], b, R[s * n + lane], p, mu);
                if (lane < rank) zv = axmby(a, Z[r * n + lane], b, Z[s * n + lane], p, mu);
                else if (lane == rank) zv = mulm(b, scale[s], p, mu);
                const uint32_t sc = mulm(a, scale[r], p, mu);
                if (lane < n) R[r * n + lane] = rv;
                if (lane <= rank) Z[r * n + lane] = zv;
                if (lane == 0) scale[r] = sc;
            }
            if (tid == 0) {
                pivot_of[s] = rank;
                smin[slot ^ 1] = 0xffffffffu;
            }
            next = s + 1;
            ++rank;
            __syncthreads();
        }
        // non-pivot rows: combination = Z / scale
        for (int r = wave; r < m; r += kMbWaves) {
            if (pivot_of[r] >= 0) continue;
            const uint32_t inv = scale[r] == 1 ? 1u : invm(scale[r], p, mu);
            if (lane < rank) Z[r * n + lane] = mulm(Z[r * n + lane], inv, p, mu);
        }
        // Qt: pivot rows first, then the others, each in increasing order
        for (int r = tid; r < m; r += kMbThreads) {
            int pos = pivot_of[r];
            if (pos < 0) {
                int before = 0;
                for (int q = 0; q < r; ++q) before += pivot_of[q] < 0 ? 1 : 0;
                pos = rank + before;
            }
            newpos[r] = pos;
            oldrow[pos] = r;
            tmpmap[pos] = rowmap[r];
            tmp64[pos] = shift[r];
            tmp64b[pos] = degree[r];
        }
        __syncthreads();
        for (int i = tid; i < m; i += kMbThreads) {
            rowmap[i] = tmpmap[i];
            shift[i] = tmp64[i];
            degree[i] = tmp64b[i];
        }
        __syncthreads();
        // ---- sigma rows rank.. -= combination . pivot rows, coefficients 0..lim (:509-517) ----
        {
            const int nrows = m - rank;
            const size_t total = (size_t)(lim + 1) * nrows * m;
            for (size_t o = tid; o < total; o += kMbThreads) {
                const int j = (int)(o % m);
                const size_t q = o / m;
                const int sidx = rank + (int)(q % nrows), l = (int)(q / nrows);
                const uint32_t *Y = Z + (size_t)oldrow[sidx] * n;
                const uint32_t *wl = W + (size_t)l * mm + j;
                Acc acc;
                for (int t = 0; t < rank; ++t) acc.add(Y[t], wl[(size_t)rowmap[t] * m]);
                const uint32_t sub = acc.get(p, mu, r64);
                uint32_t *dst = W + (size_t)l * mm + (size_t)rowmap[sidx] * m + j;
                const uint32_t v = *dst;
                *dst = v >= sub ? v - sub : v + (p - sub);
            }
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
        // ---- pivot rows times x (:573-580) ----
        {
            const int nmd = s_maxdeg;
            for (int o = tid; o < rank * m; o += kMbThreads) {
                const int t = o / m, j = o - t * m;
                uint32_t *col = W + (size_t)rowmap[t] * m + j;
                for (int l = nmd - 1; l >= 0; --l) col[(size_t)(l + 1) * mm] = col[(size_t)l * mm];
                col[0] = 0;
            }
        }
        __syncthreads();
    }
    // sigma.resize(max_degree+1), shift2 = shift (:592-597)
    const int md = s_maxdeg;
    for (size_t o = tid; o < (size_t)(md + 1) * mm; o += kMbThreads) {
        const size_t l = o / mm, e = o - l * mm, i = e / m, j = e - i * m;
        out.base[view_index(out, e, l)] = W[(l * m + rowmap[i]) * m + j];
    }
    for (int i = tid; i < m; i += kMbThreads) shift_io[i] = shift[i];
    if (tid == 0) *deg_out = (uint64_t)md;
}

unsigned grid_for(size_t total)
{
    const size_t g = (total + 255) / 256;
    return (unsigned)(g < 4096 ? (g ? g : 1) : 4096);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device bytes of the three buffers each recursion level keeps live, summed
// down the deepest path.
size_t pm_level_bytes(size_t m, size_t n, size_t order)
{
    if (order <= kMbThreshold) return 0;
    const size_t o1 = order >> 1, o2 = order - o1;
    const size_t own = align256(m * m * (o1 + 1) * 4) + align256(m * n * o2 * 4) + align256(m * m * (o2 + 1) * 4);
    return own + std::max(pm_level_bytes(m, n, o1), pm_level_bytes(m, n, o2));
}

struct MbScratch {
    uint32_t *S, *W;
    uint64_t *shift, *deg;
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

// M_Basis on a device view of the serie (coefficients 0..order-1) into sigma.
lbx_status run_m_basis(PmRun &R, const DevView &serie, const DevView &sigma, size_t order, size_t *deg)
{
    const size_t m = R.m, n = R.n;
    const size_t g = serie.rows * serie.cols * order;
    hipLaunchKernelGGL(gather_coeff_major, dim3(grid_for(g)), dim3(256), 0, R.s, serie, order, R.mb.S);
    LBX_TRY_HIP(hipGetLastError());
    const uint64_t mu = ~(uint64_t)0 / R.p;
    const uint32_t r32 = (uint32_t)(((uint64_t)1 << 32) % R.p), r64 = (uint32_t)((uint64_t)r32 * r32 % R.p);
    const size_t lds = 2 * m * n * 4;
    LBX_TRY_HIP(hipFuncSetAttribute((const void *)m_basis_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    hipLaunchKernelGGL(m_basis_kernel, dim3(1), dim3(kMbThreads), lds, R.s, R.mb.S, R.mb.W, sigma, R.mb.shift,
                       R.mb.deg, (int)order, (int)m, (int)n, R.p, mu, r64);
    LBX_TRY_HIP(hipGetLastError());
    uint64_t d = 0;
    LBX_TRY_HIP(hipMemcpyAsync(&d, R.mb.deg, 8, hipMemcpyDeviceToHost, R.s));
    LBX_TRY_HIP(hipStreamSynchronize(R.s));
    *deg = (size_t)d;
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

lbx_status pm_basis(PmRun &R, const DevView &serie, const DevView &sigma, size_t order, size_t *deg)
{
    if (order <= kMbThreshold) return run_m_basis(R, serie, sigma, order, deg);
    const size_t m = R.m, n = R.n, o1 = order >> 1, o2 = order - o1;
    const size_t mark = R.used;
    uint32_t *s1 = R.take(m * m * (o1 + 1) * 4), *ser2 = R.take(m * n * o2 * 4), *s2 = R.take(m * m * (o2 + 1) * 4);
    size_t d1 = 0, d2 = 0;
    DevView serie1 = serie;
    serie1.size = o1;  // serie1.copy(serie, 0, ord1-1) (:258) as a view
    LBX_TRY(pm_basis(R, serie1, polfirst_view(s1, m, m, o1 + 1, o1 + 1), o1, &d1));
    // serie2 = coefficients ord1..order-1 of sigma1.serie: midproductgen(serie2, sigma1, serie, true,
    // ord1+1, ord1+ord2) (:274) with c.size() = ord2 > 4, the FFT middle product
    const DevView a = polfirst_view(s1, m, m, o1 + 1, d1 + 1);
    DevView b = serie;
    b.size = order;
    const DevView c = polfirst_view(ser2, m, n, o2, o2);
    const MidSpec mid{true, o1 + 1};
    const unsigned L = log2_ceil(order);
    if (R.p < (1u << 31) && (R.p - 1) % ((size_t)1 << L) == 0)
        LBX_TRY(fft_prime_product(R.ctx, R.p, L, a, b, c, m, m, n, o2, R.method, R.s, &mid));
    else
        LBX_TRY(mp_product(R.ctx, &R.p, 1, a, b, c, m, m, n, R.method, R.s, &mid, L));
    LBX_TRY(pm_basis(R, polfirst_view(ser2, m, n, o2, o2), polfirst_view(s2, m, m, o2 + 1, o2 + 1), o2, &d2));
    // sigma = sigma2.sigma1 (:297), then resize(d1+d2+1)
    DevView out = sigma;
    out.size = d1 + d2 + 1;
    LBX_TRY(product(R, polfirst_view(s2, m, m, o2 + 1, d2 + 1), polfirst_view(s1, m, m, o1 + 1, d1 + 1), out, m, m,
                    m));
    R.used = mark;
    *deg = d1 + d2;
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
    if (m > (size_t)kMbMaxM || n > (size_t)kMbMaxN) return LBX_ERR_UNSUPPORTED;
    if (order > ((size_t)1 << kMaxLdsLog) - 1 || order > 0x7fffffff) return LBX_ERR_UNSUPPORTED;
    for (size_t i = 0; i < m; ++i)
        if (shift[i] >= kShiftLimit) return LBX_ERR_UNSUPPORTED;

    std::lock_guard<std::mutex> lock(ctx->mu);
    DeviceGuard guard(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    const size_t base_order = divide ? std::min(order, kMbThreshold) : order;
    const size_t b_S = align256(m * n * base_order * 4), b_W = align256((base_order + 1) * m * m * 4);
    const size_t b_sh = align256(m * 8), b_deg = 256;
    const size_t arena = divide ? pm_level_bytes(m, n, order) : 0;
    void *scratch = nullptr;
    LBX_TRY(ensure(ctx, "order_basis", b_S + b_W + b_sh + b_deg + arena, &scratch));
    uint8_t *base = (uint8_t *)scratch;
    PmRun R{ctx, p, m, n, method, s, {}, base + b_S + b_W + b_sh + b_deg, 0, arena};
    R.mb.S = (uint32_t *)base;
    R.mb.W = (uint32_t *)(base + b_S);
    R.mb.shift = (uint64_t *)(base + b_S + b_W);
    R.mb.deg = (uint64_t *)(base + b_S + b_W + b_sh);

    Staged ss, sg;
    lbx_status st = stage_in(serie, 4, true, &ss);
    if (st == LBX_OK) st = stage_in(sigma, 4, true, &sg);
    std::vector<uint64_t> sh(shift, shift + m);
    size_t d = 0;
    if (st == LBX_OK) st = from_hip(hipMemcpyAsync(R.mb.shift, sh.data(), m * 8, hipMemcpyHostToDevice, s));
    if (st == LBX_OK) {
        const DevView dser = devview(serie, ss.dev), dsig = devview(sigma, sg.dev);
        st = divide ? pm_basis(R, dser, dsig, order, &d) : run_m_basis(R, dser, dsig, order, &d);
    }
    if (st == LBX_OK) st = from_hip(hipMemcpyAsync(sh.data(), R.mb.shift, m * 8, hipMemcpyDeviceToHost, s));
    if (st == LBX_OK && sg.owned)
        st = from_hip(hipMemcpyAsync(sigma->data, sg.dev, storage_words(sigma) * 4, hipMemcpyDeviceToHost, s));
    const lbx_status sync = from_hip(hipStreamSynchronize(s));
    if (st == LBX_OK) st = sync;
    release(ss);
    release(sg);
    if (st == LBX_OK) {
        for (size_t i = 0; i < m; ++i) shift[i] = (size_t)sh[i];
        *degree = d;
        sigma->size = d + 1;
    }
    return st;
}

}  // namespace
}  // namespace lbx

using namespace lbx;

extern "C" lbx_status lbx_order_basis_pm_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *sigma, const lbx_polmat *serie,
                                             size_t order, size_t *shift, size_t *degree, int method, void *stream)
{
    return order_basis(ctx, p, sigma, serie, order, shift, degree, method, stream, true);
}

extern "C" lbx_status lbx_order_basis_m_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *sigma, const lbx_polmat *serie,
                                            size_t order, size_t *shift, size_t *degree, void *stream)
{
    return order_basis(ctx, p, sigma, serie, order, shift, degree, LBX_METHOD_AUTO, stream, false);
}
