// Middle products of polynomial matrices: the coefficients hdeg-1, hdeg, ...
// of a.b, which OrderBasis::PM_Basis takes at every level
// (order-basis.h:280,820 -> polynomial-matrix-domain.h:77-104).
//
//  * lbx_polymat_midproduct_u32 follows PolynomialMatrixFFTMulDomain::midproduct
//    (matpoly-mult-fft-wordsize.inl:90-127): the FFT-prime midproduct of
//    matpoly-mult-fft-wordsize-fast.inl:209-316 when p < 2^31 and 2^L | p-1,
//    otherwise the same transform product modulo FFT primes with CRT
//    (matpoly-mult-fft-multiprecision.inl:93-117, 427-520).  Both give the
//    cyclic correlation of the reversed small operand with the other one over
//    pts = 2^ceil(log2 deg) points; its first deg-hdeg+1 values are the middle
//    product, the later ones (up to c.size) wrap around.
//  * lbx_polymat_midproduct_naive_u32 follows PolynomialMatrixNaiveMulDomain::
//    midproduct (matpoly-mult-naive.h:85-125), the route midproductgen takes
//    for c.size() <= 4: coefficients hdeg-1 .. min(sa+sb-1, deg)-1 of a.b by
//    direct summation, the other coefficients of c untouched.
#include <algorithm>

#include "lbx_ctx.h"

namespace lbx {
namespace {

__global__ void __launch_bounds__(256) reverse_gather_kernel(DevView in, size_t entries, size_t hdeg, uint32_t *out)
{
    const size_t total = entries * hdeg;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t e = i / hdeg, t = i - e * hdeg, src = hdeg - 1 - t;
        out[i] = src < in.size ? in.base[view_index(in, e, src)] : 0u;
    }
}

// One thread per output coefficient (entry e, index t): sum over q < k and
// over the admissible u of a(i,q,u) b(q,j,kk-u), kk = t + hdeg - 1, with
// products added into (acc:u64, hi:u32) and reduced once at the end in 64-bit
// arithmetic (any p < 2^32; r32 = 2^32 mod p, r64 = 2^64 mod p).
__global__ void __launch_bounds__(256) midproduct_naive_kernel(DevView a, DevView b, DevView c, int k, size_t hdeg,
                                                               size_t count, uint32_t p, uint32_t r32, uint32_t r64)
{
    const size_t n = c.cols, total = c.rows * n * count;
    const size_t sa = a.size, sb = b.size;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        const size_t e = idx / count, t = idx - e * count;
        const size_t i = e / n, j = e - i * n, kk = t + hdeg - 1;
        const size_t u_lo = kk + 1 > sb ? kk + 1 - sb : 0, u_hi = kk < sa - 1 ? kk : sa - 1;
        uint64_t acc = 0;
        uint32_t hi = 0;
        for (int q = 0; q < k; ++q)
            for (size_t u = u_lo; u <= u_hi; ++u) {
                const uint64_t prod = (uint64_t)a.base[view_index(a, i * k + q, u)] *
                                      b.base[view_index(b, (size_t)q * n + j, kk - u)];
                const uint64_t s2 = acc + prod;
                hi += s2 < prod ? 1u : 0u;
                acc = s2;
            }
        const uint64_t r = ((uint64_t)hi * r64 % p + (acc >> 32) * r32 % p + (acc & 0xffffffffu)) % p;
        c.base[view_index(c, e, t)] = (uint32_t)r;
    }
}

unsigned grid_for(size_t total)
{
    const size_t g = (total + 255) / 256;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

// hdeg, deg and the transform length of midproduct (wordsize-fast.inl:211-221).
struct MidShape {
    size_t hdeg, deg, pts;
    unsigned L;
};

lbx_status mid_shape(const lbx_polmat *c, size_t n0, size_t n1, MidShape *out)
{
    out->hdeg = n0 ? n0 : c->size;
    if (!out->hdeg) return LBX_ERR_INVALID_ARG;
    out->deg = n1 ? n1 : 2 * out->hdeg - 1;
    out->pts = 1;
    out->L = 0;
    while (out->pts < out->deg) {
        out->pts <<= 1;
        ++out->L;
    }
    return LBX_OK;
}

}  // namespace

hipError_t launch_reverse_gather(const DevView &in, size_t entries, size_t hdeg, uint32_t *out, hipStream_t s)
{
    if (!entries || !hdeg) return hipSuccess;
    hipLaunchKernelGGL(reverse_gather_kernel, dim3(grid_for(entries * hdeg)), dim3(256), 0, s, in, entries, hdeg, out);
    return hipGetLastError();
}

}  // namespace lbx

using namespace lbx;

namespace {

// Stages host operands, runs `body` on device views and copies c back.
template <typename F>
lbx_status with_views(lbx_ctx *ctx, lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b, bool copy_c, void *stream,
                      F body)  // copy_c: stage c's host storage in as well as out
{
    std::lock_guard<std::mutex> lock(ctx->mu);
    DeviceGuard guard(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    StreamOrder stream_order(ctx, s);
    Staged sa_, sb_, sc_;
    lbx_status st = stage_in(ctx, "stage_a", a, 4, true, &sa_, s);
    if (st == LBX_OK) st = stage_in(ctx, "stage_b", b, 4, true, &sb_, s);
    if (st == LBX_OK) st = stage_in(ctx, "stage_c", c, 4, copy_c, &sc_, s);
    const bool host_io = sa_.owned || sb_.owned || sc_.owned;
    if (st == LBX_OK) st = body(devview(a, sa_.dev), devview(b, sb_.dev), devview(c, sc_.dev), s);
    if (st == LBX_OK && sc_.owned) {
        st = device_to_host(ctx, c->data, sc_.dev, storage_words(c) * 4, s);
    } else if (st == LBX_OK && host_io) {
        st = from_hip(hipStreamSynchronize(s));
    }
    if (host_io) (void)hipStreamSynchronize(s);
    release(sa_);
    release(sb_);
    release(sc_);
    return st;
}

lbx_status check_args(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                      int method)
{
    if (!ctx || !a || !b || !c) return LBX_ERR_INVALID_ARG;
    LBX_TRY(check_view(a));
    LBX_TRY(check_view(b));
    LBX_TRY(check_view(c));
    if (method < 0 || method > 2 || p < 2) return LBX_ERR_INVALID_ARG;
    if (a->cols != b->rows || c->rows != a->rows || c->cols != b->cols) return LBX_ERR_SIZE_MISMATCH;
    if (!a->size || !b->size || !c->size) return LBX_ERR_INVALID_ARG;
    return LBX_OK;
}

}  // namespace

extern "C" lbx_status lbx_polymat_midproduct_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a,
                                                 const lbx_polmat *b, int small_left, size_t n0, size_t n1,
                                                 int method, void *stream)
{
    return guarded([&]() -> lbx_status {
        LBX_TRY(check_args(ctx, p, c, a, b, method));
        MidShape sh;
        LBX_TRY(mid_shape(c, n0, n1, &sh));
        const size_t small = small_left ? a->size : b->size, big = small_left ? b->size : a->size;
        // wordsize-fast.inl:214-218; sizes beyond pts would overrun the padded
        // copies there, and a small operand longer than hdeg is not reversed whole
        if (sh.deg < sh.hdeg || c->size < sh.deg - sh.hdeg || big >= sh.hdeg + sh.deg) return LBX_ERR_SIZE_MISMATCH;
        if (small > sh.hdeg || big > sh.pts || c->size > sh.pts || sh.hdeg > sh.pts) return LBX_ERR_SIZE_MISMATCH;
        if (sh.L == 0) return LBX_ERR_K_NOT_POSITIVE;  // FFT(F, 0): "k must be positive" (fft.h:136-139)
        if (sh.L > (unsigned)kMaxLog) return LBX_ERR_UNSUPPORTED;
        const MidSpec mid{small_left != 0, sh.hdeg};
        const bool single_prime = p < (1u << 31) && (p - 1) % sh.pts == 0;
        // c is copied in too: a view's storage outside the window must come back unchanged
        LBX_TRY(with_views(ctx, c, a, b, true, stream, [&](const DevView &da, const DevView &db, const DevView &dc,
                                                           hipStream_t s) -> lbx_status {
            if (single_prime)
                return fft_prime_product(ctx, p, sh.L, da, db, dc, a->rows, a->cols, b->cols, c->size, method, s, &mid);
            return mp_product(ctx, &p, 1, da, db, dc, a->rows, a->cols, b->cols, method, s, &mid, sh.L);
        }));
        // CHECK_MATPOL_MIDP (polynomial-matrix-domain.h:88-90): the middle
        // coefficients, not those that wrap around past deg-hdeg
        return self_check_midproduct(ctx, p, c, a, b, sh.hdeg, std::min(c->size, sh.deg - sh.hdeg + 1), stream);
    });
}

extern "C" lbx_status lbx_polymat_midproduct_naive_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a,
                                                       const lbx_polmat *b, size_t n0, size_t n1, void *stream)
{
    return guarded([&]() -> lbx_status {
        LBX_TRY(check_args(ctx, p, c, a, b, 0));
        MidShape sh;
        LBX_TRY(mid_shape(c, n0, n1, &sh));
        const size_t top = std::min(a->size + b->size - 1, sh.deg);
        if (top < sh.hdeg) return LBX_OK;  // empty range: nothing written (naive.h:92)
        const size_t count = top - sh.hdeg + 1;
        if (count > c->size) return LBX_ERR_SIZE_MISMATCH;
        if (a->cols > 0x7fffffffu) return LBX_ERR_UNSUPPORTED;
        const uint32_t r32 = (uint32_t)(((uint64_t)1 << 32) % p), r64 = (uint32_t)((uint64_t)r32 * r32 % p);
        LBX_TRY(with_views(ctx, c, a, b, true, stream, [&](const DevView &da, const DevView &db, const DevView &dc,
                                                           hipStream_t s) -> lbx_status {
            const size_t total = a->rows * b->cols * count;
            hipLaunchKernelGGL(midproduct_naive_kernel, dim3(grid_for(total)), dim3(256), 0, s, da, db, dc,
                               (int)a->cols, sh.hdeg, count, p, r32, r64);
            return from_hip(hipGetLastError());
        }));
        return self_check_midproduct(ctx, p, c, a, b, sh.hdeg, count, stream);  // CHECK_MATPOL_MIDP
    });
}
