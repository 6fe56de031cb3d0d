// Check of a product at one point, on the device: the random-point form of
// the reference's check_mul (matpoly-mult-fft.h:87-163, which evaluates
// at 1), run after every lbx_polymat_mul_u32 when a context asks for it, as
// CHECK_MATPOL_MUL does after PolynomialMatrixMulDomain::mul
// (polynomial-matrix-domain.h:72-74).
#include <algorithm>
#include <vector>

#include "lbx_ctx.h"

namespace lbx {
namespace {

// x mod p for any 64-bit x and 2 <= p < 2^32, mu = floor((2^64-1)/p): the
// quotient estimate hi64(x mu) is low by less than 3.
__device__ __forceinline__ uint32_t reduce64(uint64_t x, uint32_t p, uint64_t mu)
{
    uint64_t r = x - __umul64hi(x, mu) * p;
    if (r >= p) r -= p;
    if (r >= p) r -= p;
    return (uint32_t)r;
}

__device__ __forceinline__ uint32_t mulm(uint32_t a, uint32_t b, uint32_t p, uint64_t mu)
{
    return reduce64((uint64_t)a * b, p, mu);
}

__device__ __forceinline__ uint32_t powm(uint32_t x, uint32_t e, uint32_t p, uint64_t mu)
{
    uint32_t r = 1 % p;
    for (; e; e >>= 1) {
        if (e & 1) r = mulm(r, x, p, mu);
        x = mulm(x, x, p, mu);
    }
    return r;
}

// out[e] = sum over t < count of v(e, t) x^t mod p, one wave per entry: lane j
// runs Horner over t = j + 64 i with x^64, then the lanes' values are weighted
// by x^j and summed.
__global__ void __launch_bounds__(256) eval_kernel(DevView v, size_t entries, size_t count, uint32_t x, uint32_t p,
                                                   uint64_t mu, uint32_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const size_t wave = ((size_t)blockIdx.x * 256 + threadIdx.x) >> 6, nw = ((size_t)gridDim.x * 256) >> 6;
    const uint32_t xl = powm(x % p, (uint32_t)lane, p, mu), x64 = powm(x % p, 64, p, mu);
    for (size_t e = wave; e < entries; e += nw) {
        uint32_t acc = 0;
        if ((size_t)lane < count) {
            for (size_t t = lane + 64 * ((count - 1 - lane) / 64);; t -= 64) {
                const uint32_t c = v.base[view_index(v, e, t)];
                acc = reduce64((uint64_t)mulm(acc, x64, p, mu) + (c >= p ? c % p : c), p, mu);
                if (t < 64) break;
            }
        }
        uint64_t s = mulm(acc, xl, p, mu);
        for (int off = 32; off; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) out[e] = reduce64(s, p, mu);
    }
}

// bad += number of (i,j) with sum_q EA[i][q] EB[q][j] != EC[i][j] mod p
__global__ void __launch_bounds__(256) compare_kernel(const uint32_t *__restrict__ ea, const uint32_t *__restrict__ eb,
                                                      const uint32_t *__restrict__ ec, size_t m, size_t k, size_t n,
                                                      uint32_t p, uint64_t mu, unsigned *__restrict__ bad)
{
    for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < m * n; o += (size_t)gridDim.x * 256) {
        const size_t i = o / n, j = o - i * n;
        uint32_t s = 0;
        for (size_t q = 0; q < k; ++q) {
            const uint32_t t = s + mulm(ea[i * k + q], eb[q * n + j], p, mu);
            s = t >= p || t < s ? t - p : t;
        }
        if (s != ec[o]) atomicAdd(bad, 1u);
    }
}

// a + b mod p for a, b < p < 2^32 (the sum may pass 2^32)
__device__ __forceinline__ uint32_t addm(uint32_t a, uint32_t b, uint32_t p)
{
    const uint32_t t = a + b;
    return (t >= p || t < a) ? t - p : t;
}

// Projections for the middle-product check (Freivalds): with random r (rows)
// and v (columns), out = r^T M over M's rows, M v over its columns, or
// r^T M v, one coefficient slot per thread.
__device__ __forceinline__ uint32_t coef(const DevView &M, size_t e, size_t t, uint32_t p)
{
    const uint32_t c = M.base[view_index(M, e, t)];
    return c >= p ? c % p : c;
}

// out[q*size + t] = sum_i r[i] M(i, q, t)
__global__ void __launch_bounds__(256) combine_rows_kernel(DevView M, const uint32_t *__restrict__ r, size_t size,
                                                           uint32_t p, uint64_t mu, uint32_t *__restrict__ out)
{
    const size_t rows = M.rows, cols = M.cols;
    for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < cols * size; o += (size_t)gridDim.x * 256) {
        const size_t q = o / size, t = o - q * size;
        uint32_t acc = 0;
        for (size_t i = 0; i < rows; ++i) acc = addm(acc, mulm(r[i], coef(M, i * cols + q, t, p), p, mu), p);
        out[o] = acc;
    }
}

// out[q*size + t] = sum_j M(q, j, t) v[j]
__global__ void __launch_bounds__(256) combine_cols_kernel(DevView M, const uint32_t *__restrict__ v, size_t size,
                                                           uint32_t p, uint64_t mu, uint32_t *__restrict__ out)
{
    const size_t rows = M.rows, cols = M.cols;
    for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < rows * size; o += (size_t)gridDim.x * 256) {
        const size_t q = o / size, t = o - q * size;
        uint32_t acc = 0;
        for (size_t j = 0; j < cols; ++j) acc = addm(acc, mulm(coef(M, q * cols + j, t, p), v[j], p, mu), p);
        out[o] = acc;
    }
}

// out[t] = sum_{i,j} r[i] M(i, j, t) v[j]
__global__ void __launch_bounds__(256) combine_both_kernel(DevView M, const uint32_t *__restrict__ r,
                                                           const uint32_t *__restrict__ v, size_t count, uint32_t p,
                                                           uint64_t mu, uint32_t *__restrict__ out)
{
    const size_t rows = M.rows, cols = M.cols;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < count; t += (size_t)gridDim.x * 256) {
        uint32_t acc = 0;
        for (size_t i = 0; i < rows; ++i) {
            uint32_t row = 0;
            for (size_t j = 0; j < cols; ++j) row = addm(row, mulm(coef(M, i * cols + j, t, p), v[j], p, mu), p);
            acc = addm(acc, mulm(r[i], row, p, mu), p);
        }
        out[t] = acc;
    }
}

// bad += number of t < count with sum_q sum_u a1[q][u] b1[q][kk-u] != c1[t],
// kk = hdeg-1+t (the coefficient kk of the scalar products, zero outside their
// range); one workgroup per t, its 256 threads over (q, u).
__global__ void __launch_bounds__(256) midcheck_kernel(const uint32_t *__restrict__ a1, size_t sa,
                                                       const uint32_t *__restrict__ b1, size_t sb, size_t k,
                                                       const uint32_t *__restrict__ c1, size_t hdeg, size_t count,
                                                       uint32_t p, uint64_t mu, unsigned *__restrict__ bad)
{
    __shared__ uint32_t part[256];
    for (size_t t = blockIdx.x; t < count; t += gridDim.x) {
        const size_t kk = hdeg - 1 + t;
        const size_t u_lo = kk + 1 > sb ? kk + 1 - sb : 0, u_hi = kk < sa - 1 ? kk : sa - 1;
        const size_t len = u_hi >= u_lo ? u_hi - u_lo + 1 : 0;
        uint32_t acc = 0;
        for (size_t idx = threadIdx.x; idx < k * len; idx += 256) {
            const size_t q = idx / len, u = u_lo + (idx - q * len);
            acc = addm(acc, mulm(a1[q * sa + u], b1[q * sb + (kk - u)], p, mu), p);
        }
        part[threadIdx.x] = acc;
        __syncthreads();
        for (int h = 128; h; h >>= 1) {
            if ((int)threadIdx.x < h) part[threadIdx.x] = addm(part[threadIdx.x], part[threadIdx.x + h], p);
            __syncthreads();
        }
        if (threadIdx.x == 0 && part[0] != c1[t]) atomicAdd(bad, 1u);
        __syncthreads();  // part is rewritten for the next t
    }
}

unsigned blocks_for(size_t threads)
{
    const size_t g = (threads + 255) / 256;
    return (unsigned)(g < 8192 ? (g ? g : 1) : 8192);
}

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

lbx_status check_product(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                         size_t deg, uint32_t x, int *equal, void *stream)
{
    if (!ctx || !a || !b || !c || !equal) return LBX_ERR_INVALID_ARG;
    LBX_TRY(check_view(a));
    LBX_TRY(check_view(b));
    LBX_TRY(check_view(c));
    if (p < 2) return LBX_ERR_INVALID_ARG;
    if (a->cols != b->rows || c->rows != a->rows || c->cols != b->cols) return LBX_ERR_SIZE_MISMATCH;
    if (!deg) deg = c->size;
    if (deg > c->size) return LBX_ERR_SIZE_MISMATCH;
    const size_t m = a->rows, k = a->cols, n = b->cols;
    std::lock_guard<std::mutex> lock(ctx->mu);
    DeviceGuard guard(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    StreamOrder stream_order(ctx, s);
    Staged sa, sb, sc;
    lbx_status st = stage_in(ctx, "check_a", a, 4, true, &sa, s);
    if (st == LBX_OK) st = stage_in(ctx, "check_b", b, 4, true, &sb, s);
    if (st == LBX_OK) st = stage_in(ctx, "check_c", c, 4, true, &sc, s);
    void *ws = nullptr;
    const size_t words = m * k + k * n + m * n;
    if (st == LBX_OK) st = ensure(ctx, "check_eval", words * 4 + 256, &ws);
    unsigned nbad = 0;
    if (st == LBX_OK) {
        uint32_t *ea = (uint32_t *)ws, *eb = ea + m * k, *ec = eb + k * n;
        unsigned *bad = (unsigned *)(ec + m * n);
        const uint64_t mu = ~(uint64_t)0 / p;
        hipLaunchKernelGGL(eval_kernel, dim3(blocks_for(m * k * 64)), dim3(256), 0, s, devview(a, sa.dev), m * k,
                           a->size, x, p, mu, ea);
        hipLaunchKernelGGL(eval_kernel, dim3(blocks_for(k * n * 64)), dim3(256), 0, s, devview(b, sb.dev), k * n,
                           b->size, x, p, mu, eb);
        hipLaunchKernelGGL(eval_kernel, dim3(blocks_for(m * n * 64)), dim3(256), 0, s, devview(c, sc.dev), m * n, deg,
                           x, p, mu, ec);
        st = from_hip(hipMemsetAsync(bad, 0, 4, s));
        if (st == LBX_OK) {
            hipLaunchKernelGGL(compare_kernel, dim3(blocks_for(m * n)), dim3(256), 0, s, ea, eb, ec, m, k, n, p, mu,
                               bad);
            st = from_hip(hipGetLastError());
        }
        if (st == LBX_OK) st = from_hip(hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, s));
        if (st == LBX_OK) st = from_hip(hipStreamSynchronize(s));
    }
    release(sa);
    release(sb);
    release(sc);
    if (st == LBX_OK) *equal = nbad == 0;
    return st;
}

lbx_status check_midproduct(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                            size_t hdeg, size_t count, int *equal, void *stream)
{
    if (!ctx || !a || !b || !c || !equal) return LBX_ERR_INVALID_ARG;
    LBX_TRY(check_view(a));
    LBX_TRY(check_view(b));
    LBX_TRY(check_view(c));
    if (p < 2 || !hdeg) return LBX_ERR_INVALID_ARG;
    if (a->cols != b->rows || c->rows != a->rows || c->cols != b->cols) return LBX_ERR_SIZE_MISMATCH;
    if (count > c->size) return LBX_ERR_SIZE_MISMATCH;
    if (!count || !a->size || !b->size) {
        *equal = 1;
        return LBX_OK;
    }
    const size_t m = a->rows, k = a->cols, n = b->cols, sa = a->size, sb = b->size;
    std::lock_guard<std::mutex> lock(ctx->mu);
    DeviceGuard guard(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    StreamOrder stream_order(ctx, s);
    // random projections r (m words) and v (n words) in [1, p)
    std::vector<uint32_t> rv(m + n);
    for (auto &x : rv) x = (uint32_t)(1 + splitmix64(ctx->check_state) % (p - 1));
    Staged stA, stB, stC;
    lbx_status st = stage_in(ctx, "check_a", a, 4, true, &stA, s);
    if (st == LBX_OK) st = stage_in(ctx, "check_b", b, 4, true, &stB, s);
    if (st == LBX_OK) st = stage_in(ctx, "check_c", c, 4, true, &stC, s);
    void *ws = nullptr;
    const size_t words = (m + n) + k * sa + k * sb + count;
    if (st == LBX_OK) st = ensure(ctx, "check_eval", words * 4 + 256, &ws);
    unsigned nbad = 0;
    if (st == LBX_OK) {
        uint32_t *r = (uint32_t *)ws, *v = r + m, *a1 = v + n, *b1 = a1 + k * sa, *c1 = b1 + k * sb;
        unsigned *bad = (unsigned *)(c1 + count);
        const uint64_t mu = ~(uint64_t)0 / p;
        st = host_to_device(ctx, r, rv.data(), rv.size() * 4, s);
        if (st == LBX_OK) {
            hipLaunchKernelGGL(combine_rows_kernel, dim3(blocks_for(k * sa)), dim3(256), 0, s, devview(a, stA.dev), r,
                               sa, p, mu, a1);
            hipLaunchKernelGGL(combine_cols_kernel, dim3(blocks_for(k * sb)), dim3(256), 0, s, devview(b, stB.dev), v,
                               sb, p, mu, b1);
            hipLaunchKernelGGL(combine_both_kernel, dim3(blocks_for(count)), dim3(256), 0, s, devview(c, stC.dev), r,
                               v, count, p, mu, c1);
            st = from_hip(hipMemsetAsync(bad, 0, 4, s));
        }
        if (st == LBX_OK) {
            hipLaunchKernelGGL(midcheck_kernel, dim3((unsigned)(count < 65536 ? count : 65536)), dim3(256), 0, s, a1,
                               sa, b1, sb, k, c1, hdeg, count, p, mu, bad);
            st = from_hip(hipGetLastError());
        }
        if (st == LBX_OK) st = from_hip(hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, s));
        if (st == LBX_OK) st = from_hip(hipStreamSynchronize(s));
    }
    release(stA);
    release(stB);
    release(stC);
    if (st == LBX_OK) *equal = nbad == 0;
    return st;
}

lbx_status self_check_midproduct(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a,
                                 const lbx_polmat *b, size_t hdeg, size_t count, void *stream)
{
    if (!ctx->check) return LBX_OK;
    int equal = 0;
    LBX_TRY(check_midproduct(ctx, p, c, a, b, hdeg, count, &equal, stream));
    return equal ? LBX_OK : LBX_ERR_CHECK_FAILED;
}

// The check lbx_polymat_mul_u32 runs when the context asks for it: c holds
// the whole product (deg+1 >= a.size+b.size-1), checked at a random point;
// truncated products are left unchecked.
lbx_status self_check(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                      void *stream)
{
    if (!ctx->check || c->size + 1 < a->size + b->size) return LBX_OK;
    uint32_t x;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        x = (uint32_t)(splitmix64(ctx->check_state) % p);
    }
    int equal = 0;
    LBX_TRY(check_product(ctx, p, c, a, b, c->size, x, &equal, stream));
    return equal ? LBX_OK : LBX_ERR_CHECK_FAILED;
}

}  // namespace lbx

using namespace lbx;

extern "C" lbx_status lbx_polymat_check_mul_u32(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a,
                                                const lbx_polmat *b, size_t deg, uint32_t x, int *equal,
                                                void *stream)
{
    return guarded([&]() -> lbx_status {
        return check_product(ctx, p, c, a, b, deg, x, equal, stream);
    });
}

extern "C" lbx_status lbx_polymat_check_midproduct_u32(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c,
                                                       const lbx_polmat *a, const lbx_polmat *b, size_t n0, size_t n1,
                                                       int *equal, void *stream)
{
    return guarded([&]() -> lbx_status {
        if (!c) return LBX_ERR_INVALID_ARG;
        const size_t hdeg = n0 ? n0 : c->size;
        if (!hdeg) return LBX_ERR_INVALID_ARG;
        const size_t deg = n1 ? n1 : 2 * hdeg - 1;
        if (deg < hdeg) return LBX_ERR_SIZE_MISMATCH;
        const size_t count = std::min(c->size, deg - hdeg + 1);
        return check_midproduct(ctx, p, c, a, b, hdeg, count, equal, stream);
    });
}

extern "C" lbx_status lbx_set_check(lbx_ctx *ctx, int on)
{
    return guarded([&]() -> lbx_status {
        if (!ctx) return LBX_ERR_INVALID_ARG;
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->check = on != 0;
        return LBX_OK;
    });
}
