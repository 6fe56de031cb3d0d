// Check of a product at one point, on the device: the random-point form of
// the reference's check_mul (matpoly-mult-fft.h:87-163, which evaluates
// at 1), run after every lbx_polymat_mul_u32 when a context asks for it, as
// CHECK_MATPOL_MUL does after PolynomialMatrixMulDomain::mul
// (polynomial-matrix-domain.h:72-74).
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

extern "C" lbx_status lbx_set_check(lbx_ctx *ctx, int on)
{
    return guarded([&]() -> lbx_status {
        if (!ctx) return LBX_ERR_INVALID_ARG;
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->check = on != 0;
        return LBX_OK;
    });
}
