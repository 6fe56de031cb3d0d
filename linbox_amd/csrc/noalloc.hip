// lbx_polymat_mul_u32_noalloc: the single-prime and constant-matrix routes of
// lbx_polymat_mul_u32 with every device buffer, twiddle tables included, cut
// from one fixed device array instead of the context's workspace, so a call
// makes no device allocation (for callers under graph capture or in loops
// that must not allocate).  The array holds kNoAllocBytes per device for the
// life of the process; every call on a device uses the same array, so calls
// that overlap on one device race, and each call rebuilds what it needs in it
// (no kernel relies on its previous contents).
#include <hip/hip_runtime.h>

#include <mutex>

#include "lbx/lbx.h"
#include "lbx_ctx.h"

namespace lbx {
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
}  // namespace lbx

// Host code must not use this name as an address: it is the host-side
// stand-in of the device array.  The device address of each device's copy is
// taken once with hipGetSymbolAddress (noalloc_attach) and used instead.
static __device__ __align__(256) uint8_t scratch[lbx::kNoAllocBytes];

namespace lbx {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

}  // namespace

// Called by lbx_init with `device` current: records the address of this
// device's copy of the array (once per device).
lbx_status noalloc_attach(int device)
{
    if (device < 0 || device >= kMaxDevices) return LBX_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return LBX_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return LBX_ERR_HIP;
    }
    g_scratch[device] = ptr;
    return LBX_OK;
}

}  // namespace lbx

using namespace lbx;

extern "C" lbx_status lbx_polymat_mul_u32_noalloc(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a,
                                                  const lbx_polmat *b, size_t max_rowdeg, int method, void *stream)
{
    return guarded([&]() -> lbx_status {
        if (!ctx || !a || !b || !c) return LBX_ERR_INVALID_ARG;
        LBX_TRY(check_view(a));
        LBX_TRY(check_view(b));
        LBX_TRY(check_view(c));
        if (method < 0 || method > 2) return LBX_ERR_INVALID_ARG;
        if (p < 2) return LBX_ERR_INVALID_ARG;
        if (a->cols != b->rows || c->rows != a->rows || c->cols != b->cols) return LBX_ERR_SIZE_MISMATCH;
        if (!a->size || !b->size) return LBX_ERR_INVALID_ARG;
        const size_t m = a->rows, k = a->cols, n = b->cols, sa = a->size, sb = b->size;
        const size_t deg = max_rowdeg ? max_rowdeg : sa + sb - 2;
        if (c->size < deg + 1) return LBX_ERR_SIZE_MISMATCH;
        unsigned L = 0;
        size_t pts = 1;
        while (pts <= deg) {
            pts <<= 1;
            ++L;
        }
        if (sa > pts || sb > pts) return LBX_ERR_SIZE_MISMATCH;
        const bool small_p = p < (1u << 31);
        const bool constant = sa + sb <= 2 && small_p;
        const bool single_prime = !constant && small_p && (p - 1) % pts == 0;
        // the few-primes and multiprecision routes keep per-modulus plans in the context
        if (!constant && !single_prime) return LBX_ERR_UNSUPPORTED;
        std::lock_guard<std::mutex> lock(ctx->mu);
        DeviceGuard guard(ctx->device);
        // device memory only: host operands would need staging buffers
        if (!is_device_ptr(a->data) || !is_device_ptr(b->data) || !is_device_ptr(c->data)) return LBX_ERR_INVALID_ARG;
        void *base = nullptr;
        {
            std::lock_guard<std::mutex> g(g_mu);
            base = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
        }
        if (!base) return LBX_ERR_HIP;
        ArenaWorkspace ws(base, kNoAllocBytes);
        hipStream_t s = (hipStream_t)stream;
        const DevView da = devview(a, a->data), db = devview(b, b->data), dc = devview(c, c->data);
        if (constant) {
            LBX_TRY(constant_product(ctx, p, da, db, dc, m, k, n, s, &ws));
            if (deg > 0) LBX_TRY_HIP(launch_fill(dc, m * n, 1, deg, 0u, s));
        } else {
            LBX_TRY(fft_prime_product(ctx, p, L, da, db, dc, m, k, n, deg + 1, method, s, nullptr, &ws));
        }
        c->size = deg + 1;
        return LBX_OK;
    });
}
