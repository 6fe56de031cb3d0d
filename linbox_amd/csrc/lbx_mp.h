// Multiprecision (CRT) route and batch scheduler internals.
#pragma once

#include "lbx/lbx.h"
#include "lbx_internal.h"

namespace lbx {
// Releases the device state the multiprecision route caches in a context.
void lbx_mp_release(lbx_ctx *ctx);
// Single-prime product on device views (lbx_api.hip); deg+1 = count.
lbx_status fft_prime_product(lbx_ctx *ctx, uint32_t p, unsigned L, const DevView &a, const DevView &b,
                             const DevView &c, size_t m, size_t k, size_t n, size_t count, int method,
                             hipStream_t s);
}  // namespace lbx
