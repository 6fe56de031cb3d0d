// Context and helpers shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "lbx/lbx.h"
#include "lbx_internal.h"

namespace lbx {

struct Buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct MpPlan;     // cached CRT constants of the multiprecision route (mp.hip)
struct BatchPlan;  // cached item lists and graph of a batch shape (batch.hip)

}  // namespace lbx

struct lbx_ctx {
    int device = 0;
    std::mutex mu;
    std::map<std::string, lbx::Buffer> bufs;
    std::map<std::tuple<uint32_t, unsigned, uint32_t>, uint2 *> twiddles;
    std::map<std::string, lbx::MpPlan *> mp_plans;
    std::map<std::string, lbx::BatchPlan *> batch_plans;
    int tile_tuned[3] = {-1, -1, -1};  // batch pointwise tile per size class, -1 = not tuned yet
    hipStream_t aux = nullptr;         // second stream for chunk pipelining (default priority)
    std::vector<hipEvent_t> events;    // pool of dependency events
    int profiling = 0;  // 0 off, 1 stages, 2 stages and kernels
    // recorded (label, start, end) event triples, resolved by lbx_profile_get
    std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> prof_pending;
    std::vector<std::string> prof_labels;
    std::vector<float> prof_ms;
    std::vector<const char *> prof_label_ptrs;
};

#define LBX_TRY_HIP(expr)                                 \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return ::lbx::from_hip(e_); \
    } while (0)
#define LBX_TRY(expr)                \
    do {                             \
        lbx_status s_ = (expr);      \
        if (s_ != LBX_OK) return s_; \
    } while (0)

namespace lbx {

lbx_status from_hip(hipError_t e);

class DeviceGuard {
public:
    explicit DeviceGuard(int dev)
    {
        (void)hipGetDevice(&prev_);
        if (prev_ != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() { (void)hipSetDevice(prev_); }

private:
    int prev_ = 0;
};

// Named device workspace, grown on demand and kept between calls.
lbx_status ensure(lbx_ctx *ctx, const char *name, size_t bytes, void **out);
bool is_device_ptr(const void *p);
size_t storage_words(const lbx_polmat *v);
lbx_status check_view(const lbx_polmat *v);
DevView devview(const lbx_polmat *v, void *dev_base);

// Host operands are staged through device copies of their whole storage
// (word_bytes per coefficient slot).
struct Staged {
    lbx_polmat v;
    void *dev = nullptr;
    bool owned = false;
};
lbx_status stage_in(const lbx_polmat *src, size_t word_bytes, bool copy_in, Staged *out);
void release(Staged &s);

// Midproduct variant of a transform product (matpoly-mult-fft-wordsize-fast.inl:209-316):
// the first hdeg coefficients of the small operand (a if small_left) are
// reversed, that operand is transformed with the root w and the other with
// w^-1, and the inverse transform uses w.  Requires small operand size <= hdeg.
struct MidSpec {
    bool small_left;
    size_t hdeg;
};

// Single-prime product on device views (count = deg+1 coefficients of c), or
// with mid != nullptr the midproduct (count = c.size coefficients).
lbx_status fft_prime_product(lbx_ctx *ctx, uint32_t p, unsigned L, const DevView &a, const DevView &b,
                             const DevView &c, size_t m, size_t k, size_t n, size_t count, int method,
                             hipStream_t s, const MidSpec *mid = nullptr);
// Product (mid == nullptr; c holds sa+sb-1 coefficients) or midproduct with
// 2^L points over Z/PZ (P of nl little-endian limbs) through FFT primes and
// CRT (matpoly-mult-fft-multiprecision.inl:141-277, 427-520).
lbx_status mp_product(lbx_ctx *ctx, const uint32_t *Pw, int nl, const DevView &a, const DevView &b, const DevView &c,
                      size_t m, size_t k, size_t n, int method, hipStream_t s, const MidSpec *mid = nullptr,
                      unsigned mid_L = 0);
// Product of constant matrices (one coefficient each).
lbx_status constant_product(lbx_ctx *ctx, uint32_t p, const DevView &a, const DevView &b, const DevView &c, size_t m,
                            size_t k, size_t n, hipStream_t s);
// Frees the multiprecision plans cached in a context.
void mp_release(lbx_ctx *ctx);
// Frees the batch plans (graphs, descriptors) cached in a context.
void batch_release(lbx_ctx *ctx);
// Independent polfirst products with device operands through the batch
// kernels, the context lock held; LBX_ERR_UNSUPPORTED (nothing launched) when a
// job is outside their range.
lbx_status batch_products(lbx_ctx *ctx, const lbx_job *jobs, size_t njobs, hipStream_t s);
// The pointwise form LBX_METHOD_AUTO resolves to for m x k x n at 2^L points.
int choose_method(int method, size_t m, size_t k, size_t n, unsigned L);

// Root of unity of order 2^k (fft-utils.h:83-104), its inverse (fft.h:119-130)
// and the cached device twiddle table of a root (fft-integral.inl:1614-1650).
lbx_status primitive_root(uint32_t p, unsigned k, uint32_t *w_out);
uint32_t invroot(uint32_t p, unsigned L, uint32_t w);
lbx_status twiddles(lbx_ctx *ctx, uint32_t p, unsigned L, uint32_t w, const uint2 **out);

}  // namespace lbx
