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

struct MpPlan;      // cached CRT constants of the multiprecision route (mp.hip)
struct BatchPlan;   // cached item lists and graph of a batch shape (batch.hip)
struct HostStager;  // pinned buffers and copy threads for host operands (staging.hip)

}  // namespace lbx

struct lbx_ctx {
    int device = 0;
    std::mutex mu;
    std::map<std::string, lbx::Buffer> bufs;
    std::map<std::tuple<uint32_t, unsigned, uint32_t, int>, uint2 *> twiddles;
    std::map<std::string, lbx::MpPlan *> mp_plans;
    std::map<std::string, lbx::BatchPlan *> batch_plans;
    int tile_tuned[3] = {-1, -1, -1};  // batch pointwise tile per size class, -1 = not tuned yet
    // batch INT8 choice per class of INT8-capable jobs (min(m,k,n) in [16,32) / >= 32):
    // form (0 plain, 1 INT8 MFMA) and points per MFMA item, -1 = not tuned yet
    int form_tuned[2] = {-1, -1};
    int ppi_tuned[2] = {-1, -1};
    // last tuning measurements (ms) of those classes: plain, best INT8
    float form_ms[2][2] = {{0, 0}, {0, 0}};
    std::string tuning_json;  // lbx_batch_tuning's last answer
    hipStream_t aux = nullptr;         // second stream for chunk pipelining (default priority)
    std::vector<hipEvent_t> events;    // pool of dependency events
    // stream order of the shared workspace: the stream of the last call and
    // an event recorded at its end (see StreamOrder)
    hipStream_t last_stream = nullptr;
    hipEvent_t last_done = nullptr;
    lbx::HostStager *stager = nullptr;  // created by the first large host copy
    int profiling = 0;  // 0 off, 1 stages, 2 stages and kernels
    bool check = false;           // check every whole product at a random point (lbx_set_check)
    uint64_t check_state = 0x4C42582D43484B31ull;  // point generator state
    // recorded (label, start, end) event triples, resolved by lbx_profile_get
    std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> prof_pending;
    std::vector<std::string> prof_labels;
    std::vector<float> prof_ms;
    std::vector<const char *> prof_label_ptrs;
};

namespace lbx {
// Runs the body of a C ABI entry point: no C++ exception may reach a C caller.
// The library throws nothing itself; the standard library throws only when an
// allocation fails (bad_alloc, length_error), which is reported as such.
template <class F>
lbx_status guarded(F &&body) noexcept
{
    try {
        return body();
    } catch (...) {
        return LBX_ERR_OUT_OF_MEMORY;
    }
}
}  // namespace lbx

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

// Every call on a context uses the context's one set of workspace buffers,
// batch plans and events.  A call on a stream other than the previous call's
// first makes its stream wait for the previous call's work (an event recorded
// on that stream at the end of the call), so two streams never use the
// workspace at the same time.  Construct after taking ctx->mu and selecting the
// device; the destructor records the event.  Not used by calls on a stream
// under graph capture (a capture cannot wait for work outside its graph).
class StreamOrder {
public:
    StreamOrder(lbx_ctx *ctx, hipStream_t s) : ctx_(ctx), s_(s)
    {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
            (void)hipGetLastError();
            cs = hipStreamCaptureStatusNone;
        }
        active_ = cs == hipStreamCaptureStatusNone;
        if (active_ && ctx->last_done && ctx->last_stream != s) (void)hipStreamWaitEvent(s, ctx->last_done, 0);
    }
    ~StreamOrder()
    {
        if (!active_) return;
        if (!ctx_->last_done && hipEventCreateWithFlags(&ctx_->last_done, hipEventDisableTiming) != hipSuccess) {
            ctx_->last_done = nullptr;
            return;
        }
        if (hipEventRecord(ctx_->last_done, s_) == hipSuccess) ctx_->last_stream = s_;
    }
    StreamOrder(const StreamOrder &) = delete;
    StreamOrder &operator=(const StreamOrder &) = delete;

private:
    lbx_ctx *ctx_;
    hipStream_t s_;
    bool active_ = false;
};

// Root of unity of order 2^k (fft-utils.h:83-104), its inverse (fft.h:119-130)
// and the cached device twiddle table of a root (fft-integral.inl:1614-1650)
// in the format of the transforms that read it.
lbx_status primitive_root(uint32_t p, unsigned k, uint32_t *w_out);
uint32_t invroot(uint32_t p, unsigned L, uint32_t w);
lbx_status twiddles(lbx_ctx *ctx, uint32_t p, unsigned L, uint32_t w, TwFormat fmt, const uint2 **out);

// Named device workspace, grown on demand and kept between calls.
lbx_status ensure(lbx_ctx *ctx, const char *name, size_t bytes, void **out);
bool is_device_ptr(const void *p);
size_t storage_words(const lbx_polmat *v);
lbx_status check_view(const lbx_polmat *v);
DevView devview(const lbx_polmat *v, void *dev_base);

// Host operands are staged through device copies of their whole storage
// (word_bytes per coefficient slot), in the context's workspace buffer `slot`
// (kept between calls like the rest of the workspace).
struct Staged {
    lbx_polmat v;
    void *dev = nullptr;
    bool owned = false;  // dev is a staging copy of host storage
};
// With copy_in the host storage is copied on stream s (queued ahead of the
// call's kernels; the host memory may be reused on return).
lbx_status stage_in(lbx_ctx *ctx, const char *slot, const lbx_polmat *src, size_t word_bytes, bool copy_in,
                    Staged *out, hipStream_t s);
void release(Staged &s);
// bytes from host to device memory on stream s (returns once the host memory
// has been read), and from device to host memory after the work queued on s
// (returns once the host memory holds them); large copies go through the
// context's pinned buffers.
lbx_status host_to_device(lbx_ctx *ctx, void *dev, const void *host, size_t bytes, hipStream_t s);
lbx_status device_to_host(lbx_ctx *ctx, void *host, const void *dev, size_t bytes, hipStream_t s);
void stager_release(lbx_ctx *ctx);
// c = a.b at the point x over c's first deg coefficients (*equal); the
// context's random-point check after a product (LBX_ERR_CHECK_FAILED)
lbx_status check_product(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                         size_t deg, uint32_t x, int *equal, void *stream);
lbx_status self_check(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                      void *stream);
// count coefficients of c against coefficients hdeg-1.. of a.b (zero beyond
// its degree), by random projections of rows and columns (*equal); the
// context's check after a middle product (LBX_ERR_CHECK_FAILED)
lbx_status check_midproduct(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                            size_t hdeg, size_t count, int *equal, void *stream);
lbx_status self_check_midproduct(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a,
                                 const lbx_polmat *b, size_t hdeg, size_t count, void *stream);

// Where one call's device buffers come from.  The context's workspace keeps
// named buffers between calls (ensure) and caches twiddle tables; an arena
// hands out consecutive pieces of memory the caller owns and builds the
// twiddle tables in it on the device, so a call through it allocates nothing.
// Every buffer a product needs is requested before its first launch, so an
// arena that is too small fails before any work is queued.
class Workspace {
public:
    virtual ~Workspace() = default;
    virtual lbx_status get(const char *name, size_t bytes, void **out) = 0;
    virtual lbx_status twiddle_table(uint32_t p, unsigned L, uint32_t w, TwFormat fmt, const uint2 **out,
                                     hipStream_t s) = 0;
    // chunked head/tail pipelining on a second stream (and its events)
    virtual bool may_pipeline() const = 0;
    // stage and kernel timing events (lbx_profile_enable)
    virtual bool timing() const = 0;
};

class CtxWorkspace final : public Workspace {
public:
    explicit CtxWorkspace(lbx_ctx *ctx) : ctx_(ctx) {}
    lbx_status get(const char *name, size_t bytes, void **out) override { return ensure(ctx_, name, bytes, out); }
    lbx_status twiddle_table(uint32_t p, unsigned L, uint32_t w, TwFormat fmt, const uint2 **out, hipStream_t) override
    {
        return twiddles(ctx_, p, L, w, fmt, out);
    }
    bool may_pipeline() const override { return true; }
    bool timing() const override { return ctx_->profiling != 0; }

private:
    lbx_ctx *ctx_;
};

class ArenaWorkspace final : public Workspace {
public:
    ArenaWorkspace(void *base, size_t cap) : base_((uint8_t *)base), cap_(cap) {}
    lbx_status get(const char *, size_t bytes, void **out) override;
    lbx_status twiddle_table(uint32_t p, unsigned L, uint32_t w, TwFormat fmt, const uint2 **out,
                             hipStream_t s) override;
    bool may_pipeline() const override { return false; }
    bool timing() const override { return false; }
    size_t used() const { return used_; }

private:
    uint8_t *base_;
    size_t cap_, used_ = 0;
};

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
// Buffers come from `ws` (the context's workspace when null).
lbx_status fft_prime_product(lbx_ctx *ctx, uint32_t p, unsigned L, const DevView &a, const DevView &b,
                             const DevView &c, size_t m, size_t k, size_t n, size_t count, int method,
                             hipStream_t s, const MidSpec *mid = nullptr, Workspace *ws = nullptr);
// Product (mid == nullptr; c holds sa+sb-1 coefficients) or midproduct with
// 2^L points over Z/PZ (P of nl little-endian limbs) through FFT primes and
// CRT (matpoly-mult-fft-multiprecision.inl:141-277, 427-520).
// A product writes `count` coefficients of c (0: sa+sb-1; fewer: the cyclic
// product mod x^pts - 1 with pts the least power of two >= count, truncated,
// as for max_rowdeg).
lbx_status mp_product(lbx_ctx *ctx, const uint32_t *Pw, int nl, const DevView &a, const DevView &b, const DevView &c,
                      size_t m, size_t k, size_t n, int method, hipStream_t s, const MidSpec *mid = nullptr,
                      unsigned mid_L = 0, size_t count = 0);
// Product of constant matrices (one coefficient each).
lbx_status constant_product(lbx_ctx *ctx, uint32_t p, const DevView &a, const DevView &b, const DevView &c, size_t m,
                            size_t k, size_t n, hipStream_t s, Workspace *ws = nullptr);
// Records the device address of the fixed workspace of
// lbx_polymat_mul_u32_noalloc on `device` (current); once per device.
lbx_status noalloc_attach(int device);
// Frees the multiprecision plans cached in a context.
void mp_release(lbx_ctx *ctx);
// Frees the batch plans (graphs, descriptors) cached in a context.
void batch_release(lbx_ctx *ctx);
// Independent polfirst products with device operands through the batch
// kernels, the context lock held; LBX_ERR_UNSUPPORTED (nothing launched) when a
// job is outside their range.
lbx_status batch_products(lbx_ctx *ctx, const lbx_job *jobs, size_t njobs, hipStream_t s);
// Whether the plain form takes point-major operands (split transforms) for m x k x n.
bool plain_point_major(size_t m, size_t k, size_t n);
// The pointwise form LBX_METHOD_AUTO resolves to for m x k x n at 2^L points.
int choose_method(int method, size_t m, size_t k, size_t n, unsigned L);


}  // namespace lbx
