// C ABI (include/lbx/lbx.h): argument checking, dispatch between the
// single-prime, constant-matrix and multiprecision routes, device workspace,
// twiddle-table cache, host/device pointer handling and stage timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "lbx/lbx.h"
#include "lbx_ctx.h"

namespace lbx {

lbx_status from_hip(hipError_t e)
{
    if (e == hipSuccess) return LBX_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return LBX_ERR_OUT_OF_MEMORY;
    return LBX_ERR_HIP;
}


lbx_status ensure(lbx_ctx *ctx, const char *name, size_t bytes, void **out)
{
    Buffer &b = ctx->bufs[name];
    if (b.bytes < bytes) {
        if (b.ptr) {
            (void)hipDeviceSynchronize();
            (void)hipFree(b.ptr);
            b.ptr = nullptr;
            b.bytes = 0;
        }
        LBX_TRY_HIP(hipMalloc(&b.ptr, bytes ? bytes : 16));
        b.bytes = bytes;
    }
    *out = b.ptr;
    return LBX_OK;
}

bool is_device_ptr(const void *p)
{
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice;
}

unsigned v2(uint32_t x)
{
    unsigned v = 0;
    while (x && !(x & 1)) {
        x >>= 1;
        ++v;
    }
    return v;
}

// fft-utils.h:65-78
bool is_primitive_root(uint32_t p, uint32_t w, unsigned k)
{
    uint64_t x = w;
    for (unsigned i = 0; i < k; ++i) {
        if (x == 1) return false;
        x = x * x % p;
    }
    return true;
}

// fft-utils.h:83-104
lbx_status primitive_root(uint32_t p, unsigned k, uint32_t *w_out)
{
    const unsigned val2p = v2(p - 1);
    const uint32_t m = (p - 1) >> val2p;
    if (k > val2p) return LBX_ERR_K_TOO_LARGE;
    for (uint32_t t = 2;; ++t) {
        uint64_t w = powmod_h(t, m, p);
        for (unsigned i = 0; i < val2p - k; ++i) w = w * w % p;
        if (is_primitive_root(p, (uint32_t)w, k)) {
            *w_out = (uint32_t)w;
            return LBX_OK;
        }
    }
}

// fft.h:119-130
uint32_t invroot(uint32_t p, unsigned L, uint32_t w)
{
    if (L == 1) return p - 1;
    const uint32_t x = powmod_h(w, ((uint64_t)1 << (L - 1)) - 1, p);
    return x ? p - x : 0;
}

// Twiddle table in the pow_w layout of init_powers (fft-integral.inl:1614-1650),
// cached per (p, L, root, format).
lbx_status twiddles(lbx_ctx *ctx, uint32_t p, unsigned L, uint32_t w, TwFormat fmt, const uint2 **out)
{
    auto key = std::make_tuple(p, L, w, (int)fmt);
    auto it = ctx->twiddles.find(key);
    if (it != ctx->twiddles.end()) {
        *out = it->second;
        return LBX_OK;
    }
    const size_t n = (size_t)1 << L;
    std::vector<uint32_t> pw(n > 1 ? n - 1 : 1);
    pw[0] = 1 % p;
    for (size_t i = 1; i < n / 2; ++i) pw[i] = (uint32_t)((uint64_t)pw[i - 1] * w % p);
    size_t idx = n / 2;
    for (size_t kk = 2; kk <= n / 2; kk <<= 1)
        for (size_t i = 0; i < n / 2; i += kk, idx++) pw[idx] = pw[i];
    std::vector<uint2> tw(pw.size());
    const uint32_t pinv = inv_mod32(p);
    for (size_t i = 0; i < pw.size(); ++i) {
        if (fmt == kTwMont) {
            const uint32_t wm = mont_wm(pw[i], p);
            tw[i] = make_uint2(wm, mont_wq(wm, pinv));
        } else {
            tw[i] = make_uint2(pw[i], shoup_pre(pw[i], p));
        }
    }
    uint2 *d = nullptr;
    LBX_TRY_HIP(hipMalloc(&d, sizeof(uint2) * tw.size()));
    LBX_TRY_HIP(hipMemcpy(d, tw.data(), sizeof(uint2) * tw.size(), hipMemcpyHostToDevice));
    ctx->twiddles[key] = d;
    *out = d;
    return LBX_OK;
}

size_t storage_words(const lbx_polmat *v) { return v->rows * v->cols * v->storage_size; }

lbx_status check_view(const lbx_polmat *v)
{
    if (!v || !v->data) return LBX_ERR_INVALID_ARG;
    if (v->layout < 0 || v->layout > 2) return LBX_ERR_INVALID_ARG;
    if (!v->rows || !v->cols) return LBX_ERR_INVALID_ARG;
    if (v->shift + v->size > v->storage_size) return LBX_ERR_SIZE_MISMATCH;
    return LBX_OK;
}

DevView devview(const lbx_polmat *v, void *dev_base)
{
    DevView d;
    d.base = (uint32_t *)dev_base;
    d.layout = v->layout;
    d.rows = v->rows;
    d.cols = v->cols;
    d.stor = v->storage_size;
    d.shift = v->shift;
    d.size = v->size;
    return d;
}

// Stage timer for the FFT_PROFILER analogue: events are recorded on the
// stream between stages and resolved later, so timing adds no host sync.
class Stages {
public:
    Stages(lbx_ctx *ctx, hipStream_t s, bool enabled) : ctx_(ctx), s_(s), on_(enabled && ctx->profiling)
    {
        if (on_) last_ = record();
    }
    void mark(const char *label_of_finished_stage)
    {
        if (!on_) return;
        hipEvent_t ev = record();
        ctx_->prof_pending.emplace_back(label_of_finished_stage, last_, ev);
        last_ = ev;
    }

private:
    hipEvent_t record()
    {
        hipEvent_t ev = nullptr;
        (void)hipEventCreate(&ev);
        (void)hipEventRecord(ev, s_);
        return ev;
    }
    lbx_ctx *ctx_;
    hipStream_t s_;
    bool on_;
    hipEvent_t last_ = nullptr;
};

// Kernel timer (profiling level 2): events around one launch on the stream it
// is queued on, so concurrent launches on other streams do not blur it.
class KernelTimer {
public:
    KernelTimer(lbx_ctx *ctx, hipStream_t s, const char *label, bool enabled) : ctx_(ctx), s_(s), label_(label)
    {
        if (enabled && ctx_->profiling >= 2 && hipEventCreate(&start_) == hipSuccess)
            (void)hipEventRecord(start_, s_);
    }
    ~KernelTimer()
    {
        if (!start_) return;
        hipEvent_t end = nullptr;
        (void)hipEventCreate(&end);
        (void)hipEventRecord(end, s_);
        ctx_->prof_pending.emplace_back(label_, start_, end);
    }

private:
    lbx_ctx *ctx_;
    hipStream_t s_;
    const char *label_;
    hipEvent_t start_ = nullptr;
};

// AUTO picks the INT8 form where it measured faster on gfx950 (both forms are
// bit-exact): for large products (m*k*n >= 2^18) and for short transforms
// (at most 2^9 points); in between, the split transforms the INT8 form needs
// cost more than the VALU pointwise they replace (16^3 at 2^14 points: 0.25 ms
// INT8, 0.08 ms plain; 32^3 at 2^13: 0.20 vs 0.13 ms; 64^3 at 2^11: 0.14 vs
// 0.16 ms; 16^3 at 2^9: 0.032 vs 0.036 ms).  MFMA tiles are 32x32 with a
// 64-deep inner step, so every dimension must be at least 16.
int choose_method(int method, size_t m, size_t k, size_t n, unsigned L)
{
    if (method == LBX_METHOD_PLAIN_U64 || method == LBX_METHOD_INT8_MFMA) return method;
    if (m < 16 || n < 16 || k < 16) return LBX_METHOD_PLAIN_U64;
    const bool big = (double)m * k * n >= (double)(1u << 18) || L <= 9;
    return big ? LBX_METHOD_INT8_MFMA : LBX_METHOD_PLAIN_U64;
}

// The plain form's operand layout: point-major blocks from the split
// transforms (each operand word read once per output tile) for large matrices,
// polfirst transform output (one point per lane, operands re-read once per
// 8-wide register tile) otherwise.  Measured on gfx950 (2013265921, whole
// products): 128^3 at 2^14 points 5.58 ms point-major against 7.23 ms; 128^3 at
// 2^11 0.865 against 0.863 ms; 64^3 at 2^13 0.749 against 0.679 ms; 32^3 at 2^13
// 0.289 against 0.192 ms -- the split transforms cost more than the re-reads
// below ~100 rows.  LBX_PLAIN_LAYOUT = "point" or "entry" overrides the choice
// (measurement and tests).
bool plain_point_major(size_t m, size_t k, size_t n)
{
    const char *e = std::getenv("LBX_PLAIN_LAYOUT");
    if (e && !std::strcmp(e, "point")) return true;
    if (e && !std::strcmp(e, "entry")) return false;
    return m >= 96 && n >= 96 && k >= 64;
}

// Bytes of head -> tail intermediate per chunk.  The default, 2 GiB, gives one
// chunk per operand up to 128x128 entries of 2^14 points: smaller chunks
// measured slower there (the hand-off through the last-level cache gains less
// than the extra launches and partial waves cost).  The LBX_CHUNK_BYTES
// environment variable overrides it (tuning and tests).
static size_t chunk_bytes()
{
    const char *e = std::getenv("LBX_CHUNK_BYTES");
    if (e && *e) {
        const long long v = std::atoll(e);
        if (v >= 4096) return (size_t)v;
    }
    return (size_t)2 << 30;
}

// Whether several chunks of one direction overlap over two streams (the
// default) or run one after the other on the caller's stream
// (LBX_CHUNK_PIPE=0: tuning and tests).
static bool chunk_pipeline()
{
    const char *e = std::getenv("LBX_CHUNK_PIPE");
    return !(e && !std::strcmp(e, "0"));
}

// Two-stream chunk pipeline: chunk c's producer runs on the caller's stream s,
// its consumer on ctx->aux after the producer's event; the producer of chunk
// c+2 (same scratch slot) waits for the consumer of chunk c.  Events 2c and
// 2c+1 of the pool mark "produced" and "consumed".
static lbx_status event_at(lbx_ctx *ctx, size_t i, hipEvent_t *ev)
{
    while (ctx->events.size() <= i) {
        hipEvent_t e;
        LBX_TRY_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->events.push_back(e);
    }
    *ev = ctx->events[i];
    return LBX_OK;
}

static lbx_status pipeline_begin(lbx_ctx *ctx, hipStream_t s)
{
    if (!ctx->aux) LBX_TRY_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    hipEvent_t ev;
    LBX_TRY(event_at(ctx, 0, &ev));  // fork: aux starts after everything queued on s
    LBX_TRY_HIP(hipEventRecord(ev, s));
    LBX_TRY_HIP(hipStreamWaitEvent(ctx->aux, ev, 0));
    return LBX_OK;
}

static lbx_status pipeline_reuse(lbx_ctx *ctx, hipStream_t s, size_t chunk)
{
    if (chunk < 2) return LBX_OK;
    hipEvent_t ev;
    LBX_TRY(event_at(ctx, 2 + 2 * (chunk - 2) + 1, &ev));
    LBX_TRY_HIP(hipStreamWaitEvent(s, ev, 0));
    return LBX_OK;
}

static lbx_status pipeline_handoff(lbx_ctx *ctx, hipStream_t s, size_t chunk)
{
    hipEvent_t ev;
    LBX_TRY(event_at(ctx, 2 + 2 * chunk, &ev));
    LBX_TRY_HIP(hipEventRecord(ev, s));
    LBX_TRY_HIP(hipStreamWaitEvent(ctx->aux, ev, 0));
    return LBX_OK;
}

static lbx_status pipeline_done(lbx_ctx *ctx, size_t chunk)
{
    hipEvent_t ev;
    LBX_TRY(event_at(ctx, 2 + 2 * chunk + 1, &ev));
    LBX_TRY_HIP(hipEventRecord(ev, ctx->aux));
    return LBX_OK;
}

static lbx_status pipeline_end(lbx_ctx *ctx, hipStream_t s, size_t nchunks)
{
    if (!nchunks) return LBX_OK;
    hipEvent_t ev;  // join: s continues after the last consumer (aux is in order)
    LBX_TRY(event_at(ctx, 2 + 2 * (nchunks - 1) + 1, &ev));
    LBX_TRY_HIP(hipStreamWaitEvent(s, ev, 0));
    return LBX_OK;
}

// The pipeline above when `on`; otherwise (a single chunk per operand, where
// there is nothing to overlap) producer and consumer both run on s in order.
struct Pipe {
    lbx_ctx *ctx;
    hipStream_t s;
    bool on;
    hipStream_t consumer() const { return on ? ctx->aux : s; }
    lbx_status begin() { return on ? pipeline_begin(ctx, s) : LBX_OK; }
    lbx_status reuse(size_t c) { return on ? pipeline_reuse(ctx, s, c) : LBX_OK; }
    lbx_status handoff(size_t c) { return on ? pipeline_handoff(ctx, s, c) : LBX_OK; }
    lbx_status done(size_t c) { return on ? pipeline_done(ctx, c) : LBX_OK; }
    lbx_status end(size_t n) { return on ? pipeline_end(ctx, s, n) : LBX_OK; }
};

// Forward transforms, pointwise products and inverse transform of one
// single-prime product on device views.  deg+1 coefficients of c are written.
// Every device buffer is taken from the workspace before the first launch.
lbx_status fft_prime_product(lbx_ctx *ctx, uint32_t p, unsigned L, const DevView &a_in, const DevView &b_in,
                             const DevView &c, size_t m, size_t k, size_t n, size_t count, int method,
                             hipStream_t s, const MidSpec *mid, Workspace *ws_in)
{
    CtxWorkspace ctx_ws(ctx);
    Workspace &ws = ws_in ? *ws_in : ctx_ws;
    const size_t N = (size_t)1 << L;
    if (L > (unsigned)kMaxLog) return LBX_ERR_UNSUPPORTED;
    uint32_t w;
    LBX_TRY(primitive_root(p, L, &w));
    const uint32_t wi = invroot(p, L, w);
    // product: both operands forward with w, inverse with w^-1; midproduct
    // (wordsize-fast.inl:277-313): small operand with w, other with w^-1,
    // inverse with w, small operand reversed over its first hdeg coefficients.
    // Forward transforms read Montgomery tables, inverse ones Shoup tables.
    const uint2 *twf, *twc;
    LBX_TRY(ws.twiddle_table(p, L, w, kTwMont, &twf, s));
    LBX_TRY(ws.twiddle_table(p, L, mid ? w : wi, kTwShoup, &twc, s));
    DevView a = a_in, b = b_in;
    const uint2 *twa = twf, *twb = twf;
    void *rev = nullptr;
    if (mid) {
        const uint2 *twfi;
        LBX_TRY(ws.twiddle_table(p, L, wi, kTwMont, &twfi, s));
        (mid->small_left ? twb : twa) = twfi;
        const size_t se = mid->small_left ? m * k : k * n;
        LBX_TRY(ws.get("mid_rev", se * mid->hdeg * 4, &rev));
    }
    const uint32_t inv_n = invmod_h((uint32_t)(N % p), p);
    method = choose_method(method, m, k, n, L);
    const bool int8 = method == LBX_METHOD_INT8_MFMA;
    const bool split = L >= 8 && L <= (unsigned)kMaxLdsLog && (int8 || plain_point_major(m, k, n));

    // ---- buffers ----
    void *al = nullptr, *bl = nullptr, *cmat = nullptr, *scr = nullptr, *ahat = nullptr, *bhat = nullptr,
         *chat = nullptr;
    // split-transform chunking: entries go in chunks of chunk_bytes() of head -> tail
    // intermediate; with several chunks, heads run on the caller's stream and tails
    // on a second one, two chunks in flight
    const size_t mpad = (m + 31) & ~(size_t)31, npad = (n + 31) & ~(size_t)31;
    size_t rpc = 0, cchunk = 0, slot_words = 0, nslots = 1;
    bool fwd_pipe = false, inv_pipe = false;
    if (split) {
        const size_t target = std::max<size_t>(64, chunk_bytes() / (N * 4));  // entries per chunk
        rpc = std::max<size_t>(1, target / k);  // limb rows per chunk
        const size_t slot_entries = rpc * k;
        cchunk = std::max(slot_entries, target);  // product entries per inverse chunk
        if (!ws.may_pipeline()) {  // one chunk each way, so one slot and no second stream
            rpc = std::max(mpad, npad);
            cchunk = m * n;
        }
        const bool pipe_ok = chunk_pipeline();
        fwd_pipe = pipe_ok && ((mpad + rpc - 1) / rpc > 1 || (npad + rpc - 1) / rpc > 1);
        inv_pipe = pipe_ok && m * n > cchunk;
        // a slot holds the largest chunk either direction uses; one slot suffices
        // when nothing is pipelined
        slot_words = std::max(std::min(rpc, std::max(m, n)) * k, std::min(cchunk, m * n)) * N;
        nslots = (fwd_pipe || inv_pipe) ? 2 : 1;
        LBX_TRY(ws.get("alimb", limb_block_bytes(m, k) * N, &al));
        LBX_TRY(ws.get("blimb", limb_block_bytes(n, k) * N, &bl));
        LBX_TRY(ws.get("cmat", m * n * N * 4, &cmat));
        LBX_TRY(ws.get("scratch", nslots * slot_words * 4, &scr));
    } else {
        LBX_TRY(ws.get("ahat", m * k * N * 4, &ahat));
        LBX_TRY(ws.get("bhat", k * n * N * 4, &bhat));
        LBX_TRY(ws.get("chat", m * n * N * 4, &chat));
        if (method == LBX_METHOD_INT8_MFMA) {
            LBX_TRY(ws.get("alimb", limb_block_bytes(m, k) * N, &al));
            LBX_TRY(ws.get("blimb", limb_block_bytes(n, k) * N, &bl));
            LBX_TRY(ws.get("cmat", m * n * N * 4, &cmat));
        }
    }

    // ---- launches ----
    if (mid) {
        const DevView &sm = mid->small_left ? a_in : b_in;
        const size_t se = mid->small_left ? m * k : k * n;
        LBX_TRY_HIP(launch_reverse_gather(sm, se, mid->hdeg, (uint32_t *)rev, s));
        const DevView rv{(uint32_t *)rev, LBX_POLFIRST, sm.rows, sm.cols, mid->hdeg, 0, mid->hdeg};
        (mid->small_left ? a : b) = rv;
    }
    const bool timing = ws.timing();
    Stages st(ctx, s, timing);
    if (split) {
        // split transforms: the last six forward stages write per-point blocks
        // (digits for the INT8 form, words for the plain one) and the first
        // six inverse stages read the point-major product
        const unsigned BL = kSplitBL;
        uint32_t *slot[2] = {(uint32_t *)scr, (uint32_t *)scr + (nslots - 1) * slot_words};
        Pipe fp{ctx, s, fwd_pipe};
        LBX_TRY(fp.begin());
        size_t chunk = 0;
        // forward: rows of A (entries i*k+q), then columns of B (entries q*n+j)
        for (int op = 0; op < 2; ++op) {
            const size_t rows = op == 0 ? m : n, rpadx = op == 0 ? mpad : npad;
            for (size_t r0 = 0; r0 < rpadx; r0 += rpc, ++chunk) {
                const size_t nr = std::min(rpc, rpadx - r0);
                const size_t valid = r0 < rows ? std::min(nr, rows - r0) : 0;
                uint32_t *buf = slot[chunk & 1];
                LBX_TRY(fp.reuse(chunk));
                if (valid) {
                    const EntryMap map = op == 0 ? EntryMap{r0 * k, valid * k, 0} : EntryMap{r0, valid, n};
                    KernelTimer kt(ctx, s, op == 0 ? "kernel dif_head a" : "kernel dif_head b", timing);
                    LBX_TRY_HIP(launch_ntt_forward_head(op == 0 ? a : b, valid * k, map, L, BL, op == 0 ? twa : twb, p, buf, s));
                }
                LBX_TRY(fp.handoff(chunk));
                {
                    KernelTimer kt(ctx, fp.consumer(),
                                   int8 ? (op == 0 ? "kernel dif_tail_limbs a" : "kernel dif_tail_limbs b")
                                        : (op == 0 ? "kernel dif_tail_words a" : "kernel dif_tail_words b"),
                                   timing);
                    LBX_TRY_HIP(launch_dif_tail_limbs(buf, (uint8_t *)(op == 0 ? al : bl), rows, r0, nr, valid, k,
                                                      op == 0, L, BL, op == 0 ? twa : twb, p, fp.consumer(), int8));
                }
                LBX_TRY(fp.done(chunk));
            }
        }
        LBX_TRY(fp.end(chunk));
        st.mark("direct FFT_DIF");
        if (int8) {
            KernelTimer kt(ctx, s, "kernel pw_mfma", timing);
            LBX_TRY_HIP(launch_pointwise_mfma((const uint8_t *)al, (const uint8_t *)bl, (uint32_t *)cmat, m, k, n, N,
                                              p, inv_n, s));
        } else {
            KernelTimer kt(ctx, s, "kernel pw_plain_pm", timing);
            LBX_TRY_HIP(launch_pointwise_plain_pm((const uint32_t *)al, (const uint32_t *)bl, (uint32_t *)cmat, m, k,
                                                  n, N, p, inv_n, s));
        }
        st.mark("Pointwise mult");
        Pipe ip{ctx, s, inv_pipe};
        LBX_TRY(ip.begin());
        chunk = 0;
        for (size_t e0 = 0; e0 < m * n; e0 += cchunk, ++chunk) {
            const size_t ne = std::min(cchunk, m * n - e0);
            uint32_t *buf = slot[chunk & 1];
            LBX_TRY(ip.reuse(chunk));
            {
                KernelTimer kt(ctx, s, "kernel dit_head", timing);
                LBX_TRY_HIP(launch_dit_head((const uint32_t *)cmat, m * n, e0, ne, L, BL, twc, p, buf, s));
            }
            LBX_TRY(ip.handoff(chunk));
            {
                KernelTimer kt(ctx, ip.consumer(), "kernel dit_tail", timing);
                LBX_TRY_HIP(launch_ntt_inverse_tail(buf, ne, L, BL, twc, p, c, e0, count, 1, ip.consumer()));
            }
            LBX_TRY(ip.done(chunk));
        }
        LBX_TRY(ip.end(chunk));
        st.mark("inverse FFT_DIT");
        return LBX_OK;
    }
    {
        KernelTimer kt(ctx, s, "kernel dif a", timing);
        LBX_TRY_HIP(launch_ntt_forward(a, m * k, L, twa, p, (uint32_t *)ahat, s));
    }
    {
        KernelTimer kt(ctx, s, "kernel dif b", timing);
        LBX_TRY_HIP(launch_ntt_forward(b, k * n, L, twb, p, (uint32_t *)bhat, s));
    }
    st.mark("direct FFT_DIF");
    if (method == LBX_METHOD_INT8_MFMA) {
        LBX_TRY_HIP(launch_to_limbs((const uint32_t *)ahat, (uint8_t *)al, m, k, true, N, p, s));
        LBX_TRY_HIP(launch_to_limbs((const uint32_t *)bhat, (uint8_t *)bl, n, k, false, N, p, s));
        st.mark("Polfirst to Matfirst");
        LBX_TRY_HIP(launch_pointwise_mfma((const uint8_t *)al, (const uint8_t *)bl, (uint32_t *)cmat, m, k, n, N, p,
                                          inv_n, s));
        st.mark("Pointwise mult");
        LBX_TRY_HIP(launch_point_to_entry((const uint32_t *)cmat, (uint32_t *)chat, m * n, N, s));
        st.mark("Matfirst to Polfirst");
    } else {
        {
            KernelTimer kt(ctx, s, "kernel pw_plain", timing);
            LBX_TRY_HIP(launch_pointwise_plain((const uint32_t *)ahat, (const uint32_t *)bhat, (uint32_t *)chat, m, k,
                                               n, N, p, inv_n, s));
        }
        st.mark("Pointwise mult");
    }
    {
        KernelTimer kt(ctx, s, "kernel dit", timing);
        LBX_TRY_HIP(launch_ntt_inverse((uint32_t *)chat, m * n, L, twc, p, c, count, 1, s));
    }
    st.mark("inverse FFT_DIT");
    return LBX_OK;
}

// Constant matrices (a.size()+b.size() <= 2): one pointwise product.
lbx_status constant_product(lbx_ctx *ctx, uint32_t p, const DevView &a, const DevView &b, const DevView &c,
                            size_t m, size_t k, size_t n, hipStream_t s, Workspace *ws_in)
{
    CtxWorkspace ctx_ws(ctx);
    Workspace &ws = ws_in ? *ws_in : ctx_ws;
    void *ahat, *bhat, *chat;
    LBX_TRY(ws.get("ahat", m * k * 4, &ahat));
    LBX_TRY(ws.get("bhat", k * n * 4, &bhat));
    LBX_TRY(ws.get("chat", m * n * 4, &chat));
    LBX_TRY_HIP(launch_gather(a, m * k, 1, (uint32_t *)ahat, 1, s));
    LBX_TRY_HIP(launch_gather(b, k * n, 1, (uint32_t *)bhat, 1, s));
    LBX_TRY_HIP(launch_pointwise_plain((const uint32_t *)ahat, (const uint32_t *)bhat, (uint32_t *)chat, m, k, n, 1,
                                       p, 1, s));
    LBX_TRY_HIP(launch_scatter((const uint32_t *)chat, 1, m * n, 1, c, s));
    return LBX_OK;
}

lbx_status ArenaWorkspace::get(const char *, size_t bytes, void **out)
{
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (need > cap_ - used_) return LBX_ERR_WORKSPACE_TOO_SMALL;
    *out = base_ + used_;
    used_ += need ? need : 256;
    return LBX_OK;
}

lbx_status ArenaWorkspace::twiddle_table(uint32_t p, unsigned L, uint32_t w, TwFormat fmt, const uint2 **out,
                                         hipStream_t s)
{
    const size_t n = (size_t)1 << L;
    void *d;
    LBX_TRY(get("twiddles", sizeof(uint2) * (n > 1 ? n - 1 : 1), &d));
    LBX_TRY_HIP(launch_build_twiddles((uint2 *)d, p, L, w, fmt, s));
    *out = (const uint2 *)d;
    return LBX_OK;
}

lbx_status stage_in(lbx_ctx *ctx, const char *slot, const lbx_polmat *src, size_t word_bytes, bool copy_in,
                    Staged *out, hipStream_t s)
{
    out->v = *src;
    if (is_device_ptr(src->data)) {
        out->dev = src->data;
        return LBX_OK;
    }
    const size_t bytes = storage_words(src) * word_bytes;
    LBX_TRY(ensure(ctx, slot, bytes, &out->dev));
    out->owned = true;
    if (copy_in) LBX_TRY(host_to_device(ctx, out->dev, src->data, bytes, s));
    return LBX_OK;
}

void release(Staged &s)
{
    s.dev = nullptr;  // staging buffers belong to the context
}

}  // namespace lbx

using namespace lbx;

extern "C" {

lbx_status lbx_init(lbx_ctx **out, int device)
{
    return guarded([&]() -> lbx_status {
        if (!out) return LBX_ERR_INVALID_ARG;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return LBX_ERR_NO_DEVICE;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) return LBX_ERR_NO_DEVICE;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return LBX_ERR_NO_DEVICE;
        {
            DeviceGuard g(device);
            const lbx_status st = noalloc_attach(device);
            if (st != LBX_OK) return st;
        }
        lbx_ctx *ctx = new lbx_ctx();
        ctx->device = device;
        if (const char *e = std::getenv("LBX_CHECK_MUL")) ctx->check = std::atoi(e) != 0;
        *out = ctx;
        return LBX_OK;
    });
}

lbx_status lbx_trim(lbx_ctx *ctx)
{
    return guarded([&]() -> lbx_status {
        if (!ctx) return LBX_ERR_INVALID_ARG;
        std::lock_guard<std::mutex> lock(ctx->mu);
        DeviceGuard g(ctx->device);
        (void)hipDeviceSynchronize();
        for (auto &kv : ctx->bufs)
            if (kv.second.ptr) (void)hipFree(kv.second.ptr);
        ctx->bufs.clear();
        stager_release(ctx);
        return LBX_OK;
    });
}

lbx_status lbx_destroy(lbx_ctx *ctx)
{
    return guarded([&]() -> lbx_status {
        if (!ctx) return LBX_ERR_INVALID_ARG;
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            DeviceGuard g(ctx->device);
            (void)hipDeviceSynchronize();
            for (auto &kv : ctx->bufs)
                if (kv.second.ptr) (void)hipFree(kv.second.ptr);
            for (auto &kv : ctx->twiddles) (void)hipFree(kv.second);
            for (auto ev : ctx->events) (void)hipEventDestroy(ev);
            if (ctx->last_done) (void)hipEventDestroy(ctx->last_done);
            if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
            mp_release(ctx);
            batch_release(ctx);
            stager_release(ctx);
        }
        delete ctx;
        return LBX_OK;
    });
}

const char *lbx_strerror(lbx_status st)
{
    switch (st) {
    case LBX_OK: return "success";
    case LBX_ERR_BAD_FFT_PRIME: return "LinBox ERROR: bad FFT Prime";
    case LBX_ERR_K_TOO_LARGE: return "FFT: k is larger that log2(p-1)";
    case LBX_ERR_K_NOT_POSITIVE: return "FFT: k must be positive";
    case LBX_ERR_NOT_PRIMITIVE_ROOT: return "w is not a primitive k-root of unity";
    case LBX_ERR_SIZE_MISMATCH: return "size mismatch: incompatible dimensions or view resized larger than available";
    case LBX_ERR_NOT_ENOUGH_PRIMES: return "not enough FFT primes";
    case LBX_ERR_INVALID_ARG: return "invalid argument";
    case LBX_ERR_PRIME_TOO_LARGE: return "modulus out of the supported range";
    case LBX_ERR_HIP: return "HIP runtime error";
    case LBX_ERR_OUT_OF_MEMORY: return "device out of memory";
    case LBX_ERR_UNSUPPORTED: return "unsupported size";
    case LBX_ERR_NO_DEVICE: return "no gfx950 device";
    case LBX_ERR_WORKSPACE_TOO_SMALL: return "the product needs more device memory than the fixed workspace holds";
    case LBX_ERR_CHECK_FAILED: return "the product check at a random point failed";
    }
    return "unknown status";
}

lbx_status lbx_profile_enable(lbx_ctx *ctx, int enable)
{
    return guarded([&]() -> lbx_status {
        if (!ctx) return LBX_ERR_INVALID_ARG;
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->profiling = enable < 0 ? 0 : (enable > 2 ? 2 : enable);
        return LBX_OK;
    });
}

int lbx_profile_get(lbx_ctx *ctx, const char **labels, float *ms, int max)
{
    // no C++ exception may reach a C caller (only a failed allocation throws here)
    try {
        if (!ctx) return 0;
        std::lock_guard<std::mutex> lock(ctx->mu);
        DeviceGuard guard(ctx->device);
        // average per label over every call recorded since the previous get
        std::vector<std::string> order;
        std::map<std::string, std::pair<double, int>> acc;
        std::vector<hipEvent_t> used;
        for (auto &rec : ctx->prof_pending) {
            const std::string &label = std::get<0>(rec);
            hipEvent_t a = std::get<1>(rec), b = std::get<2>(rec);
            (void)hipEventSynchronize(b);
            float t = 0;
            (void)hipEventElapsedTime(&t, a, b);
            if (!acc.count(label)) order.push_back(label);
            acc[label].first += t;
            acc[label].second += 1;
            used.push_back(a);
            used.push_back(b);
        }
        std::sort(used.begin(), used.end());
        used.erase(std::unique(used.begin(), used.end()), used.end());
        for (auto ev : used) (void)hipEventDestroy(ev);
        ctx->prof_pending.clear();
        if (!order.empty()) {
            ctx->prof_labels.clear();
            ctx->prof_ms.clear();
            for (auto &l : order) {
                ctx->prof_labels.push_back(l);
                ctx->prof_ms.push_back((float)(acc[l].first / acc[l].second));
            }
        }
        int n = (int)ctx->prof_ms.size();
        if (n > max) n = max;
        ctx->prof_label_ptrs.clear();
        for (auto &l : ctx->prof_labels) ctx->prof_label_ptrs.push_back(l.c_str());
        for (int i = 0; i < n; ++i) {
            if (labels) labels[i] = ctx->prof_label_ptrs[i];
            if (ms) ms[i] = ctx->prof_ms[i];
        }
        return n;
    } catch (...) {
        return 0;
    }
}

static lbx_status mul_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                          size_t max_rowdeg, int method, void *stream);

lbx_status lbx_polymat_mul_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                               size_t max_rowdeg, int method, void *stream)
{
    return guarded([&]() -> lbx_status {
        const lbx_status st = mul_u32(ctx, p, c, a, b, max_rowdeg, method, stream);
        return st == LBX_OK ? self_check(ctx, p, c, a, b, stream) : st;
    });
}

static lbx_status mul_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a, const lbx_polmat *b,
                          size_t max_rowdeg, int method, void *stream)
{
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
    // word-size kernels keep residues in [0,p) with 2p < 2^32; larger p go
    // through the CRT route, whose primes are below 2^31
    const bool small_p = p < (1u << 31);
    const bool constant = sa + sb <= 2 && small_p;
    const bool single_prime = !constant && small_p && sa + sb > 2 && (p - 1) % pts == 0;
    if (!constant && !single_prime) {
        lbx_status st = lbx_polymat_mul_mp(ctx, &p, 1, c, a, b, max_rowdeg, method, stream);
        return st;
    }

    std::lock_guard<std::mutex> lock(ctx->mu);
    DeviceGuard guard(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    StreamOrder stream_order(ctx, s);
    Staged sa_, sb_, sc_;
    // host storage of c is copied in only when the call leaves part of it as it is
    const bool c_in = !(c->shift == 0 && deg + 1 >= c->storage_size);
    lbx_status st = stage_in(ctx, "stage_a", a, 4, true, &sa_, s);
    if (st == LBX_OK) st = stage_in(ctx, "stage_b", b, 4, true, &sb_, s);
    if (st == LBX_OK) st = stage_in(ctx, "stage_c", c, 4, c_in, &sc_, s);
    const bool host_io = sa_.owned || sb_.owned || sc_.owned;
    if (st == LBX_OK) {
        const DevView da = devview(a, sa_.dev), db = devview(b, sb_.dev), dc = devview(c, sc_.dev);
        if (constant) {
            // the FFT domain's c.resize(deg+1) with deg = max_rowdeg: the
            // product of constants is coefficient 0, coefficients 1..deg are zero
            st = constant_product(ctx, p, da, db, dc, m, k, n, s);
            if (st == LBX_OK && deg > 0) st = from_hip(launch_fill(dc, m * n, 1, deg, 0u, s));
        } else
            st = fft_prime_product(ctx, p, L, da, db, dc, m, k, n, deg + 1, method, s);
    }
    if (st == LBX_OK && sc_.owned) {
        st = device_to_host(ctx, c->data, sc_.dev, storage_words(c) * 4, s);
    } else if (st == LBX_OK && host_io) {
        st = from_hip(hipStreamSynchronize(s));
    }
    if (host_io) (void)hipStreamSynchronize(s);
    release(sa_);
    release(sb_);
    release(sc_);
    if (st == LBX_OK) c->size = deg + 1;
    return st;
}

lbx_status lbx_polymat_mul_fft_u32(lbx_ctx *ctx, uint32_t p, unsigned lpts, size_t m, size_t k, size_t n,
                                   const uint32_t *a, const uint32_t *b, uint32_t *c, int method, void *stream)
{
    return guarded([&]() -> lbx_status {
        if (!ctx || !a || !b || !c || !m || !k || !n) return LBX_ERR_INVALID_ARG;
        if (method < 0 || method > 2) return LBX_ERR_INVALID_ARG;
        const size_t pts = (size_t)1 << lpts;
        if (p < 2 || p >= (1u << 31)) return LBX_ERR_PRIME_TOO_LARGE;
        if ((p - 1) % pts != 0) return LBX_ERR_BAD_FFT_PRIME;
        if (lpts == 0) return LBX_ERR_K_NOT_POSITIVE;
        lbx_polmat va{(void *)a, LBX_POLFIRST, m, k, pts, pts, 0};
        lbx_polmat vb{(void *)b, LBX_POLFIRST, k, n, pts, pts, 0};
        lbx_polmat vc{(void *)c, LBX_POLFIRST, m, n, pts, pts, 0};
        std::lock_guard<std::mutex> lock(ctx->mu);
        DeviceGuard guard(ctx->device);
        hipStream_t s = (hipStream_t)stream;
        StreamOrder stream_order(ctx, s);
        Staged sa_, sb_, sc_;
        lbx_status st = stage_in(ctx, "stage_a", &va, 4, true, &sa_, s);
        if (st == LBX_OK) st = stage_in(ctx, "stage_b", &vb, 4, true, &sb_, s);
        if (st == LBX_OK) st = stage_in(ctx, "stage_c", &vc, 4, false, &sc_, s);
        if (st == LBX_OK)
            st = fft_prime_product(ctx, p, lpts, devview(&va, sa_.dev), devview(&vb, sb_.dev), devview(&vc, sc_.dev), m,
                                   k, n, pts, method, s);
        if (st == LBX_OK && sc_.owned) {
            st = device_to_host(ctx, c, sc_.dev, m * n * pts * 4, s);
        }
        if (sa_.owned || sb_.owned || sc_.owned) (void)hipStreamSynchronize(s);
        release(sa_);
        release(sb_);
        release(sc_);
        return st;
    });
}

static lbx_status ntt_entry(lbx_ctx *ctx, uint32_t p, unsigned lpts, uint32_t root, uint32_t *data, size_t count,
                            uint32_t *root_used, void *stream, bool inverse)
{
    if (!ctx || !data) return LBX_ERR_INVALID_ARG;
    if (p < 3 || p >= (1u << 31)) return LBX_ERR_PRIME_TOO_LARGE;
    if (!lpts) return LBX_ERR_K_NOT_POSITIVE;  // fft.h:135-136
    uint32_t w = root;
    if (w == 0) {
        LBX_TRY(primitive_root(p, lpts, &w));
    } else if (!is_primitive_root(p, w, lpts)) {
        return LBX_ERR_NOT_PRIMITIVE_ROOT;
    }
    if (lpts > (unsigned)kMaxLog) return LBX_ERR_UNSUPPORTED;
    if (root_used) *root_used = w;
    if (!count) return LBX_OK;
    const size_t N = (size_t)1 << lpts;
    std::lock_guard<std::mutex> lock(ctx->mu);
    DeviceGuard guard(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    StreamOrder stream_order(ctx, s);
    const uint2 *tw;
    LBX_TRY(twiddles(ctx, p, lpts, w, inverse ? kTwShoup : kTwMont, &tw));
    lbx_polmat v{(void *)data, LBX_POLFIRST, count, 1, N, N, 0};
    Staged sd;
    LBX_TRY(stage_in(ctx, "stage_a", &v, 4, true, &sd, s));
    const DevView dv = devview(&v, sd.dev);
    hipError_t e = inverse ? launch_ntt_inverse((uint32_t *)sd.dev, count, lpts, tw, p, dv, N, 1, s)
                           : launch_ntt_forward(dv, count, lpts, tw, p, (uint32_t *)sd.dev, s);
    lbx_status st = from_hip(e);
    if (st == LBX_OK && sd.owned) {
        st = device_to_host(ctx, data, sd.dev, count * N * 4, s);
    }
    if (sd.owned) (void)hipStreamSynchronize(s);
    release(sd);
    return st;
}

lbx_status lbx_ntt_forward_u32(lbx_ctx *ctx, uint32_t p, unsigned lpts, uint32_t root, uint32_t *data, size_t count,
                               uint32_t *root_used, void *stream)
{
    return guarded([&]() -> lbx_status {
        return ntt_entry(ctx, p, lpts, root, data, count, root_used, stream, false);
    });
}

lbx_status lbx_ntt_inverse_u32(lbx_ctx *ctx, uint32_t p, unsigned lpts, uint32_t root, uint32_t *data, size_t count,
                               uint32_t *root_used, void *stream)
{
    return guarded([&]() -> lbx_status {
        return ntt_entry(ctx, p, lpts, root, data, count, root_used, stream, true);
    });
}

}  // extern "C"
