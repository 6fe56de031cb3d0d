// Copies of host operands to and from the device.
//
// A LinBox caller's polynomial matrices live in pageable host memory.  The
// runtime's own pageable copies run at about 26 GB/s to the device and 16 GB/s
// back, against 57 GB/s for DMA from pinned memory, and pinning the caller's
// memory in place (hipHostRegister) is not available everywhere.  Large copies
// therefore go through a ring of pinned buffers owned by the context: a few
// host threads copy one chunk between the caller's memory and a pinned buffer
// while the DMA engine moves the previous chunk, so the two overlap.  The DMAs
// are queued on the call's stream, behind the work the call already waits for.
#include <cstdlib>
#include <new>

#include "copy_pool.h"
#include "lbx_ctx.h"

namespace lbx {

namespace {

constexpr size_t kChunk = 16u << 20;     // bytes per pinned buffer
constexpr int kSlots = 4;                // pinned buffers in the ring
constexpr size_t kDirectMax = 4u << 20;  // smaller copies go straight through the runtime

}  // namespace

struct HostStager {
    char *pinned = nullptr;       // kSlots buffers of kChunk bytes
    hipEvent_t ev[kSlots] = {};   // recorded after the last DMA that used each buffer
    bool queued[kSlots] = {};
    CopyPool *pool = nullptr;
};

namespace {

// Host threads that copy between the caller's memory and the pinned buffers:
// LBX_COPY_THREADS (1..16, the calling thread included), else 4.
int copy_threads()
{
    int t = 4;
    if (const char *e = std::getenv("LBX_COPY_THREADS")) t = std::atoi(e);
    return t < 1 ? 1 : (t > 16 ? 16 : t);
}

lbx_status get_stager(lbx_ctx *ctx, HostStager **out)
{
    if (!ctx->stager) {
        HostStager *hs = new (std::nothrow) HostStager();
        if (!hs) return LBX_ERR_OUT_OF_MEMORY;
        hipError_t e = hipHostMalloc((void **)&hs->pinned, kChunk * kSlots, hipHostMallocDefault);
        for (int i = 0; i < kSlots && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&hs->ev[i], hipEventDisableTiming);
        if (e != hipSuccess) {
            for (auto ev : hs->ev)
                if (ev) (void)hipEventDestroy(ev);
            if (hs->pinned) (void)hipHostFree(hs->pinned);
            delete hs;
            return from_hip(e);
        }
        // no exception leaves the C ABI: a pool the system grants fewer threads
        // splits its copies over those
        hs->pool = new (std::nothrow) CopyPool(copy_threads() - 1);
        if (!hs->pool) {
            for (auto ev : hs->ev) (void)hipEventDestroy(ev);
            (void)hipHostFree(hs->pinned);
            delete hs;
            return LBX_ERR_OUT_OF_MEMORY;
        }
        ctx->stager = hs;
    }
    *out = ctx->stager;
    return LBX_OK;
}

// Waits until the DMA that last used buffer `slot` is done.
lbx_status wait_slot(HostStager *hs, int slot)
{
    if (hs->queued[slot]) {
        LBX_TRY_HIP(hipEventSynchronize(hs->ev[slot]));
        hs->queued[slot] = false;
    }
    return LBX_OK;
}

// After a failed step no DMA may still read or write a pinned buffer whose
// use is not recorded: wait for the stream and mark every buffer free.
lbx_status drain(HostStager *hs, hipStream_t s, lbx_status st)
{
    (void)hipStreamSynchronize(s);
    for (bool &q : hs->queued) q = false;
    return st;
}

lbx_status ring_to_device(HostStager *hs, char *dev, const char *host, size_t bytes, hipStream_t s)
{
    for (size_t off = 0, i = 0; off < bytes; off += kChunk, ++i) {
        const int slot = (int)(i % kSlots);
        const size_t len = bytes - off < kChunk ? bytes - off : kChunk;
        char *buf = hs->pinned + (size_t)slot * kChunk;
        LBX_TRY(wait_slot(hs, slot));
        hs->pool->copy(buf, host + off, len);
        LBX_TRY_HIP(hipMemcpyAsync(dev + off, buf, len, hipMemcpyHostToDevice, s));
        LBX_TRY_HIP(hipEventRecord(hs->ev[slot], s));
        hs->queued[slot] = true;
    }
    return LBX_OK;
}

lbx_status ring_to_host(HostStager *hs, char *host, const char *dev, size_t bytes, hipStream_t s)
{
    const size_t nchunks = (bytes + kChunk - 1) / kChunk;
    auto queue = [&](size_t i) -> lbx_status {
        const int slot = (int)(i % kSlots);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        // DMAs on s run in order, so a buffer's previous DMA (on s) is done
        // before this one writes it; the host has already read it
        LBX_TRY_HIP(hipMemcpyAsync(hs->pinned + (size_t)slot * kChunk, dev + off, len, hipMemcpyDeviceToHost, s));
        LBX_TRY_HIP(hipEventRecord(hs->ev[slot], s));
        hs->queued[slot] = true;
        return LBX_OK;
    };
    for (size_t i = 0; i < nchunks && i < (size_t)kSlots; ++i) LBX_TRY(queue(i));
    for (size_t i = 0; i < nchunks; ++i) {
        const int slot = (int)(i % kSlots);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        LBX_TRY(wait_slot(hs, slot));
        hs->pool->copy(host + off, hs->pinned + (size_t)slot * kChunk, len);
        if (i + kSlots < nchunks) LBX_TRY(queue(i + kSlots));
    }
    return LBX_OK;
}

}  // namespace

lbx_status host_to_device(lbx_ctx *ctx, void *dev, const void *host, size_t bytes, hipStream_t s)
{
    if (!bytes) return LBX_OK;
    if (bytes <= kDirectMax) {
        LBX_TRY_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s));
        LBX_TRY_HIP(hipStreamSynchronize(s));  // the caller may reuse its memory on return
        return LBX_OK;
    }
    HostStager *hs;
    LBX_TRY(get_stager(ctx, &hs));
    // on return the caller's memory has been read; the last DMAs run ahead of
    // the call's kernels on s
    const lbx_status st = ring_to_device(hs, (char *)dev, (const char *)host, bytes, s);
    return st == LBX_OK ? st : drain(hs, s, st);
}

lbx_status device_to_host(lbx_ctx *ctx, void *host, const void *dev, size_t bytes, hipStream_t s)
{
    if (!bytes) return LBX_OK;
    if (bytes <= kDirectMax) {
        LBX_TRY_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
        LBX_TRY_HIP(hipStreamSynchronize(s));
        return LBX_OK;
    }
    HostStager *hs;
    LBX_TRY(get_stager(ctx, &hs));
    const lbx_status st = ring_to_host(hs, (char *)host, (const char *)dev, bytes, s);
    return st == LBX_OK ? st : drain(hs, s, st);
}

void stager_release(lbx_ctx *ctx)
{
    HostStager *hs = ctx->stager;
    if (!hs) return;
    for (int i = 0; i < kSlots; ++i) {
        if (hs->queued[i]) (void)hipEventSynchronize(hs->ev[i]);
        (void)hipEventDestroy(hs->ev[i]);
    }
    (void)hipHostFree(hs->pinned);
    delete hs->pool;
    delete hs;
    ctx->stager = nullptr;
}

}  // namespace lbx
