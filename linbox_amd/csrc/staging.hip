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
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "lbx_ctx.h"

namespace lbx {

namespace {

constexpr size_t kChunk = 16u << 20;     // bytes per pinned buffer
constexpr int kSlots = 4;                // pinned buffers in the ring
constexpr size_t kDirectMax = 4u << 20;  // smaller copies go straight through the runtime
constexpr size_t kSplitMin = 1u << 20;   // smaller host copies stay on the calling thread

}  // namespace

// memcpy split over a fixed set of worker threads and the calling thread.
class CopyPool {
public:
    explicit CopyPool(int workers)
    {
        for (int i = 0; i < workers; ++i) threads_.emplace_back([this, i] { work(i + 1); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        go_.notify_all();
        for (auto &t : threads_) t.join();
    }
    CopyPool(const CopyPool &) = delete;
    CopyPool &operator=(const CopyPool &) = delete;

    void copy(void *dst, const void *src, size_t bytes)
    {
        if (threads_.empty() || bytes < kSplitMin) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            dst_ = (char *)dst;
            src_ = (const char *)src;
            bytes_ = bytes;
            pending_ = (int)threads_.size();
            ++gen_;
        }
        go_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return pending_ == 0; });
    }

private:
    // part q of parts() pieces, boundaries on 4 KiB pages
    void part(int q)
    {
        const size_t parts = threads_.size() + 1;
        const size_t a = (bytes_ * q / parts) & ~(size_t)4095, b = q + 1 == (int)parts ? bytes_
                                                                                       : (bytes_ * (q + 1) / parts) & ~(size_t)4095;
        if (b > a) std::memcpy(dst_ + a, src_ + a, b - a);
    }

    void work(int q)
    {
        unsigned seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            go_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            lk.unlock();
            part(q);
            lk.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }

    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable go_, done_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0;
    unsigned gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

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
        HostStager *hs = new HostStager();
        hipError_t e = hipHostMalloc((void **)&hs->pinned, kChunk * kSlots, hipHostMallocDefault);
        for (int i = 0; i < kSlots && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&hs->ev[i], hipEventDisableTiming);
        if (e != hipSuccess) {
            for (auto ev : hs->ev)
                if (ev) (void)hipEventDestroy(ev);
            if (hs->pinned) (void)hipHostFree(hs->pinned);
            delete hs;
            return from_hip(e);
        }
        hs->pool = new CopyPool(copy_threads() - 1);
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
    for (size_t off = 0, i = 0; off < bytes; off += kChunk, ++i) {
        const int slot = (int)(i % kSlots);
        const size_t len = bytes - off < kChunk ? bytes - off : kChunk;
        char *buf = hs->pinned + (size_t)slot * kChunk;
        LBX_TRY(wait_slot(hs, slot));
        hs->pool->copy(buf, (const char *)host + off, len);
        LBX_TRY_HIP(hipMemcpyAsync((char *)dev + off, buf, len, hipMemcpyHostToDevice, s));
        LBX_TRY_HIP(hipEventRecord(hs->ev[slot], s));
        hs->queued[slot] = true;
    }
    // the caller's memory has been read; the last DMAs run ahead of the
    // call's kernels on s
    return LBX_OK;
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
    const size_t nchunks = (bytes + kChunk - 1) / kChunk;
    auto queue = [&](size_t i) -> lbx_status {
        const int slot = (int)(i % kSlots);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        // DMAs on s run in order, so a buffer's previous DMA (on s) is done
        // before this one writes it; the host has already read it
        LBX_TRY_HIP(hipMemcpyAsync(hs->pinned + (size_t)slot * kChunk, (const char *)dev + off, len,
                                   hipMemcpyDeviceToHost, s));
        LBX_TRY_HIP(hipEventRecord(hs->ev[slot], s));
        hs->queued[slot] = true;
        return LBX_OK;
    };
    for (size_t i = 0; i < nchunks && i < (size_t)kSlots; ++i) LBX_TRY(queue(i));
    for (size_t i = 0; i < nchunks; ++i) {
        const int slot = (int)(i % kSlots);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        LBX_TRY(wait_slot(hs, slot));
        hs->pool->copy((char *)host + off, hs->pinned + (size_t)slot * kChunk, len);
        if (i + kSlots < nchunks) LBX_TRY(queue(i + kSlots));
    }
    return LBX_OK;
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
