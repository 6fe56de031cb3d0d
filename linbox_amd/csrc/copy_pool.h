// memcpy split over a fixed set of host threads and the calling thread: the
// host side of the pinned-buffer copies of host operands (staging.hip).
// Plain C++, so the host tests build it on its own.
#pragma once

#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace lbx {

class CopyPool {
public:
    static constexpr size_t kSplitMin = 1u << 20;  // smaller copies stay on the calling thread

    // Starts up to `workers` threads; if the system refuses one, the pool keeps
    // those already running (copies are split over however many there are).
    explicit CopyPool(int workers) noexcept
    {
        try {
            threads_.reserve(workers > 0 ? (size_t)workers : 0);
            for (int i = 0; i < workers; ++i) threads_.emplace_back([this, i] { work(i + 1); });
        } catch (...) {
        }
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
    // part q of (workers + 1) pieces, boundaries on 4 KiB pages
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

}  // namespace lbx
