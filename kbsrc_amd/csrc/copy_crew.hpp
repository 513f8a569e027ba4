// A crew of host threads for large memcpy (no device code: also built on its own by the CPU tests).
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>

namespace kbsrc::gpu {

// Host memcpy of large buffers on a crew of worker threads made once (copies into fresh pageable memory also take
// their page faults in parallel). One copy at a time: a caller that finds the crew busy copies alone. A copy is cut
// into parts; the parts are claimed through one atomic word holding the copy's number and the next part, so a
// worker that wakes late can never claim a part of a later copy with the addresses of an earlier one.
class CopyCrew {
public:
    void copy(void* dst, const void* src, size_t bytes) {
        std::unique_lock<std::mutex> busy(busy_, std::try_to_lock);
        if (!busy.owns_lock() || bytes < kMinSplit || n_ == 0) {
            std::memcpy(dst, src, bytes);
            return;
        }
        const size_t parts = std::min<size_t>(n_ + 1, (bytes + kMinPart - 1) / kMinPart);
        Job j;
        j.dst = static_cast<char*>(dst);
        j.src = static_cast<const char*>(src);
        j.bytes = bytes;
        j.per = ((bytes + parts - 1) / parts + 4095) & ~(size_t)4095;
        j.parts = (uint32_t)((bytes + j.per - 1) / j.per);
        {
            std::lock_guard<std::mutex> lock(m_);
            j.gen = ++gen_;
            job_ = j;
            left_ = j.parts;
            next_.store((uint64_t)j.gen << 32);
        }
        cv_.notify_all();
        run_parts(j);
        std::unique_lock<std::mutex> lock(m_);
        done_.wait(lock, [&] { return left_ == 0; });
    }
    static CopyCrew& get() {
        static CopyCrew* c = new CopyCrew;   // kept until the process ends; its threads wait for work
        return *c;
    }
    explicit CopyCrew(unsigned workers = default_workers()) {
        for (unsigned t = 0; t < workers; ++t) {
            try {
                std::thread(&CopyCrew::work, this).detach();
                ++n_;
            } catch (...) {
                break;
            }
        }
    }
    static unsigned default_workers() {
        const unsigned hw = std::thread::hardware_concurrency();
        return std::min(8u, std::max(1u, hw)) - 1;
    }

private:
    static constexpr size_t kMinSplit = 4u << 20, kMinPart = 1u << 20;
    struct Job {
        char* dst = nullptr;
        const char* src = nullptr;
        size_t bytes = 0, per = 0;
        uint32_t parts = 0, gen = 0;
    };
    // claims and copies parts of job j while they last (j is the copy whose number the claim word holds)
    void run_parts(const Job& j) {
        for (;;) {
            uint64_t v = next_.load();
            do {
                if ((uint32_t)(v >> 32) != j.gen || (uint32_t)v >= j.parts) return;
            } while (!next_.compare_exchange_weak(v, v + 1));
            const size_t o = (size_t)(uint32_t)v * j.per;
            std::memcpy(j.dst + o, j.src + o, std::min(j.per, j.bytes - o));
            std::lock_guard<std::mutex> lock(m_);
            if (--left_ == 0) done_.notify_all();
        }
    }
    void work() {
        uint32_t seen = 0;
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lock(m_);
                cv_.wait(lock, [&] { return gen_ != seen; });
                seen = gen_;
                j = job_;
            }
            run_parts(j);
        }
    }
    std::mutex busy_, m_;
    std::condition_variable cv_, done_;
    unsigned n_ = 0;
    uint32_t gen_ = 0;
    Job job_;
    uint32_t left_ = 0;
    std::atomic<uint64_t> next_{0};
};

}  // namespace kbsrc::gpu
