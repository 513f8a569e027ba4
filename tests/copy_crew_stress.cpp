// Stress of the host copy crew (kbsrc_amd/csrc/copy_crew.hpp): several threads copy buffers of random sizes at once,
// some through the crew and some alone when it is busy; every copy must equal its source. Built by
// tests/test_copy_crew.py, also under ThreadSanitizer.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "copy_crew.hpp"

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 30;
    auto* crew = new kbsrc::gpu::CopyCrew(5);   // its threads outlive main's return: never freed
    std::atomic<int> bad{0};
    std::vector<std::thread> callers;
    for (int t = 0; t < 3; ++t) {
        callers.emplace_back([&, t] {
            std::mt19937_64 rng(1234 + t);
            const size_t cap = (size_t)12 << 20;
            std::vector<unsigned char> src(cap + 4096), dst(cap + 4096);
            for (size_t i = 0; i < src.size(); ++i) src[i] = (unsigned char)(i * 131 + t);
            for (int it = 0; it < iters; ++it) {
                const size_t n = 1 + rng() % cap, off = rng() % 4096;
                src[off + rng() % n] ^= 0x5a;   // a change between copies
                std::memset(dst.data(), 0, dst.size());
                crew->copy(dst.data() + off, src.data() + off, n);
                if (std::memcmp(dst.data() + off, src.data() + off, n) != 0) bad.fetch_add(1);
                for (size_t i = 0; i < off; ++i)
                    if (dst[i]) {
                        bad.fetch_add(1);
                        break;
                    }
            }
        });
    }
    for (auto& c : callers) c.join();
    std::printf("copies checked, mismatches %d\n", bad.load());
    return bad.load() ? 1 : 0;
}
