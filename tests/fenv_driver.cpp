// Test driver (not part of the library): calls kbsrc::resample with the caller in a given floating-point rounding
// mode and reports whether the mode survived, as the reference's tests/fenv_driver.cpp does for its engines.
//
// usage: kbsrc_fenv_driver <nearest|down|up|zero> <exact|fast> <rate_in> <rate_out> <channels> <length> <beta>
//                          <in.f64> <out.f64>
//   in.f64: interleaved little-endian doubles; out.f64: planar output, channel after channel.
// Prints "mode_kept=<0|1> threw=<0|1> frames=<n>". Exit 0, 3 when resample threw std::invalid_argument, 2 on usage
// or I/O errors.
#include <cfenv>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>
#include <xmmintrin.h>

#include "kbsrc/kbsrc.hpp"

namespace {

struct Mode {   // fegetround, and the SSE and x87 rounding-control bits the hardware holds
    int fe = -1;
    unsigned sse = 0, x87 = 0;
    static Mode now() {
        Mode m;
        m.fe = std::fegetround();
        unsigned short cw = 0;
        __asm__ volatile("fnstcw %0" : "=m"(cw));
        m.sse = _mm_getcsr() & 0x6000u;
        m.x87 = cw & 0x0C00u;
        return m;
    }
    bool same(const Mode& o) const { return fe == o.fe && sse == o.sse && x87 == o.x87; }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 10) {
        std::fprintf(stderr, "usage: kbsrc_fenv_driver <nearest|down|up|zero> <exact|fast> <rate_in> <rate_out> "
                             "<channels> <length> <beta> <in.f64> <out.f64>\n");
        return 2;
    }
    const std::string mode = argv[1], prec = argv[2];
    const int fe = mode == "nearest" ? FE_TONEAREST : mode == "down" ? FE_DOWNWARD : mode == "up" ? FE_UPWARD
                   : mode == "zero" ? FE_TOWARDZERO : -1;
    kbsrc::Options opt;
    opt.precision = prec == "fast" ? kbsrc::Precision::Fast : kbsrc::Precision::Exact;
    const int rate_in = std::atoi(argv[3]), rate_out = std::atoi(argv[4]), channels = std::atoi(argv[5]);
    opt.length = std::strtod(argv[6], nullptr);   // parsed before the mode changes: strtod follows it
    opt.beta = std::strtod(argv[7], nullptr);
    if (fe < 0 || channels < 1 || (prec != "exact" && prec != "fast")) return 2;
    std::vector<double> in;
    if (FILE* f = std::fopen(argv[8], "rb")) {
        double buf[512];
        for (size_t n; (n = std::fread(buf, sizeof(double), 512, f)) > 0;) in.insert(in.end(), buf, buf + n);
        std::fclose(f);
    } else {
        return 2;
    }
    if (std::fesetround(fe) != 0) return 2;
    const Mode before = Mode::now();
    std::vector<std::vector<double>> out;
    bool threw = false;
    try {
        out = kbsrc::resample(in.data(), (long long)(in.size() / channels), channels, rate_in, rate_out, opt);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    const Mode after = Mode::now();
    std::fesetround(FE_TONEAREST);
    std::printf("mode_kept=%d threw=%d frames=%zu\n", before.same(after) && before.fe == fe ? 1 : 0, threw ? 1 : 0,
                out.empty() ? (size_t)0 : out[0].size());
    FILE* f = std::fopen(argv[9], "wb");
    if (!f) return 2;
    for (const auto& c : out) std::fwrite(c.data(), sizeof(double), c.size(), f);
    std::fclose(f);
    return threw ? 3 : 0;
}
