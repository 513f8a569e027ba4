// Rounding-mode probe for the tests (not part of the library): sets the calling thread's floating-point rounding
// mode, calls kbsrc::resample, and reports whether the caller's mode came back unchanged.
//
//   kbsrc_fenv_driver key=value ...
//     round=nearest|down|up|zero  precision=exact|fast  from=<Hz> to=<Hz>  channels=<n>  len=<s>  beta=<b>
//     src=<file of interleaved little-endian float64>  dst=<file for the planar float64 output>
//     [tamper=1: after the call, change the SSE rounding field alone, as a library that failed to restore it would]
//
// Prints "restored=<0|1> rejected=<0|1> nout=<frames>". The mode counts as restored when fegetround(), the x87
// and SSE rounding-control fields of the environment (fegetenv) and the direction in which SSE double division
// actually rounds all read the same after the call as before it. Exit 0, 3 when the call was rejected
// (std::invalid_argument), 2 for a bad request.
#include <cfenv>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#include <xmmintrin.h>

#include "kbsrc/kbsrc.hpp"

// This file changes and inspects the rounding mode: no floating-point operation may be moved across fesetround or
// evaluated in a mode other than the one in force where it is written.
#pragma STDC FENV_ACCESS ON

namespace {

// The rounding state a thread can observe.
struct RoundingState {
    int fe;                // fegetround()
    unsigned x87, sse;     // rounding-control fields of the x87 control word and of MXCSR (fenv_t on x86-64)
    int direction[3];      // -1 / +1: whether q = a / b came out below or above the exact quotient, for three
                           // quotients whose round-to-nearest, down, up and toward-zero results all differ in sign
                           // pattern; the sign of fma(q, b, -a) is exact whatever the mode
    bool operator==(const RoundingState& o) const {
        return fe == o.fe && x87 == o.x87 && sse == o.sse && direction[0] == o.direction[0] &&
               direction[1] == o.direction[1] && direction[2] == o.direction[2];
    }
};

__attribute__((noinline)) RoundingState observe() {
    RoundingState s{};
    s.fe = std::fegetround();
    std::fenv_t env;
    std::fegetenv(&env);
    s.x87 = (env.__control_word >> 10) & 3u;
    s.sse = (env.__mxcsr >> 13) & 3u;
    static const double num[3] = {1.0, -1.0, 5.0};
    for (int k = 0; k < 3; ++k) {
        volatile double a = num[k], b = 3.0;
        volatile double q = a / b;
        volatile double r = std::fma(q, (double)b, -(double)a);
        s.direction[k] = r > 0 ? 1 : -1;
    }
    return s;
}

std::map<std::string, std::string> parse(int argc, char** argv) {
    std::map<std::string, std::string> kv;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const size_t eq = a.find('=');
        if (eq == std::string::npos || eq == 0) throw std::runtime_error("expected key=value, got " + a);
        kv[a.substr(0, eq)] = a.substr(eq + 1);
    }
    for (const char* need : {"round", "precision", "from", "to", "channels", "len", "beta", "src", "dst"})
        if (!kv.count(need)) throw std::runtime_error(std::string("missing ") + need);
    return kv;
}

}  // namespace

int main(int argc, char** argv) {
    std::map<std::string, std::string> kv;
    std::vector<double> in;
    int mode = 0;
    kbsrc::Options opt;
    int from = 0, to = 0, channels = 0;
    try {
        kv = parse(argc, argv);
        const std::map<std::string, int> modes = {
            {"nearest", FE_TONEAREST}, {"down", FE_DOWNWARD}, {"up", FE_UPWARD}, {"zero", FE_TOWARDZERO}};
        mode = modes.at(kv["round"]);
        if (kv["precision"] != "exact" && kv["precision"] != "fast") throw std::runtime_error("precision");
        opt.precision = kv["precision"] == "fast" ? kbsrc::Precision::Fast : kbsrc::Precision::Exact;
        from = std::stoi(kv["from"]);
        to = std::stoi(kv["to"]);
        channels = std::stoi(kv["channels"]);
        opt.length = std::stod(kv["len"]);   // numbers are read before the mode changes
        opt.beta = std::stod(kv["beta"]);
        if (channels < 1) throw std::runtime_error("channels");
        std::ifstream f(kv["src"], std::ios::binary);
        if (!f) throw std::runtime_error("cannot read " + kv["src"]);
        const std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        in.resize(bytes.size() / sizeof(double));
        if (!in.empty()) std::copy(bytes.begin(), bytes.begin() + in.size() * sizeof(double),
                                   reinterpret_cast<char*>(in.data()));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "kbsrc_fenv_driver: %s\n", e.what());
        return 2;
    }
    if (std::fesetround(mode) != 0) return 2;
    const RoundingState before = observe();
    std::vector<std::vector<double>> y;
    bool rejected = false;
    try {
        y = kbsrc::resample(in.data(), (long long)(in.size() / channels), channels, from, to, opt);
    } catch (const std::invalid_argument&) {
        rejected = true;
    }
    if (kv.count("tamper") && kv["tamper"] == "1")
        _MM_SET_ROUNDING_MODE(_MM_GET_ROUNDING_MODE() == _MM_ROUND_NEAREST ? _MM_ROUND_DOWN : _MM_ROUND_NEAREST);
    const RoundingState after = observe();
    std::fesetround(FE_TONEAREST);
    const bool restored = before == after && after.fe == mode;
    std::printf("restored=%d rejected=%d nout=%zu\n", restored ? 1 : 0, rejected ? 1 : 0, y.empty() ? 0 : y[0].size());
    std::ofstream out(kv["dst"], std::ios::binary);
    for (const auto& c : y) out.write(reinterpret_cast<const char*>(c.data()), (std::streamsize)(c.size() * 8));
    if (!out) return 2;
    return rejected ? 3 : 0;
}
