// KBSRC public C++ API, served by the MI355X engine.
// Same shape as the reference's include/kbsrc/kbsrc.hpp:8-45 (Precision, Options, Stats, resample, version), so
// callers of the CPU converter compile unchanged against this header.
#pragma once

#include <vector>

namespace kbsrc {

enum class Precision {
    Fast,        // float64 pipeline
    Exact,       // double-double (~106-bit) pipeline
    LongDouble,  // x87 80-bit in the reference; not available on the GPU engine (rejected as out of range)
};

struct Options {
    Precision precision = Precision::Exact;
    double length = 0.2;    // one-sided kernel length in seconds
    double beta = 40.0;     // Kaiser window beta
    int threads = 0;        // the reference's worker count (>= 0); accepted, not used by the GPU engine
    int block = 0;          // block size factor nb, 0 = automatic (nb changes the rounding of the output)
    bool lpc_edges = true;  // Burg LPC extrapolation at the signal edges
};

struct Stats {
    long long p = 0, q = 0;     // reduced rate ratio P/Q
    long long taps = 0;         // kernel taps at the kernel rate
    int block = 0, threads = 0;
    bool general_fold = false;  // near-unity ratio path
    double design_seconds = 0, process_seconds = 0;
    double passband = 0, stopband = 0;  // Hz
    const char* isa = "";       // engine that ran: "gfx950"
};

// Resamples interleaved input (frames x channels) from rate_in to rate_out on the GPU. Returns planar output,
// ceil(frames * rate_out / rate_in) samples per channel, bit-identical to the reference converter for the same
// inputs and options. Throws std::invalid_argument on bad arguments (the reference's messages),
// std::runtime_error on device errors, std::bad_alloc when memory runs out. Exact runs its host-side
// double-double work in round-to-nearest whatever the caller's rounding mode, and restores that mode.
std::vector<std::vector<double>> resample(const double* interleaved, long long frames, int channels, int rate_in,
                                          int rate_out, const Options& options = {}, Stats* stats = nullptr);

const char* version();

}  // namespace kbsrc
