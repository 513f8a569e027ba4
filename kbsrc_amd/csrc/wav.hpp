// WAV reading and writing for the kbsrc command-line tool: the reference's format rules and messages
// (src/io/wav.hpp:9-23, src/io/wav.cpp:134-278). Host code; not on the GPU path.
#pragma once

#include <string>
#include <vector>

namespace kbsrc::io {

struct Audio {
    int rate = 0, ch = 0;
    long long frames = 0;
    std::vector<double> d;  // interleaved
    int fmt_bits = 0;
    bool fmt_float = false;
};

// PCM 8/16/24/32-bit and IEEE float 32/64 (plain or WAVE_FORMAT_EXTENSIBLE) -> interleaved f64.
// Throws std::runtime_error with the reference's messages.
Audio read_wav(const char* path);

// Fails (std::runtime_error) when a header in this format cannot describe ch channels at rate.
void check_wav_output(int ch, int rate, const std::string& format);

// format: f64, f32, s32, s24 or s16. Integer formats round in the current rounding mode (TPDF dither from
// mt19937_64(12345) when asked), clip to the format's range and count the clipped samples. y: planar.
void write_wav(const char* path, const std::vector<std::vector<double>>& y, int ch, long long frames, int rate,
               const std::string& format, bool dither, long long& clipped);

}  // namespace kbsrc::io
