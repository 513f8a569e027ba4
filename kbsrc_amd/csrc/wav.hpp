// WAV files for the kbsrc command-line tool (host code, off the GPU path). What a file contains and which inputs are
// refused, with which message, is the reference's (src/io/wav.hpp:9-23, src/io/wav.cpp:109-253): PCM of 1-4 bytes
// and IEEE float of 4 or 8 bytes per sample in, plain or WAVE_FORMAT_EXTENSIBLE; f64, f32, s32, s24 and s16 out,
// with the same header bytes. Errors are std::runtime_error.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace kbsrc::io {

// An output sample format: its name on the command line, bits per sample and whether it is IEEE float.
struct SampleFormat {
    const char* name;
    int bits;
    bool floating;
    int bytes() const { return bits / 8; }
};
// nullptr when the name is not one of f64, f32, s32, s24, s16
const SampleFormat* find_format(const std::string& name);

// Reads a WAV file's header on construction (every check, in the reference's order) and decodes its samples to
// interleaved float64 on demand, in file order.
class WavReader {
public:
    explicit WavReader(const char* path);
    ~WavReader();
    WavReader(const WavReader&) = delete;
    WavReader& operator=(const WavReader&) = delete;
    int rate() const { return rate_; }
    int channels() const { return ch_; }
    long long frames() const { return frames_; }
    void read(double* dst, long long frames);   // the next `frames` frames

private:
    using Decoder = void (*)(const uint8_t* src, size_t samples, double* dst);
    std::FILE* file_ = nullptr;
    Decoder decode_ = nullptr;
    int rate_ = 0, ch_ = 0, sample_bytes_ = 0;
    long long frames_ = 0, consumed_ = 0;
    std::vector<uint8_t> raw_;
};

// Fails when a header in this format cannot describe ch channels at rate: the frame size must fit 16 bits and the
// byte rate 32 bits.
void check_wav_output(int ch, int rate, const std::string& format);

// The header for `frames` frames of ch channels at rate in this format, the data written by the caller, and the
// pad byte after an odd-sized data chunk on finish().
class WavWriter {
public:
    WavWriter(const char* path, int ch, long long frames, int rate, const std::string& format);
    ~WavWriter();
    WavWriter(const WavWriter&) = delete;
    WavWriter& operator=(const WavWriter&) = delete;
    void write(const void* bytes, size_t n);
    void finish();
    size_t frame_bytes() const { return frame_bytes_; }

private:
    std::FILE* file_ = nullptr;
    size_t frame_bytes_ = 0;
    unsigned long long data_bytes_ = 0;
};

// A whole file from planar output y (y[c * frames + i]). Integer formats scale, add TPDF dither when asked (two
// uniform draws in [-0.5, 0.5) per sample from mt19937_64 seeded 12345, in frame then channel order), round in the
// caller's rounding mode, clip to the format's range (clipped counts the clipped samples) and write NaN as 0.
void write_wav(const char* path, const double* y, int ch, long long frames, int rate, const std::string& format,
               bool dither, long long& clipped);

}  // namespace kbsrc::io
