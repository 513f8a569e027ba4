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

// The same decoder, streamed: the constructor parses the header (same checks and messages), read() decodes the
// next frames into interleaved f64 in chunks of 1 Mi samples (src/io/wav.cpp:150-170).
class WavReader {
public:
    explicit WavReader(const char* path);
    ~WavReader();
    WavReader(const WavReader&) = delete;
    WavReader& operator=(const WavReader&) = delete;
    int rate() const { return rate_; }
    int channels() const { return ch_; }
    long long frames() const { return frames_; }
    void read(double* dst, long long frames);

private:
    void* f_ = nullptr;
    int rate_ = 0, ch_ = 0, tag_ = 0, bits_ = 0, bps_ = 0;
    long long frames_ = 0, done_ = 0;
    std::vector<unsigned char> buf_;
};

// Streamed writing: the header for frames x ch samples in this format, then data written by the caller.
// The constructor checks the format and writes the header; write() appends data; finish() pads an odd data
// length and closes the file.
class WavWriter {
public:
    WavWriter(const char* path, int ch, long long frames, int rate, const std::string& format);
    ~WavWriter();
    WavWriter(const WavWriter&) = delete;
    WavWriter& operator=(const WavWriter&) = delete;
    void write(const void* bytes, size_t n);
    void finish();
    size_t frame_bytes() const { return align_; }

private:
    void* f_ = nullptr;
    size_t align_ = 0, data_bytes_ = 0;
};

// Fails (std::runtime_error) when a header in this format cannot describe ch channels at rate.
void check_wav_output(int ch, int rate, const std::string& format);

// format: f64, f32, s32, s24 or s16. Integer formats round in the current rounding mode (TPDF dither from
// mt19937_64(12345) when asked), clip to the format's range and count the clipped samples. y: planar.
void write_wav(const char* path, const std::vector<std::vector<double>>& y, int ch, long long frames, int rate,
               const std::string& format, bool dither, long long& clipped);

}  // namespace kbsrc::io
