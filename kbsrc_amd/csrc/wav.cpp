// WAV I/O (host). Header layout, decoding and quantisation follow the reference byte for byte
// (src/io/wav.cpp:134-278): the undithered s16/s24/s32 output must be bit-identical.
#include "wav.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>

#pragma clang fp contract(off)

namespace kbsrc::io {
namespace {

[[noreturn]] void fail(const char* m) { throw std::runtime_error(m); }

struct FileCloser {
    FILE* f = nullptr;
    ~FileCloser() {
        if (f) fclose(f);
    }
};

unsigned rd16(const uint8_t* p) { return (unsigned)p[0] | (unsigned)p[1] << 8; }
uint32_t rd32(const uint8_t* p) { return (uint32_t)rd16(p) | (uint32_t)rd16(p + 2) << 16; }

int bytes_of(const std::string& fmt) { return fmt == "f64" ? 8 : fmt == "s24" ? 3 : fmt == "s16" ? 2 : 4; }

// One sample to a B-byte little-endian integer: scale, optional TPDF dither, round in the current mode, clip.
template <int B> void quantise(uint8_t* p, double y, bool dither, std::mt19937_64& rng,
                               std::uniform_real_distribution<double>& U, long long& clipped) {
    constexpr double sc = B == 2 ? 32768.0 : B == 3 ? 8388608.0 : 2147483648.0;
    double v = y * sc;
    if (dither) {
        const double d1 = U(rng);
        const double d2 = U(rng);
        v += d1 + d2;
    }
    v = std::nearbyint(v);
    if (v > sc - 1) {
        v = sc - 1;
        ++clipped;
    } else if (v < -sc) {
        v = -sc;
        ++clipped;
    }
    const int64_t iv = std::isnan(v) ? 0 : (int64_t)v;   // NaN is written as zero
    for (int t = 0; t < B; ++t) p[t] = (uint8_t)((iv >> (8 * t)) & 255);
}

}  // namespace

WavReader::WavReader(const char* path) {
    FileCloser fc{fopen(path, "rb")};
    FILE* f = fc.f;
    if (!f) fail("cannot open input");
    fseeko(f, 0, SEEK_END);
    const long long size = ftello(f);
    fseeko(f, 0, SEEK_SET);
    uint8_t riff[12];
    if (size < 12 || fread(riff, 1, 12, f) != 12 || memcmp(riff, "RIFF", 4) || memcmp(riff + 8, "WAVE", 4))
        fail("not a RIFF/WAVE file");
    long long pos = 12, data_pos = -1, data_len = 0;
    int tag = 0, ch = 0, bits = 0, align = 0;
    uint32_t rate = 0;
    bool have_fmt = false;
    while (pos + 8 <= size) {
        uint8_t hdr[8];
        fseeko(f, pos, SEEK_SET);
        if (fread(hdr, 1, 8, f) != 8) break;
        const uint32_t len = rd32(hdr + 4);
        const long long body = pos + 8;
        if (!memcmp(hdr, "fmt ", 4)) {
            uint8_t fm[40] = {};
            const size_t got = fread(fm, 1, std::min<uint32_t>(len, 40), f);
            if (got < 16) fail("bad fmt chunk");
            have_fmt = true;
            tag = (int)rd16(fm);
            ch = (int)rd16(fm + 2);
            rate = rd32(fm + 4);
            align = (int)rd16(fm + 12);
            bits = (int)rd16(fm + 14);
            if (tag == 0xFFFE && got >= 26) tag = (int)rd16(fm + 24);   // EXTENSIBLE: sub-format GUID's tag
        } else if (!memcmp(hdr, "data", 4)) {
            data_pos = body;
            data_len = std::min<long long>(len, size - body);
            if (len == 0xFFFFFFFFu || len == 0) data_len = size - body;   // streamed files
            break;
        }
        pos = body + len + (len & 1);
    }
    if (!have_fmt || data_pos < 0) fail("missing fmt/data chunk");
    if (!ch) fail("invalid WAV channel count");
    if (!rate || rate > INT_MAX) fail("invalid WAV sample rate");
    if (align <= 0 || align % ch) fail("invalid WAV block alignment");
    if (!(tag == 1 || (tag == 3 && (bits == 32 || bits == 64)))) fail("unsupported WAV format (need PCM or IEEE float)");
    const int bps = align / ch;
    if (tag == 1 && (bps < 1 || bps > 4 || (bits + 7) / 8 != bps)) fail("unsupported PCM width");
    if (tag == 3 && bits != 8 * bps) fail("IEEE float sample size does not match the block alignment");
    rate_ = (int)rate;
    ch_ = ch;
    tag_ = tag;
    bits_ = bits;
    bps_ = bps;
    frames_ = data_len / align;
    fseeko(f, data_pos, SEEK_SET);
    buf_.resize(((size_t)1 << 20) * bps);
    f_ = f;
    fc.f = nullptr;
}

WavReader::~WavReader() {
    if (f_) fclose(static_cast<FILE*>(f_));
}

void WavReader::read(double* dst, long long frames) {
    FILE* f = static_cast<FILE*>(f_);
    if (frames < 0 || frames > frames_ - done_) fail("read error");
    const size_t n = (size_t)frames * ch_, chunk = (size_t)1 << 20, bps = bps_;
    for (size_t k0 = 0; k0 < n; k0 += chunk) {
        const size_t m = std::min(chunk, n - k0);
        if (fread(buf_.data(), bps, m, f) != m) fail("read error");
        const uint8_t* s = buf_.data();
        double* o = dst + k0;
        if (tag_ == 3 && bits_ == 64) {
            memcpy(o, s, 8 * m);
        } else if (tag_ == 3) {
            for (size_t k = 0; k < m; ++k) {
                float v;
                memcpy(&v, s + 4 * k, 4);
                o[k] = v;
            }
        } else if (bps == 1) {
            for (size_t k = 0; k < m; ++k) o[k] = ((int)s[k] - 128) / 128.0;
        } else if (bps == 2) {
            for (size_t k = 0; k < m; ++k) o[k] = (int16_t)(s[2 * k] | s[2 * k + 1] << 8) / 32768.0;
        } else if (bps == 3) {
            for (size_t k = 0; k < m; ++k) {
                const uint8_t* p = s + 3 * k;
                int32_t v = p[0] | p[1] << 8 | p[2] << 16;
                if (v & 0x800000) v -= 1 << 24;
                o[k] = v / 8388608.0;
            }
        } else {
            for (size_t k = 0; k < m; ++k) {
                const uint8_t* p = s + 4 * k;
                o[k] = (int32_t)(p[0] | p[1] << 8 | p[2] << 16 | (uint32_t)p[3] << 24) / 2147483648.0;
            }
        }
    }
    done_ += frames;
}

Audio read_wav(const char* path) {
    WavReader r(path);
    Audio a;
    a.rate = r.rate();
    a.ch = r.channels();
    a.frames = r.frames();
    a.d.resize((size_t)a.frames * a.ch);
    r.read(a.d.data(), a.frames);
    return a;
}

void check_wav_output(int ch, int rate, const std::string& fmt) {
    const long long frame = (long long)ch * bytes_of(fmt);
    if (ch < 1 || frame > 0xFFFF) fail("too many channels for a WAV file in this format");
    if (rate < 1 || rate * frame > 0xFFFFFFFFLL) fail("sample rate too high for a WAV file with this many channels");
}

WavWriter::WavWriter(const char* path, int ch, long long frames, int rate, const std::string& fmt) {
    check_wav_output(ch, rate, fmt);
    const bool is_float = fmt == "f64" || fmt == "f32";
    const int bits = fmt == "f64" ? 64 : fmt == "f32" ? 32 : fmt == "s32" ? 32 : fmt == "s24" ? 24 : 16;
    const int bps = bits / 8, align = ch * bps;
    if (frames < 0 || frames > (0xFFFFFFFFLL - 80) / align) fail("output too large for WAV (4 GB limit)");
    const size_t data_bytes = (size_t)frames * align;
    std::vector<uint8_t> h;
    auto put16 = [&](unsigned v) { h.push_back(v & 255); h.push_back(v >> 8 & 255); };
    auto put32 = [&](uint32_t v) { put16(v & 0xFFFF); put16(v >> 16); };
    auto tag4 = [&](const char* t) { h.insert(h.end(), t, t + 4); };
    const bool extensible = ch > 2 || (!is_float && bits > 16);
    const unsigned fmt_len = extensible ? 40 : (is_float ? 18 : 16);
    const uint32_t riff_len = 4 + 8 + fmt_len + (is_float ? 12 : 0) + 8 + data_bytes + (data_bytes & 1);
    tag4("RIFF"); put32(riff_len); tag4("WAVE"); tag4("fmt "); put32(fmt_len);
    put16(extensible ? 0xFFFE : is_float ? 3 : 1); put16(ch); put32(rate); put32((uint32_t)rate * align);
    put16(align); put16(bits);
    if (fmt_len >= 18) put16(extensible ? 22 : 0);
    if (extensible) {
        put16(bits);
        put32(ch == 1 ? 4 : ch == 2 ? 3 : 0);   // speaker mask: mono = FC, stereo = FL|FR
        static const uint8_t kGuidTail[14] = {0x00, 0x00, 0x00, 0x00, 0x10, 0x00, 0x80,
                                              0x00, 0x00, 0xAA, 0x00, 0x38, 0x9B, 0x71};
        put16(is_float ? 3 : 1);
        h.insert(h.end(), kGuidTail, kGuidTail + 14);
    }
    if (is_float) { tag4("fact"); put32(4); put32((uint32_t)frames); }
    tag4("data"); put32((uint32_t)data_bytes);
    FileCloser fc{fopen(path, "wb")};
    if (!fc.f) fail("cannot open output");
    if (fwrite(h.data(), 1, h.size(), fc.f) != h.size()) fail("write error");
    align_ = (size_t)align;
    data_bytes_ = data_bytes;
    f_ = fc.f;
    fc.f = nullptr;
}

WavWriter::~WavWriter() {
    if (f_) fclose(static_cast<FILE*>(f_));
}

void WavWriter::write(const void* bytes, size_t n) {
    if (n && fwrite(bytes, 1, n, static_cast<FILE*>(f_)) != n) fail("write error");
}

void WavWriter::finish() {
    FILE* f = static_cast<FILE*>(f_);
    if (data_bytes_ & 1) fputc(0, f);
    f_ = nullptr;
    if (fclose(f)) fail("write error");
}

void write_wav(const char* path, const std::vector<std::vector<double>>& y, int ch, long long frames, int rate,
               const std::string& fmt, bool dither, long long& clipped) {
    WavWriter w(path, ch, frames, rate, fmt);
    const size_t align = w.frame_bytes();
    const int bps = (int)(align / ch);
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-0.5, 0.5);
    clipped = 0;
    const long long chunk = std::max(1, std::min(1 << 16, (int)((4 << 20) / align)));
    std::vector<uint8_t> buf((size_t)chunk * align);
    for (long long f0 = 0; f0 < frames; f0 += chunk) {
        const long long nf = std::min(chunk, frames - f0);
        uint8_t* p = buf.data();
        for (long long i = f0; i < f0 + nf; ++i)
            for (int c = 0; c < ch; ++c, p += bps) {
                const double x = y[c][i];
                if (fmt == "f64") memcpy(p, &x, 8);
                else if (fmt == "f32") { const float v = (float)x; memcpy(p, &v, 4); }
                else if (bps == 2) quantise<2>(p, x, dither, rng, U, clipped);
                else if (bps == 3) quantise<3>(p, x, dither, rng, U, clipped);
                else quantise<4>(p, x, dither, rng, U, clipped);
            }
        w.write(buf.data(), (size_t)(p - buf.data()));
    }
    w.finish();
}

}  // namespace kbsrc::io
