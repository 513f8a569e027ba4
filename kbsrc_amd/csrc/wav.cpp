// WAV reading and writing for the command-line tool. The bytes written and the inputs refused (with their messages,
// checked in the same order) are the reference's (src/io/wav.cpp:109-253); undithered s16/s24/s32 and f32/f64 files
// are byte-identical to its output.
#include "wav.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>

#pragma clang fp contract(off)

namespace kbsrc::io {
namespace {

[[noreturn]] void fail(const char* m) { throw std::runtime_error(m); }

constexpr SampleFormat kFormats[] = {
    {"f64", 64, true}, {"f32", 32, true}, {"s32", 32, false}, {"s24", 24, false}, {"s16", 16, false},
};

// little-endian field of n bytes at p
uint64_t le(const uint8_t* p, int n) {
    uint64_t v = 0;
    while (n-- > 0) v = v << 8 | p[n];
    return v;
}

long long file_size(std::FILE* f) {
    fseeko(f, 0, SEEK_END);
    const long long n = ftello(f);
    fseeko(f, 0, SEEK_SET);
    return n;
}

// ---- decoders: n samples of one encoding to float64. Integer PCM is scaled by 2^-(8B-1) (8-bit PCM is unsigned
// around 128), which is exact.
template <int B> void decode_pcm(const uint8_t* src, size_t n, double* dst) {
    constexpr double unit = 1.0 / (double)(1ull << (8 * B - 1));
    for (size_t k = 0; k < n; ++k, src += B) {
        const uint32_t u = (uint32_t)le(src, B);
        const int32_t v = B == 1 ? (int32_t)u - 128 : (int32_t)(u << (32 - 8 * B)) >> (32 - 8 * B);
        dst[k] = v * unit;
    }
}
void decode_f32(const uint8_t* src, size_t n, double* dst) {
    for (size_t k = 0; k < n; ++k) {
        float v;
        std::memcpy(&v, src + 4 * k, 4);
        dst[k] = v;
    }
}
void decode_f64(const uint8_t* src, size_t n, double* dst) { std::memcpy(dst, src, 8 * n); }

// What the fmt chunk says (the last one before the data chunk counts).
struct FmtFields {
    bool seen = false;
    unsigned tag = 0, channels = 0, block_align = 0, bits = 0;
    uint32_t rate = 0;
};

// Walks the RIFF chunks after the 12-byte RIFF/WAVE header: fmt fields, and the data chunk's offset and length (a
// declared length of 0 or 0xFFFFFFFF means "to the end of the file", as a streamed writer leaves it).
struct Layout {
    FmtFields fmt;
    long long data_at = -1, data_len = 0;
};
Layout walk_chunks(std::FILE* f, long long size) {
    Layout L;
    for (long long at = 12; at + 8 <= size;) {
        uint8_t head[8];
        fseeko(f, at, SEEK_SET);
        if (std::fread(head, 1, 8, f) != 8) break;
        const uint32_t len = (uint32_t)le(head + 4, 4);
        const long long body = at + 8;
        if (std::memcmp(head, "fmt ", 4) == 0) {
            uint8_t b[40] = {};
            const size_t got = std::fread(b, 1, std::min<uint32_t>(len, sizeof b), f);
            if (got < 16) fail("bad fmt chunk");
            FmtFields& m = L.fmt;
            m.seen = true;
            m.tag = (unsigned)le(b, 2);
            m.channels = (unsigned)le(b + 2, 2);
            m.rate = (uint32_t)le(b + 4, 4);
            m.block_align = (unsigned)le(b + 12, 2);
            m.bits = (unsigned)le(b + 14, 2);
            if (m.tag == 0xFFFE && got >= 26) m.tag = (unsigned)le(b + 24, 2);   // the sub-format's tag
        } else if (std::memcmp(head, "data", 4) == 0) {
            L.data_at = body;
            L.data_len = (len == 0 || len == 0xFFFFFFFFu) ? size - body : std::min<long long>(len, size - body);
            break;
        }
        at = body + len + (len & 1);
    }
    return L;
}

// ---- header of an output file, assembled chunk by chunk
class Bytes {
public:
    Bytes& u16(unsigned v) { return put(v, 2); }
    Bytes& u32(uint32_t v) { return put(v, 4); }
    Bytes& id(const char* four) {
        b_.insert(b_.end(), four, four + 4);
        return *this;
    }
    Bytes& raw(const uint8_t* p, size_t n) {
        b_.insert(b_.end(), p, p + n);
        return *this;
    }
    Bytes& chunk(const char* four, const Bytes& payload) {
        id(four).u32((uint32_t)payload.size());
        return raw(payload.b_.data(), payload.size());
    }
    size_t size() const { return b_.size(); }
    const std::vector<uint8_t>& bytes() const { return b_; }

private:
    Bytes& put(uint64_t v, int n) {
        for (int i = 0; i < n; ++i) b_.push_back((uint8_t)(v >> (8 * i)));
        return *this;
    }
    std::vector<uint8_t> b_;
};

// KSDATAFORMAT_SUBTYPE_{PCM,IEEE_FLOAT}: the format tag in the first two bytes, then this fixed GUID tail
constexpr uint8_t kSubtypeTail[14] = {0x00, 0x00, 0x00, 0x00, 0x10, 0x00, 0x80,
                                      0x00, 0x00, 0xAA, 0x00, 0x38, 0x9B, 0x71};

std::vector<uint8_t> header(const SampleFormat& f, int ch, int rate, long long frames) {
    const unsigned align = (unsigned)ch * f.bytes();
    const uint32_t data = (uint32_t)((unsigned long long)frames * align);
    // WAVE_FORMAT_EXTENSIBLE for more than two channels and for integer samples wider than 16 bits; otherwise the
    // 16-byte PCM fmt, or the 18-byte float fmt with its fact chunk
    const bool ext = ch > 2 || (!f.floating && f.bits > 16);
    Bytes fmt;
    fmt.u16(ext ? 0xFFFE : f.floating ? 3 : 1).u16(ch).u32(rate).u32((uint32_t)rate * align).u16(align).u16(f.bits);
    if (ext) {
        const uint32_t mask = ch == 1 ? 0x4 : ch == 2 ? 0x3 : 0;   // front centre; front left and right
        fmt.u16(22).u16(f.bits).u32(mask).u16(f.floating ? 3 : 1).raw(kSubtypeTail, sizeof kSubtypeTail);
    } else if (f.floating) {
        fmt.u16(0);
    }
    Bytes body;
    body.id("WAVE").chunk("fmt ", fmt);
    if (f.floating) {
        Bytes fact;
        fact.u32((uint32_t)frames);
        body.chunk("fact", fact);
    }
    body.id("data").u32(data);
    Bytes file;
    file.id("RIFF").u32((uint32_t)(body.size() + data + (data & 1))).raw(body.bytes().data(), body.size());
    return file.bytes();
}

// One integer sample: round in the caller's mode, clip to [-2^(bits-1), 2^(bits-1) - 1], little-endian bytes.
struct IntegerSample {
    int bytes;
    double full;   // 2^(bits-1)
    void store(uint8_t* p, double v, long long& clipped) const {
        v = std::nearbyint(v);
        if (v > full - 1) {
            v = full - 1;
            ++clipped;
        } else if (v < -full) {
            v = -full;
            ++clipped;
        }
        const uint64_t q = v == v ? (uint64_t)(int64_t)v : 0;   // NaN is written as zero
        for (int i = 0; i < bytes; ++i) p[i] = (uint8_t)(q >> (8 * i));
    }
};

}  // namespace

const SampleFormat* find_format(const std::string& name) {
    for (const SampleFormat& f : kFormats)
        if (name == f.name) return &f;
    return nullptr;
}

WavReader::WavReader(const char* path) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) fail("cannot open input");
    file_ = f;   // closed by the destructor, also when a check below throws
    struct CloseOnThrow {
        WavReader* r;
        bool armed = true;
        ~CloseOnThrow() {
            if (armed && r->file_) {
                std::fclose(r->file_);
                r->file_ = nullptr;
            }
        }
    } guard{this};
    const long long size = file_size(f);
    uint8_t riff[12];
    if (size < 12 || std::fread(riff, 1, 12, f) != 12 || std::memcmp(riff, "RIFF", 4) ||
        std::memcmp(riff + 8, "WAVE", 4))
        fail("not a RIFF/WAVE file");
    const Layout L = walk_chunks(f, size);
    const FmtFields& m = L.fmt;
    if (!m.seen || L.data_at < 0) fail("missing fmt/data chunk");
    if (m.channels == 0) fail("invalid WAV channel count");
    if (m.rate == 0 || m.rate > (uint32_t)INT_MAX) fail("invalid WAV sample rate");
    if (m.block_align == 0 || m.block_align % m.channels) fail("invalid WAV block alignment");
    const bool pcm = m.tag == 1, ieee = m.tag == 3 && (m.bits == 32 || m.bits == 64);
    if (!pcm && !ieee) fail("unsupported WAV format (need PCM or IEEE float)");
    // a PCM sample takes the fewest whole bytes that hold its bits (12-bit audio in 2 bytes decodes as 16-bit); a
    // float sample exactly 4 or 8
    const int width = (int)(m.block_align / m.channels);
    if (pcm && (width < 1 || width > 4 || (int)(m.bits + 7) / 8 != width)) fail("unsupported PCM width");
    if (ieee && (int)m.bits != 8 * width) fail("IEEE float sample size does not match the block alignment");
    static constexpr Decoder kPcm[4] = {decode_pcm<1>, decode_pcm<2>, decode_pcm<3>, decode_pcm<4>};
    decode_ = ieee ? (width == 8 ? decode_f64 : decode_f32) : kPcm[width - 1];
    rate_ = (int)m.rate;
    ch_ = (int)m.channels;
    sample_bytes_ = width;
    frames_ = L.data_len / m.block_align;
    fseeko(f, L.data_at, SEEK_SET);
    guard.armed = false;
}

WavReader::~WavReader() {
    if (file_) std::fclose(file_);
}

void WavReader::read(double* dst, long long frames) {
    if (frames < 0 || frames > frames_ - consumed_) fail("read error");
    constexpr size_t kBlock = (size_t)1 << 18;   // samples decoded per file read
    if (raw_.empty()) raw_.resize(kBlock * (size_t)sample_bytes_);
    const size_t total = (size_t)frames * (size_t)ch_;
    for (size_t done = 0; done < total;) {
        const size_t n = std::min(kBlock, total - done);
        if (std::fread(raw_.data(), (size_t)sample_bytes_, n, file_) != n) fail("read error");
        decode_(raw_.data(), n, dst + done);
        done += n;
    }
    consumed_ += frames;
}

void check_wav_output(int ch, int rate, const std::string& format) {
    const SampleFormat* f = find_format(format);
    const long long frame = (long long)ch * (f ? f->bytes() : 4);
    if (ch < 1 || frame > 0xFFFF) fail("too many channels for a WAV file in this format");
    if (rate < 1 || rate * frame > 0xFFFFFFFFLL) fail("sample rate too high for a WAV file with this many channels");
}

WavWriter::WavWriter(const char* path, int ch, long long frames, int rate, const std::string& format) {
    check_wav_output(ch, rate, format);
    const SampleFormat* f = find_format(format);
    if (!f) fail("unsupported output format");
    const long long align = (long long)ch * f->bytes();
    // the RIFF size field holds the whole file less 8 bytes, header included (at most 80 bytes)
    if (frames < 0 || frames > (0xFFFFFFFFLL - 80) / align) fail("output too large for WAV (4 GB limit)");
    const std::vector<uint8_t> h = header(*f, ch, rate, frames);
    std::FILE* out = std::fopen(path, "wb");
    if (!out) fail("cannot open output");
    file_ = out;
    frame_bytes_ = (size_t)align;
    data_bytes_ = (unsigned long long)frames * (unsigned long long)align;
    write(h.data(), h.size());
}

WavWriter::~WavWriter() {
    if (file_) std::fclose(file_);
}

void WavWriter::write(const void* bytes, size_t n) {
    if (n && std::fwrite(bytes, 1, n, file_) != n) fail("write error");
}

void WavWriter::finish() {
    if (data_bytes_ & 1) std::fputc(0, file_);
    std::FILE* f = file_;
    file_ = nullptr;
    if (std::fclose(f) != 0) fail("write error");
}

void write_wav(const char* path, const double* y, int ch, long long frames, int rate, const std::string& format,
               bool dither, long long& clipped) {
    WavWriter w(path, ch, frames, rate, format);
    const SampleFormat& f = *find_format(format);
    const size_t fb = w.frame_bytes();
    const long long per_write = std::max<long long>(1, (long long)((1u << 20) / fb));   // about 1 MiB per write
    std::vector<uint8_t> buf((size_t)per_write * fb);
    const IntegerSample q{f.bytes(), std::ldexp(1.0, f.bits - 1)};
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> tpdf(-0.5, 0.5);
    clipped = 0;
    for (long long f0 = 0; f0 < frames; f0 += per_write) {
        const long long nf = std::min(per_write, frames - f0);
        uint8_t* p = buf.data();
        for (long long i = f0; i < f0 + nf; ++i)
            for (int c = 0; c < ch; ++c, p += f.bytes()) {
                const double x = y[(size_t)c * (size_t)frames + (size_t)i];
                if (f.floating && f.bits == 64) {
                    std::memcpy(p, &x, 8);
                } else if (f.floating) {
                    const float v = (float)x;
                    std::memcpy(p, &v, 4);
                } else {
                    double v = x * q.full;
                    if (dither) {
                        const double d1 = tpdf(rng), d2 = tpdf(rng);
                        v = v + (d1 + d2);
                    }
                    q.store(p, v, clipped);
                }
            }
        w.write(buf.data(), (size_t)(p - buf.data()));
    }
    w.finish();
}

}  // namespace kbsrc::io
