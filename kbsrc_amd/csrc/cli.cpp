// kbsrc command-line tool over the GPU engine. Flags, messages and exit codes follow the reference CLI
// (app/main.cpp:84-165): 0 on success, 1 for a rejected conversion (no output file), 2 for usage errors.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include <thread>

#include "kbsrc/kbsrc.hpp"
#include "kbsrc_gpu.h"
#include "wav.hpp"

namespace {

struct UsageError : std::exception {
    std::string msg;
    explicit UsageError(std::string m) : msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

int parse_int(const std::string& v, const char* name) {
    try {
        size_t used = 0;
        const int r = std::stoi(v, &used);
        if (used == v.size()) return r;
    } catch (...) {
    }
    throw UsageError(std::string("invalid value for ") + name + ": " + v);
}

double parse_double(const std::string& v, const char* name) {
    try {
        size_t used = 0;
        const double r = std::stod(v, &used);
        if (used == v.size() && std::isfinite(r)) return r;
    } catch (...) {
    }
    throw UsageError(std::string("invalid value for ") + name + ": " + v);
}

// ---- streamed conversion: the file is decoded and uploaded while the GPU designs the kernel, and the output is
// packed on the GPU and written while the next chunk is packed and downloaded (src/io/wav.cpp:150-170,235-250).

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void kb_ok(int rc) {
    if (rc == KBSRC_OK) return;
    const std::string m = kbsrc_last_error();
    if (rc == KBSRC_EINVAL) throw std::invalid_argument(m);
    if (rc == KBSRC_ENOMEM) throw std::bad_alloc();
    throw std::runtime_error(m);
}

struct DevMem {
    void* p = nullptr;
    explicit DevMem(size_t n) { hip_ok(hipMalloc(&p, n ? n : 8), "device allocation"); }
    ~DevMem() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
};

struct Pinned {
    void* p = nullptr;
    explicit Pinned(size_t n) { hip_ok(hipHostMalloc(&p, n ? n : 8, hipHostMallocDefault), "pinned allocation"); }
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
    Pinned(const Pinned&) = delete;
    Pinned& operator=(const Pinned&) = delete;
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream"); }
    ~Stream() { (void)hipStreamDestroy(s); }
};

struct Event {
    hipEvent_t e = nullptr;
    bool used = false;
    Event() { hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event"); }
    ~Event() { (void)hipEventDestroy(e); }
    void record(hipStream_t s) { hip_ok(hipEventRecord(e, s), "event"); used = true; }
    void wait() {
        if (used) hip_ok(hipEventSynchronize(e), "event");
    }
};

struct PlanHandle {
    kbsrc_plan* p = nullptr;
    ~PlanHandle() { kbsrc_plan_destroy(p); }
};

int pack_code(const std::string& f) {
    return f == "f64" ? KBSRC_PACK_F64 : f == "f32" ? KBSRC_PACK_F32 : f == "s32" ? KBSRC_PACK_S32
           : f == "s24" ? KBSRC_PACK_S24 : KBSRC_PACK_S16;
}

struct Times {
    double read = 0, design = 0, resample = 0, write = 0;
};

// The whole conversion for differing rates and a GPU precision. Errors surface in the reference's order: reading,
// the input rate, the output format, then the conversion itself (app/main.cpp:117-140).
void convert_streamed(kbsrc::io::WavReader& r, const char* out_path, int rate, const std::string& format,
                      bool dither, const kbsrc_options& o, kbsrc_stats& st, long long& nout, long long& clipped,
                      Times& t) {
    const double t0 = now();
    const int ch = r.channels();
    const long long frames = r.frames();
    bool early_ok = r.rate() <= 10000000;
    if (early_ok) {
        try {
            kbsrc::io::check_wav_output(ch, rate, format);
        } catch (const std::exception&) {
            early_ok = false;
        }
    }
    PlanHandle plan;
    int plan_rc = KBSRC_OK;
    std::string plan_err;
    std::thread design;
    struct Joiner {
        std::thread& t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } joiner{design};
    if (early_ok)
        design = std::thread([&] {
            plan_rc = kbsrc_plan_create(frames, ch, r.rate(), rate, &o, &plan.p);
            if (plan_rc) plan_err = kbsrc_last_error();
        });
    // decode into pinned chunks and upload them behind the decoder
    const long long n = frames * ch;
    DevMem d_in((size_t)n * sizeof(double));
    {
        Stream cs;
        const long long chunk = std::max<long long>(1, ((long long)1 << 21) / ch);   // frames per chunk (16 MiB)
        Pinned buf0((size_t)chunk * ch * sizeof(double)), buf1((size_t)chunk * ch * sizeof(double));
        Pinned* buf[2] = {&buf0, &buf1};
        Event ev[2];
        int b = 0;
        for (long long f0 = 0; f0 < frames; f0 += chunk, b ^= 1) {
            const long long nf = std::min(chunk, frames - f0);
            ev[b].wait();
            double* h = static_cast<double*>(buf[b]->p);
            r.read(h, nf);
            hip_ok(hipMemcpyAsync(static_cast<double*>(d_in.p) + f0 * ch, h, sizeof(double) * nf * ch,
                                  hipMemcpyHostToDevice, cs.s), "input upload");
            ev[b].record(cs.s);
        }
        hip_ok(hipStreamSynchronize(cs.s), "input upload");
    }
    t.read = now() - t0;
    if (r.rate() > 10000000) throw std::runtime_error("input sample rate above 10000000 Hz");
    kbsrc::io::check_wav_output(ch, rate, format);
    design.join();
    if (plan_rc) {
        if (plan_rc == KBSRC_EINVAL) throw std::invalid_argument(plan_err);
        if (plan_rc == KBSRC_ENOMEM) throw std::bad_alloc();
        throw std::runtime_error(plan_err);
    }
    kb_ok(kbsrc_plan_stats(plan.p, &st));
    t.design = st.design_seconds;
    long long v[KBSRC_PLAN_FIELDS];
    kb_ok(kbsrc_plan_query(plan.p, v));
    nout = v[2];
    const long long xstride = v[22];
    const double t1 = now();
    Stream s;
    DevMem d_out((size_t)ch * nout * sizeof(double));
    {
        DevMem d_pad((size_t)ch * xstride * sizeof(double));
        kb_ok(kbsrc_plan_stage_device(plan.p, static_cast<const double*>(d_in.p), static_cast<double*>(d_pad.p), s.s));
        d_in.release();
        kb_ok(kbsrc_plan_execute(plan.p, static_cast<const double*>(d_pad.p), static_cast<double*>(d_out.p), s.s));
        hip_ok(hipStreamSynchronize(s.s), "engine");
    }
    const double t2 = now();
    t.resample = t2 - t1;
    clipped = 0;
    if (dither) {   // TPDF dither keeps the reference's mt19937_64 sequence: quantised on the host
        std::vector<double> flat((size_t)ch * nout);
        hip_ok(hipMemcpy(flat.data(), d_out.p, sizeof(double) * flat.size(), hipMemcpyDeviceToHost), "download");
        std::vector<std::vector<double>> y(ch);
        for (int c = 0; c < ch; ++c) y[c].assign(flat.begin() + (size_t)c * nout, flat.begin() + (size_t)(c + 1) * nout);
        kbsrc::io::write_wav(out_path, y, ch, nout, rate, format, true, clipped);
    } else {
        kbsrc::io::WavWriter w(out_path, ch, nout, rate, format);
        const size_t fb = w.frame_bytes();
        const long long chunk = std::max<long long>(1, ((long long)4 << 20) / (long long)fb);   // 4 MiB of output
        DevMem dbytes0(chunk * fb), dbytes1(chunk * fb), dclip(sizeof(unsigned long long));
        Pinned hb0(chunk * fb), hb1(chunk * fb);
        void* dbytes[2] = {dbytes0.p, dbytes1.p};
        void* hb[2] = {hb0.p, hb1.p};
        size_t len[2] = {0, 0};
        Event ev[2];
        hip_ok(hipMemsetAsync(dclip.p, 0, sizeof(unsigned long long), s.s), "clip counter");
        const int fmt = pack_code(format);
        int b = 0;
        for (long long f0 = 0; f0 < nout; f0 += chunk, b ^= 1) {
            const long long nf = std::min(chunk, nout - f0);
            kb_ok(kbsrc_pack(static_cast<const double*>(d_out.p), ch, nout, f0, nf, fmt, dbytes[b],
                             static_cast<unsigned long long*>(dclip.p), s.s));
            len[b] = (size_t)nf * fb;
            hip_ok(hipMemcpyAsync(hb[b], dbytes[b], len[b], hipMemcpyDeviceToHost, s.s), "output download");
            ev[b].record(s.s);
            if (f0 > 0) {   // write the previous chunk while this one is packed and downloaded
                ev[b ^ 1].wait();
                w.write(hb[b ^ 1], len[b ^ 1]);
            }
        }
        if (nout > 0) {
            ev[b ^ 1].wait();
            w.write(hb[b ^ 1], len[b ^ 1]);
        }
        unsigned long long c = 0;
        hip_ok(hipMemcpy(&c, dclip.p, sizeof c, hipMemcpyDeviceToHost), "clip counter");
        clipped = (long long)c;
        w.finish();
    }
    t.write = now() - t2;
}

void help() {
    std::printf(
        "kbsrc %s - Kaiser-Bessel sample-rate converter (MI355X engine)\n"
        "\n"
        "usage: kbsrc [options] input.wav output.wav\n"
        "\n"
        "  -r, --rate N        output sample rate in Hz (default 44100)\n"
        "  -m, --mode M        exact (default) or fast\n"
        "  -f, --format F      f64 (default), f32, s32, s24 or s16\n"
        "      --dither        TPDF dither for integer formats\n"
        "  -t, --threads N     accepted for compatibility (the GPU engine has no host workers)\n"
        "      --len S         one-sided kernel length in seconds (default 0.2)\n"
        "      --beta B        Kaiser beta (default 40)\n"
        "      --block N       block size factor (default: automatic)\n"
        "      --no-lpc        zero-pad edges instead of LPC extrapolation\n"
        "  -v, --verbose       print plan and timing\n"
        "  -h, --help          show this help\n"
        "      --version       show version\n",
        kbsrc::version());
}

int run(int argc, char** argv) {
    kbsrc::Options opt;
    int rate = 44100;
    std::string format = "f64", in_path, out_path;
    bool dither = false, verbose = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&]() -> std::string {
            if (i + 1 >= argc) throw UsageError("missing value for " + a);
            return argv[++i];
        };
        if (a == "-r" || a == "--rate") rate = parse_int(value(), "--rate");
        else if (a == "-m" || a == "--mode") {
            const std::string m = value();
            if (m == "exact") opt.precision = kbsrc::Precision::Exact;
            else if (m == "fast") opt.precision = kbsrc::Precision::Fast;
            else if (m == "ld") opt.precision = kbsrc::Precision::LongDouble;
            else throw UsageError("mode must be exact, fast or ld");
        } else if (a == "-f" || a == "--format") format = value();
        else if (a == "-t" || a == "--threads") opt.threads = parse_int(value(), "--threads");
        else if (a == "--block") opt.block = parse_int(value(), "--block");
        else if (a == "--len") opt.length = parse_double(value(), "--len");
        else if (a == "--beta") opt.beta = parse_double(value(), "--beta");
        else if (a == "--dither") dither = true;
        else if (a == "--no-lpc") opt.lpc_edges = false;
        else if (a == "-v" || a == "--verbose") verbose = true;
        else if (a == "-h" || a == "--help") { help(); return 0; }
        else if (a == "--version") { std::printf("kbsrc %s\n", kbsrc::version()); return 0; }
        else if (a.size() > 1 && a[0] == '-') throw UsageError("unknown option " + a);
        else if (in_path.empty()) in_path = a;
        else if (out_path.empty()) out_path = a;
        else throw UsageError("too many arguments");
    }
    if (in_path.empty() || out_path.empty()) throw UsageError("expected input and output paths");
    if (rate < 1 || rate > 10000000) throw UsageError("--rate must be between 1 and 10000000");
    if (format != "f64" && format != "f32" && format != "s32" && format != "s24" && format != "s16")
        throw UsageError("format must be f64, f32, s32, s24 or s16");

    const double t0 = now();
    kbsrc::io::WavReader r(in_path.c_str());
    long long nout = 0, clipped = 0;
    Times t;
    kbsrc_stats st{};
    if (r.rate() != rate && opt.precision != kbsrc::Precision::LongDouble) {
        kbsrc_options o;
        kbsrc_default_options(&o);
        o.precision = opt.precision == kbsrc::Precision::Fast ? KBSRC_FAST : KBSRC_EXACT;
        o.length = opt.length;
        o.beta = opt.beta;
        o.threads = opt.threads;
        o.block = opt.block;
        o.lpc_edges = opt.lpc_edges ? 1 : 0;
        convert_streamed(r, out_path.c_str(), rate, format, dither, o, st, nout, clipped, t);
    } else {   // equal rates (a copy) and the rejected long-double mode: the library call on the whole file
        std::vector<double> d((size_t)r.frames() * r.channels());
        r.read(d.data(), r.frames());
        t.read = now() - t0;
        if (r.rate() > 10000000) throw std::runtime_error("input sample rate above 10000000 Hz");
        kbsrc::io::check_wav_output(r.channels(), rate, format);
        const double t1 = now();
        kbsrc::Stats s;
        auto y = kbsrc::resample(d.data(), r.frames(), r.channels(), r.rate(), rate, opt, &s);
        const double t2 = now();
        nout = y.empty() ? 0 : (long long)y[0].size();
        kbsrc::io::write_wav(out_path.c_str(), y, r.channels(), nout, rate, format, dither, clipped);
        t.design = s.design_seconds;
        t.resample = t2 - t1 - s.design_seconds;
        t.write = now() - t2;
        st.p = s.p; st.q = s.q; st.taps = s.taps; st.block = s.block; st.threads = s.threads;
        st.general_fold = s.general_fold; st.passband = s.passband; st.stopband = s.stopband; st.isa = s.isa;
    }
    const double t3 = now();
    if (clipped) std::fprintf(stderr, "kbsrc: warning: %lld samples clipped\n", clipped);
    if (verbose) {
        std::fprintf(stderr, "%d Hz -> %d Hz, %d ch, %lld -> %lld frames\n", r.rate(), rate, r.channels(),
                     r.frames(), nout);
        std::fprintf(stderr, "P/Q=%lld/%lld taps=%lld block=%d threads=%d engine=%s%s\n", st.p, st.q, st.taps,
                     st.block, st.threads, st.isa ? st.isa : "gfx950", st.general_fold ? " (general fold)" : "");
        if (st.stopband > 0)
            std::fprintf(stderr, "passband 0-%.1f Hz, stopband from %.1f Hz\n", st.passband, st.stopband);
        // the design runs on the GPU while the file is read: read and design overlap
        std::fprintf(stderr, "read %.3fs  design %.3fs  resample %.3fs  write %.3fs  total %.3fs\n", t.read,
                     t.design, t.resample, t.write, t3 - t0);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const UsageError& e) {
        std::fprintf(stderr, "kbsrc: %s (see --help)\n", e.what());
        return 2;
    } catch (const std::bad_alloc&) {
        std::fprintf(stderr, "kbsrc: out of memory\n");
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "kbsrc: %s\n", e.what());
        return 1;
    }
}
