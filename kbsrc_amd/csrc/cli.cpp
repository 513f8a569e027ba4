// kbsrc command-line tool over the GPU engine. The flags, the messages and the exit codes are the reference CLI's
// (app/main.cpp:84-165): 0 on success, 1 for a rejected conversion (no output file is left), 2 for a usage error.
#include <cerrno>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
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

// A command-line mistake: reported with a pointer to --help and exit status 2.
class BadUsage : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class Stopwatch {
public:
    double seconds() const { return std::chrono::duration<double>(Clock::now() - t0_).count(); }

private:
    using Clock = std::chrono::steady_clock;
    Clock::time_point t0_ = Clock::now();
};

// Option values are taken whole: leading blanks are allowed, anything after the number, an int out of range, a real
// that overflows or underflows, or a non-finite real makes the value invalid (std::stoi / std::stod acceptance).
[[noreturn]] void bad_value(const char* flag, const std::string& v) {
    throw BadUsage(std::string("invalid value for ") + flag + ": " + v);
}
int whole_int(const std::string& v, const char* flag) {
    const char* s = v.c_str();
    char* end = nullptr;
    errno = 0;
    const long x = std::strtol(s, &end, 10);
    if (end == s || *end != '\0' || errno == ERANGE || x < INT_MIN || x > INT_MAX) bad_value(flag, v);
    return (int)x;
}
double whole_real(const std::string& v, const char* flag) {
    const char* s = v.c_str();
    char* end = nullptr;
    errno = 0;
    const double x = std::strtod(s, &end);
    if (end == s || *end != '\0' || errno == ERANGE || !std::isfinite(x)) bad_value(flag, v);
    return x;
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
    const Stopwatch clock;
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
    t.read = clock.seconds();
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
    const double t1 = clock.seconds();
    Stream s;
    DevMem d_out((size_t)ch * nout * sizeof(double));
    {
        DevMem d_pad((size_t)ch * xstride * sizeof(double));
        kb_ok(kbsrc_plan_stage_device(plan.p, static_cast<const double*>(d_in.p), static_cast<double*>(d_pad.p), s.s));
        d_in.release();
        kb_ok(kbsrc_plan_execute(plan.p, static_cast<const double*>(d_pad.p), static_cast<double*>(d_out.p), s.s));
        hip_ok(hipStreamSynchronize(s.s), "engine");
    }
    const double t2 = clock.seconds();
    t.resample = t2 - t1;
    clipped = 0;
    if (dither) {   // TPDF dither keeps the reference's mt19937_64 sequence: quantised on the host
        std::vector<double> flat((size_t)ch * nout);
        hip_ok(hipMemcpy(flat.data(), d_out.p, sizeof(double) * flat.size(), hipMemcpyDeviceToHost), "download");
        kbsrc::io::write_wav(out_path, flat.data(), ch, nout, rate, format, true, clipped);
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
    t.write = clock.seconds() - t2;
}

// ---- the command line
struct Request {
    kbsrc::Options opt;
    int rate = 44100;
    std::string format = "f64", in_path, out_path;
    bool dither = false, verbose = false;
    enum { Convert, ShowHelp, ShowVersion } action = Convert;
};

// One flag: its short and long spellings, whether it takes a value, and what it does with it.
struct Flag {
    const char* brief;   // "" when there is none
    const char* full;
    bool takes_value;
    void (*apply)(Request&, const std::string& value);
};

const Flag kFlags[] = {
    {"-r", "--rate", true, [](Request& q, const std::string& v) { q.rate = whole_int(v, "--rate"); }},
    {"-m", "--mode", true,
     [](Request& q, const std::string& v) {
         if (v == "exact") q.opt.precision = kbsrc::Precision::Exact;
         else if (v == "fast") q.opt.precision = kbsrc::Precision::Fast;
         else if (v == "ld") q.opt.precision = kbsrc::Precision::LongDouble;   // refused later, as by the library
         else throw BadUsage("mode must be exact, fast or ld");
     }},
    {"-f", "--format", true, [](Request& q, const std::string& v) { q.format = v; }},
    {"-t", "--threads", true, [](Request& q, const std::string& v) { q.opt.threads = whole_int(v, "--threads"); }},
    {"", "--block", true, [](Request& q, const std::string& v) { q.opt.block = whole_int(v, "--block"); }},
    {"", "--len", true, [](Request& q, const std::string& v) { q.opt.length = whole_real(v, "--len"); }},
    {"", "--beta", true, [](Request& q, const std::string& v) { q.opt.beta = whole_real(v, "--beta"); }},
    {"", "--dither", false, [](Request& q, const std::string&) { q.dither = true; }},
    {"", "--no-lpc", false, [](Request& q, const std::string&) { q.opt.lpc_edges = false; }},
    {"-v", "--verbose", false, [](Request& q, const std::string&) { q.verbose = true; }},
    {"-h", "--help", false, [](Request& q, const std::string&) { q.action = Request::ShowHelp; }},
    {"", "--version", false, [](Request& q, const std::string&) { q.action = Request::ShowVersion; }},
};

const Flag* find_flag(const std::string& a) {
    for (const Flag& f : kFlags)
        if (a == f.full || (f.brief[0] && a == f.brief)) return &f;
    return nullptr;
}

// Arguments are taken left to right; help and version act as soon as they are met.
Request parse(int argc, char** argv) {
    Request q;
    std::vector<std::string> paths;
    for (int i = 1; i < argc && q.action == Request::Convert; ++i) {
        const std::string a = argv[i];
        if (const Flag* f = find_flag(a)) {
            if (f->takes_value && i + 1 >= argc) throw BadUsage("missing value for " + a);
            f->apply(q, f->takes_value ? std::string(argv[++i]) : std::string());
        } else if (a.size() > 1 && a[0] == '-') {
            throw BadUsage("unknown option " + a);
        } else if (a.empty() && paths.size() < 2) {
            // an empty argument names no file and fills no place
        } else if (paths.size() == 2) {
            throw BadUsage("too many arguments");
        } else {
            paths.push_back(a);
        }
    }
    if (q.action != Request::Convert) return q;
    if (paths.size() < 2) throw BadUsage("expected input and output paths");
    q.in_path = paths[0];
    q.out_path = paths[1];
    if (q.rate < 1 || q.rate > 10000000) throw BadUsage("--rate must be between 1 and 10000000");
    if (!kbsrc::io::find_format(q.format)) throw BadUsage("format must be f64, f32, s32, s24 or s16");
    return q;
}

void print_usage() {
    std::printf(
        "kbsrc %s (MI355X engine)\n"
        "Converts the sample rate of a WAV file with a Kaiser-windowed sinc kernel.\n"
        "\n"
        "usage: kbsrc [options] input.wav output.wav\n"
        "\n"
        "options:\n"
        "  -r, --rate N      rate of the output file, Hz; default 44100\n"
        "  -m, --mode M      exact (double-double, the default) or fast (float64)\n"
        "  -f, --format F    output samples: f64 (default), f32, s32, s24 or s16\n"
        "      --dither      add TPDF dither before rounding to an integer format\n"
        "  -t, --threads N   accepted; the GPU engine uses no host worker threads\n"
        "      --len S       kernel half-length in seconds; default 0.2\n"
        "      --beta B      Kaiser window beta; default 40\n"
        "      --block N     block size factor; default chosen from the rates\n"
        "      --no-lpc      pad the signal edges with zeros instead of Burg LPC extension\n"
        "  -v, --verbose     print the plan and the time of each step\n"
        "  -h, --help        print this text\n"
        "      --version     print the version\n",
        kbsrc::version());
}

int run(int argc, char** argv) {
    const Request q = parse(argc, argv);
    if (q.action == Request::ShowHelp) {
        print_usage();
        return 0;
    }
    if (q.action == Request::ShowVersion) {
        std::printf("kbsrc %s\n", kbsrc::version());
        return 0;
    }
    const kbsrc::Options& opt = q.opt;
    const int rate = q.rate;
    const std::string& format = q.format;
    const Stopwatch clock;
    kbsrc::io::WavReader r(q.in_path.c_str());
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
        convert_streamed(r, q.out_path.c_str(), rate, format, q.dither, o, st, nout, clipped, t);
    } else {   // equal rates (a copy) and the refused long-double mode: the library call on the whole file
        std::vector<double> d((size_t)r.frames() * r.channels());
        r.read(d.data(), r.frames());
        t.read = clock.seconds();
        if (r.rate() > 10000000) throw std::runtime_error("input sample rate above 10000000 Hz");
        kbsrc::io::check_wav_output(r.channels(), rate, format);
        const double t1 = clock.seconds();
        kbsrc::Stats s;
        const auto y = kbsrc::resample(d.data(), r.frames(), r.channels(), r.rate(), rate, opt, &s);
        const double t2 = clock.seconds();
        nout = y.empty() ? 0 : (long long)y[0].size();
        std::vector<double> planar;
        planar.reserve((size_t)r.channels() * (size_t)nout);
        for (const auto& c : y) planar.insert(planar.end(), c.begin(), c.end());
        kbsrc::io::write_wav(q.out_path.c_str(), planar.data(), r.channels(), nout, rate, format, q.dither, clipped);
        t.design = s.design_seconds;
        t.resample = t2 - t1 - s.design_seconds;
        t.write = clock.seconds() - t2;
        st.p = s.p; st.q = s.q; st.taps = s.taps; st.block = s.block; st.threads = s.threads;
        st.general_fold = s.general_fold; st.passband = s.passband; st.stopband = s.stopband; st.isa = s.isa;
    }
    if (clipped) std::fprintf(stderr, "kbsrc: warning: %lld samples clipped\n", clipped);
    if (q.verbose) {
        std::fprintf(stderr, "%d Hz -> %d Hz, %d ch, %lld -> %lld frames\n", r.rate(), rate, r.channels(),
                     r.frames(), nout);
        std::fprintf(stderr, "P/Q=%lld/%lld taps=%lld block=%d threads=%d engine=%s%s\n", st.p, st.q, st.taps,
                     st.block, st.threads, st.isa ? st.isa : "gfx950", st.general_fold ? " (general fold)" : "");
        if (st.stopband > 0)
            std::fprintf(stderr, "passband 0-%.1f Hz, stopband from %.1f Hz\n", st.passband, st.stopband);
        // the design runs on the GPU while the file is read: read and design overlap
        std::fprintf(stderr, "read %.3fs  design %.3fs  resample %.3fs  write %.3fs  total %.3fs\n", t.read,
                     t.design, t.resample, t.write, clock.seconds());
    }
    return 0;
}

// Exit status and message of whatever ended the run: usage mistakes 2, every other failure 1.
int report(std::exception_ptr failure) {
    try {
        std::rethrow_exception(failure);
    } catch (const BadUsage& e) {
        std::fprintf(stderr, "kbsrc: %s (see --help)\n", e.what());
        return 2;
    } catch (const std::bad_alloc&) {
        std::fprintf(stderr, "kbsrc: out of memory\n");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "kbsrc: %s\n", e.what());
    }
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (...) {
        return report(std::current_exception());
    }
}
