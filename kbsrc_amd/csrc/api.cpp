// C-ABI entry points (include/kbsrc_gpu.h) and the C++ API (include/kbsrc/kbsrc.hpp) over the GPU engine.
#include <xmmintrin.h>

#include <chrono>
#include <cmath>
#include <climits>
#include <numeric>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"
#include "kbsrc/kbsrc.hpp"
#include "kbsrc_gpu.h"

#pragma clang fp contract(off)

namespace {

using namespace kbsrc::gpu;

thread_local std::string g_error;

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// Round-to-nearest for the host-side dd work of Exact; the caller's MXCSR rounding bits come back on exit
// (the reference's guard: src/dispatch.cpp:37-67). Exception flags and FTZ/DAZ are left alone.
class NearestRounding {
public:
    explicit NearestRounding(bool active) {
        if (!active) return;
        saved_ = _mm_getcsr() & kRC;
        if (saved_) _mm_setcsr(_mm_getcsr() & ~kRC);
    }
    ~NearestRounding() {
        if (saved_) _mm_setcsr((_mm_getcsr() & ~kRC) | saved_);
    }
    NearestRounding(const NearestRounding&) = delete;
    NearestRounding& operator=(const NearestRounding&) = delete;

private:
    static constexpr unsigned kRC = 0x6000;
    unsigned saved_ = 0;
};

kbsrc_options defaults() {
    kbsrc_options o;
    o.precision = KBSRC_EXACT;
    o.length = 0.2;
    o.beta = 40.0;
    o.threads = 0;
    o.block = 0;
    o.lpc_edges = 1;
    return o;
}

void check_options(const kbsrc_options& o) {
    if (o.precision != KBSRC_FAST && o.precision != KBSRC_EXACT) throw std::invalid_argument("option out of range");
    if (o.threads < 0 || o.block < 0) throw std::invalid_argument("option out of range");
}

template <class F> int guarded(F&& f) {
    try {
        f();
        return KBSRC_OK;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return KBSRC_EINVAL;
    } catch (const std::bad_alloc&) {
        g_error = "out of memory";
        return KBSRC_ENOMEM;
    } catch (const std::exception& e) {
        g_error = e.what();
        return KBSRC_ERUNTIME;
    } catch (...) {
        g_error = "unknown error";
        return KBSRC_ERUNTIME;
    }
}

}  // namespace

struct kbsrc_plan {
    Plan p;
    kbsrc_options opt;
    std::unique_ptr<Engine> eng;
    int Jg = 0;
    double design_seconds = 0;
};

namespace {

// Planning (host) and design (taps and spectrum on the GPU) (src/resampler.cpp:92-167,187-206).
std::unique_ptr<kbsrc_plan> make(long long frames, int channels, int rate_in, int rate_out, const kbsrc_options& o,
                                 hipStream_t s) {
    check_options(o);
    const double t0 = now();
    auto pl = std::make_unique<kbsrc_plan>();
    pl->opt = o;
    pl->p = make_plan(frames, channels, rate_in, rate_out, o.precision == KBSRC_FAST, o.length, o.beta, o.block);
    const Plan& p = pl->p;
    if (p.equal) return pl;
    DeviceConsts k;
    k.init(s);
    DevBuf H;   // kernel spectrum, (hi, lo) pairs
    long long nH = 0;
    if (p.gen) {
        // near-unity ratios: the kernel on a finer grid Nk, its true response folded modulo Mb in the engine
        // (src/resampler.cpp:190-202)
        long long Nk = (long long)std::max(p.Nb, p.Mb) * 5 / 4;
        Nk += Nk & 1;
        while (!smooth7(Nk / 2)) Nk += 2;
        const __int128 A = (__int128)p.c2 * p.Nb, B = (__int128)Nk * p.fi;
        const long long g2 = (long long)std::gcd((unsigned long long)p.c2 * (unsigned long long)p.Nb,
                                                 (unsigned long long)Nk * (unsigned long long)p.fi);
        const double fk2 = (double)Nk * p.fi / p.Nb;
        const long long L2 = (long long)(int)std::floor(o.length * fk2);
        const DevBuf h2 = kernel_taps_gpu(L2, (long long)(A / g2), (long long)(B / g2), o.beta, s);
        pl->Jg = (int)std::min<long long>(Nk / 2, (long long)std::ceil((p.mn / 2.0 + 4 * p.fml) * p.Nb / p.fi));
        H = kernel_spectrum_gpu(h2.d(), L2, (int)Nk, (double)p.Nb, k, s);
        nH = Nk / 2 + 1;
    } else {
        const DevBuf h = kernel_taps_gpu(p.Lk, p.c2 / p.gc, p.fk / p.gc, o.beta, s);
        const int n = p.up ? p.Mb : p.Nb;
        H = kernel_spectrum_gpu(h.d(), p.Lk, n, (double)p.Nb, k, s);
        nH = n / 2 + 1;
    }
    pl->eng = std::make_unique<Engine>(p, H.d(), nH, pl->Jg, s);   // keeps its own copy of H
    pl->design_seconds = now() - t0;
    return pl;
}

// Staging (src/resampler.cpp:168-186) from the interleaved input in device memory (d_in) and its first and last
// edge_frames() frames on the host: Burg extensions on the host, then deinterleave, pad and place on the GPU.
void stage_dev(kbsrc_plan* pl, const double* d_in, const double* first, const double* last, double* d_padded,
               hipStream_t s) {
    const Plan& p = pl->p;
    const long long ext = p.E + 32;
    std::vector<double> head((size_t)p.ch * ext, 0.0), tail((size_t)p.ch * ext, 0.0);
    if (pl->opt.lpc_edges) {
        for (int c = 0; c < p.ch; ++c) {
            std::vector<double> hd, tl;
            edge_extensions(p, first, last, c, hd, tl);
            std::memcpy(head.data() + (size_t)c * ext, hd.data(), sizeof(double) * ext);
            std::memcpy(tail.data() + (size_t)c * ext, tl.data(), sizeof(double) * ext);
        }
    }
    DevBuf dh, dt;
    dh.upload(head.data(), sizeof(double) * head.size(), s);
    dt.upload(tail.data(), sizeof(double) * tail.size(), s);
    StageArgs a;
    a.in = d_in; a.head = dh.d(); a.tail = dt.d(); a.xx = d_padded;
    a.nin = p.nin; a.tot = p.tot; a.stride = p.xstride; a.off = p.base + p.pre; a.ext = ext; a.ch = p.ch;
    hip_check(launch_stage(a, s), "staging kernel");
    hip_check(hipStreamSynchronize(s), "staging sync");
}

// Staging from host interleaved input: one upload, then as above.
void stage(kbsrc_plan* pl, const double* in, double* d_padded, hipStream_t s) {
    const Plan& p = pl->p;
    DevBuf din;
    if (p.nin > 0) din.upload(in, sizeof(double) * (size_t)(p.nin * p.ch), s);
    else din.alloc(sizeof(double));
    const long long m = edge_frames(p);
    stage_dev(pl, din.d(), in, in + (p.nin - m) * p.ch, d_padded, s);
}

// Staging from device interleaved input: the edge frames come down to the host for the Burg recursion.
void stage_device(kbsrc_plan* pl, const double* d_in, double* d_padded, hipStream_t s) {
    const Plan& p = pl->p;
    const long long m = edge_frames(p);
    std::vector<double> first((size_t)std::max(1LL, m * p.ch)), last((size_t)std::max(1LL, m * p.ch));
    if (m > 0 && pl->opt.lpc_edges) {
        const size_t bytes = sizeof(double) * (size_t)(m * p.ch);
        hip_check(hipMemcpyAsync(first.data(), d_in, bytes, hipMemcpyDeviceToHost, s), "edge download");
        hip_check(hipMemcpyAsync(last.data(), d_in + (p.nin - m) * p.ch, bytes, hipMemcpyDeviceToHost, s),
                  "edge download");
        hip_check(hipStreamSynchronize(s), "edge sync");
    }
    stage_dev(pl, d_in, first.data(), last.data(), d_padded, s);
}

void fill_stats(const kbsrc_plan& pl, double process_seconds, kbsrc_stats* st) {
    if (!st) return;
    const Plan& p = pl.p;
    st->p = p.P;
    st->q = p.Q;
    st->taps = p.equal ? 0 : 2 * p.Lk + 1;
    st->block = p.nb;
    st->threads = 1;
    st->general_fold = pl.Jg != 0;
    st->design_seconds = pl.design_seconds;
    st->process_seconds = process_seconds;
    st->passband = p.equal ? 0 : (double)p.c2 / 2 - p.fml;
    st->stopband = p.equal ? 0 : (double)p.c2 / 2 + p.fml;
    st->isa = "gfx950";
}

// The whole conversion, host buffers in and out.
void resample_host(const double* in, long long frames, int channels, int rate_in, int rate_out,
                   const kbsrc_options& o, double* out, kbsrc_stats* st) {
    if (channels < 1 || frames < 0 || (frames > 0 && !in)) throw std::invalid_argument("bad input buffer");
    NearestRounding rn(o.precision != KBSRC_FAST);
    const double t0 = now();
    if (channels >= 1 && rate_in >= 1 && rate_out >= 1 && rate_in == rate_out) {
        // equal rates: the input is copied (src/resampler.cpp:106-111); options are still validated
        check_options(o);
        make_plan(frames, channels, rate_in, rate_out, o.precision == KBSRC_FAST, o.length, o.beta, o.block);
        for (int c = 0; c < channels; ++c)
            for (long long i = 0; i < frames; ++i) out[(size_t)c * frames + i] = in[i * channels + c];
        if (st) {
            kbsrc_plan pl;
            pl.p.P = pl.p.Q = 1;
            pl.p.equal = true;
            fill_stats(pl, 0, st);
            st->design_seconds = now() - t0;
        }
        return;
    }
    hipStream_t s = nullptr;
    hip_check(hipStreamCreate(&s), "stream");
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    auto pl = make(frames, channels, rate_in, rate_out, o, s);
    const Plan& p = pl->p;
    const double t1 = now();
    DevBuf xx((size_t)p.ch * p.xstride * sizeof(double)), dout((size_t)std::max(1LL, p.ch * p.nout) * sizeof(double));
    stage(pl.get(), in, xx.d(), s);
    pl->design_seconds += now() - t1;
    const double t2 = now();
    pl->eng->execute(xx.d(), dout.d(), s);
    if (p.ch * p.nout > 0)
        hip_check(hipMemcpyAsync(out, dout.d(), sizeof(double) * p.ch * p.nout, hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "engine sync");
    fill_stats(*pl, now() - t2, st);
}

}  // namespace

extern "C" {

void kbsrc_default_options(kbsrc_options* opt) {
    if (opt) *opt = defaults();
}

const char* kbsrc_version(void) { return "1.3.0-gfx950"; }

const char* kbsrc_last_error(void) { return g_error.c_str(); }

int kbsrc_output_frames(long long frames, int rate_in, int rate_out, long long* nout) {
    return guarded([&] {
        if (frames < 0) throw std::invalid_argument("bad input buffer");
        if (rate_in < 1 || rate_out < 1) throw std::invalid_argument("sample rates must be positive");
        const Plan p = make_plan(frames, 1, rate_in, rate_out, false, 0.2, 40.0, 0);
        *nout = p.nout;
    });
}

int kbsrc_resample(const double* in, long long frames, int channels, int rate_in, int rate_out,
                   const kbsrc_options* opt, double* out, kbsrc_stats* stats) {
    return guarded([&] { resample_host(in, frames, channels, rate_in, rate_out, opt ? *opt : defaults(), out, stats); });
}

int kbsrc_plan_create(long long frames, int channels, int rate_in, int rate_out, const kbsrc_options* opt,
                      kbsrc_plan** plan) {
    return guarded([&] {
        if (!plan) throw std::invalid_argument("bad input buffer");
        const kbsrc_options o = opt ? *opt : defaults();
        NearestRounding rn(o.precision != KBSRC_FAST);
        check_options(o);   // argument errors before any device call
        (void)make_plan(frames, channels, rate_in, rate_out, o.precision == KBSRC_FAST, o.length, o.beta, o.block);
        // the design runs on a stream of its own that does not wait for the legacy default stream, so a caller
        // may copy input on other streams meanwhile
        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{s};
        auto pl = make(frames, channels, rate_in, rate_out, o, s);
        hip_check(hipStreamSynchronize(s), "plan sync");
        *plan = pl.release();
    });
}

int kbsrc_plan_query(const kbsrc_plan* pl, long long* v) {
    return guarded([&] {
        if (!pl || !v) throw std::invalid_argument("bad input buffer");
        const Plan& p = pl->p;
        const long long a[KBSRC_PLAN_FIELDS] = {p.P, p.Q, p.nout, p.nb, p.marg, p.Nb, p.Mb, p.hop, p.nblk, p.o, p.tot,
                                                p.base, p.pre, p.E, p.L, p.Lk, p.c2, p.gen, p.up, p.equal,
                                                pl->eng ? pl->eng->tasks() : 0, pl->eng ? pl->eng->batch() : 0,
                                                p.xstride};
        std::memcpy(v, a, sizeof a);
    });
}

int kbsrc_plan_stage_device(kbsrc_plan* pl, const double* d_interleaved, double* d_padded, void* stream) {
    return guarded([&] {
        if (!pl || !d_padded || (pl->p.nin > 0 && !d_interleaved)) throw std::invalid_argument("bad input buffer");
        if (!pl->eng) throw std::invalid_argument("plan has no engine (equal rates)");
        NearestRounding rn(pl->opt.precision != KBSRC_FAST);
        stage_device(pl, d_interleaved, d_padded, static_cast<hipStream_t>(stream));
    });
}

int kbsrc_plan_stats(const kbsrc_plan* pl, kbsrc_stats* st) {
    return guarded([&] {
        if (!pl || !st) throw std::invalid_argument("bad input buffer");
        fill_stats(*pl, 0, st);
    });
}

int kbsrc_pack(const double* d_planar, int channels, long long nout, long long frame0, long long frames, int format,
               void* d_bytes, unsigned long long* d_clipped, void* stream) {
    return guarded([&] {
        if (channels < 1 || nout < 0 || frame0 < 0 || frames < 0 || frame0 > nout || frames > nout - frame0 ||
            pack_bytes(format) == 0 || (frames > 0 && (!d_planar || !d_bytes || !d_clipped)) ||
            (reinterpret_cast<uintptr_t>(d_bytes) & 7))
            throw std::invalid_argument("bad input buffer");
        PackArgs a;
        a.y = d_planar; a.nout = nout; a.frame0 = frame0; a.frames = frames; a.ch = channels; a.format = format;
        a.out = static_cast<unsigned char*>(d_bytes); a.clipped = d_clipped;
        hip_check(launch_pack(a, static_cast<hipStream_t>(stream)), "pack kernel");
    });
}

int kbsrc_plan_stage(kbsrc_plan* pl, const double* in, double* d_padded, void* stream) {
    return guarded([&] {
        if (!pl || !d_padded || (pl->p.nin > 0 && !in)) throw std::invalid_argument("bad input buffer");
        if (!pl->eng) throw std::invalid_argument("plan has no engine (equal rates)");
        NearestRounding rn(pl->opt.precision != KBSRC_FAST);
        stage(pl, in, d_padded, static_cast<hipStream_t>(stream));
    });
}

int kbsrc_plan_execute(kbsrc_plan* pl, const double* d_padded, double* d_out, void* stream) {
    return guarded([&] {
        if (!pl || !pl->eng || !d_padded || !d_out) throw std::invalid_argument("bad input buffer");
        pl->eng->execute(d_padded, d_out, static_cast<hipStream_t>(stream));
    });
}

int kbsrc_plan_execute_timed(kbsrc_plan* pl, const double* d_padded, double* d_out, void* stream, double* ms) {
    return guarded([&] {
        if (!pl || !pl->eng || !d_padded || !d_out || !ms) throw std::invalid_argument("bad input buffer");
        PassTimes t;
        pl->eng->execute(d_padded, d_out, static_cast<hipStream_t>(stream), &t);
        ms[0] = t.fwd_col; ms[1] = t.fwd_row; ms[2] = t.mid; ms[3] = t.inv_col; ms[4] = t.inv_row;
    });
}

void kbsrc_plan_destroy(kbsrc_plan* pl) { delete pl; }

int kbsrc_kernel_taps(int L, long long a, long long b, double beta, double* out) {
    return guarded([&] {
        if (L < 1 || a < 1 || b < 1 || !(beta > 0) || !out) throw std::invalid_argument("option out of range");
        NearestRounding rn(true);
        const DevBuf h = kernel_taps_gpu(L, a, b, beta, nullptr);
        hip_check(hipMemcpy(out, h.d(), sizeof(double) * 2 * ((size_t)L + 1), hipMemcpyDeviceToHost), "taps download");
    });
}

int kbsrc_kernel_spectrum(const double* taps, int Lk, int n, double scale, double* out) {
    return guarded([&] {
        if (!taps || !out || Lk < 0 || n < 2 || (n & 1) || Lk >= n || !(scale != 0))
            throw std::invalid_argument("option out of range");
        NearestRounding rn(true);
        DevBuf h;
        h.upload(taps, sizeof(double) * 2 * ((size_t)Lk + 1), nullptr);
        DeviceConsts k;
        k.init(nullptr);
        const DevBuf H = kernel_spectrum_gpu(h.d(), Lk, n, scale, k, nullptr);
        hip_check(hipMemcpy(out, H.d(), sizeof(double) * 2 * ((size_t)n / 2 + 1), hipMemcpyDeviceToHost),
                  "spectrum download");
    });
}

}  // extern "C"

namespace kbsrc {

const char* version() { return kbsrc_version(); }

std::vector<std::vector<double>> resample(const double* in, long long frames, int channels, int rate_in,
                                          int rate_out, const Options& options, Stats* stats) {
    kbsrc_options o = defaults();
    if (options.precision == Precision::LongDouble) throw std::invalid_argument("option out of range");
    o.precision = options.precision == Precision::Fast ? KBSRC_FAST : KBSRC_EXACT;
    o.length = options.length;
    o.beta = options.beta;
    o.threads = options.threads;
    o.block = options.block;
    o.lpc_edges = options.lpc_edges ? 1 : 0;
    if (channels < 1 || frames < 0 || (frames > 0 && !in)) throw std::invalid_argument("bad input buffer");
    if (rate_in < 1 || rate_out < 1) throw std::invalid_argument("sample rates must be positive");
    if (!(o.length > 0 && o.length <= 10) || !(o.beta > 0 && o.beta <= 200) || o.threads < 0 || o.block < 0)
        throw std::invalid_argument("option out of range");
    // validated in the reference's order above; the output length needs only the reduced ratio
    const Plan p = make_plan(frames, channels, rate_in, rate_in, false, o.length, o.beta, o.block);
    const long long g = std::gcd((long long)rate_in, (long long)rate_out);
    const long long nout = (long long)(((__int128)frames * (rate_out / g) + rate_in / g - 1) / (rate_in / g));
    if (frames > LLONG_MAX / 16 || nout > LLONG_MAX / 16) throw std::invalid_argument("input too long");
    (void)p;
    std::vector<double> flat((size_t)channels * nout);
    kbsrc_stats st{};
    resample_host(in, frames, channels, rate_in, rate_out, o, flat.data(), &st);
    std::vector<std::vector<double>> out(channels);
    for (int c = 0; c < channels; ++c)
        out[c].assign(flat.begin() + (size_t)c * nout, flat.begin() + (size_t)(c + 1) * nout);
    if (stats) {
        stats->p = st.p; stats->q = st.q; stats->taps = st.taps; stats->block = st.block; stats->threads = st.threads;
        stats->general_fold = st.general_fold; stats->design_seconds = st.design_seconds;
        stats->process_seconds = st.process_seconds; stats->passband = st.passband; stats->stopband = st.stopband;
        stats->isa = st.isa;
    }
    return out;
}

}  // namespace kbsrc
