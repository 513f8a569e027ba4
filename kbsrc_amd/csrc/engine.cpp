// GPU engine orchestration: FFT plans, device tables, and the pass sequence of the overlap-save engine.
#include "engine.hpp"

#include <algorithm>
#include <numeric>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#pragma clang fp contract(off)

namespace kbsrc::gpu {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
thread_local const PoolScope* t_scope = nullptr;

// Device blocks of released pooled DevBufs, kept for reuse by later scoped allocations (on this ROCm a hipFree or
// hipFreeAsync of a large block costs ~1.5 ms on the calling thread and the stream-ordered pools return freed memory
// at once, whatever their release threshold). A block carries an event recorded on the stream of its release; the
// allocation that takes it makes its own stream wait for that event. Up to kCap bytes stay cached, the most recent
// releases first; beyond that the oldest blocks are freed.
class BlockCache {
public:
    // a cached block of at least `bytes` (at most twice that) on the current device, ordered on s, or nullptr
    void* take(size_t bytes, hipStream_t s, size_t& n) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "device");
        std::lock_guard<std::mutex> lock(m_);
        size_t best = blocks_.size();
        for (size_t i = 0; i < blocks_.size(); ++i) {
            const Block& b = blocks_[i];
            if (b.dev == dev && b.n >= bytes && b.n <= 2 * bytes && (best == blocks_.size() || b.n < blocks_[best].n))
                best = i;
        }
        if (best == blocks_.size()) return nullptr;
        Block b = blocks_[best];
        blocks_.erase(blocks_.begin() + (std::ptrdiff_t)best);
        cached_ -= b.n;
        if (hipStreamWaitEvent(s, b.ready, 0) != hipSuccess) {   // cannot order it: free it once the device is idle
            (void)hipGetLastError();
            evict(b);
            return nullptr;
        }
        spare_.push_back(b.ready);
        n = b.n;
        return b.p;
    }
    // keeps p (n bytes, on the current device) for reuse once the work queued on s so far is done
    void give(void* p, size_t n, hipStream_t s) {
        int dev = 0;
        Block b{p, n, nullptr, 0};
        std::vector<Block> out;
        {
            std::lock_guard<std::mutex> lock(m_);
            if (!spare_.empty()) {
                b.ready = spare_.back();
                spare_.pop_back();
            }
        }
        bool ok = hipGetDevice(&dev) == hipSuccess &&
                  (b.ready || hipEventCreateWithFlags(&b.ready, hipEventDisableTiming) == hipSuccess) &&
                  hipEventRecord(b.ready, s) == hipSuccess;
        b.dev = dev;
        if (!ok) {   // not reusable in order: freed after everything queued on the device
            (void)hipGetLastError();
            (void)hipDeviceSynchronize();
            (void)hipFree(p);
            if (b.ready) (void)hipEventDestroy(b.ready);
            return;
        }
        {
            std::lock_guard<std::mutex> lock(m_);
            blocks_.push_back(b);
            cached_ += n;
            while (cached_ > kCap && !blocks_.empty()) {
                out.push_back(blocks_.front());
                cached_ -= blocks_.front().n;
                blocks_.erase(blocks_.begin());
            }
        }
        for (const Block& o : out) evict(o);
    }
    // frees every cached block of the current device (each once its last work is done); true if any was freed
    bool trim() {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return false;
        std::vector<Block> out;
        {
            std::lock_guard<std::mutex> lock(m_);
            for (size_t i = 0; i < blocks_.size();) {
                if (blocks_[i].dev == dev) {
                    out.push_back(blocks_[i]);
                    cached_ -= blocks_[i].n;
                    blocks_.erase(blocks_.begin() + (std::ptrdiff_t)i);
                } else {
                    ++i;
                }
            }
        }
        for (const Block& o : out) evict(o);
        return !out.empty();
    }
    static BlockCache& get() {
        static BlockCache* c = new BlockCache;   // kept until the process ends
        return *c;
    }

private:
    static constexpr size_t kCap = (size_t)4 << 30;
    struct Block {
        void* p;
        size_t n;
        hipEvent_t ready;
        int dev;
    };
    static void evict(const Block& b) {
        (void)hipEventSynchronize(b.ready);
        (void)hipFree(b.p);
        (void)hipEventDestroy(b.ready);
    }
    std::mutex m_;
    std::vector<Block> blocks_;
    std::vector<hipEvent_t> spare_;
    size_t cached_ = 0;
};
}  // namespace

PoolScope::PoolScope(hipStream_t s) : s_(s), prev_(t_scope) { t_scope = this; }
PoolScope::~PoolScope() { t_scope = prev_; }
const PoolScope* PoolScope::current() { return t_scope; }

void DevBuf::alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 16;
    const PoolScope* sc = PoolScope::current();
    size_t n = bytes;
    if (sc && (p_ = BlockCache::get().take(bytes, sc->stream(), n))) {
        n_ = n;
        pooled_ = true;
        return;
    }
    hipError_t e = hipMalloc(&p_, bytes);
    if (e != hipSuccess && BlockCache::get().trim()) {   // out of memory: give the cached blocks back, once
        (void)hipGetLastError();
        e = hipMalloc(&p_, bytes);
    }
    if (e != hipSuccess) {
        p_ = nullptr;
        (void)hipGetLastError();
        throw std::bad_alloc();
    }
    n_ = bytes;
    pooled_ = sc != nullptr;
}
void DevBuf::release() {
    if (p_) {
        if (!pooled_) {
            (void)hipFree(p_);
        } else if (const PoolScope* sc = PoolScope::current()) {
            BlockCache::get().give(p_, n_, sc->stream());
        } else {   // no order known: reusable after everything queued on the device
            (void)hipDeviceSynchronize();
            BlockCache::get().give(p_, n_, nullptr);
        }
    }
    p_ = nullptr;
    n_ = 0;
    pooled_ = false;
}
void DevBuf::upload(const void* host, size_t bytes, hipStream_t s) {
    if (bytes > n_) alloc(bytes);
    hip_check(hipMemcpyAsync(p_, host, bytes, hipMemcpyHostToDevice, s), "upload");
    hip_check(hipStreamSynchronize(s), "upload sync");
}

std::vector<double> pack_cx(const std::vector<dd>& re, const std::vector<dd>& im, bool as_dd) {
    const size_t n = re.size();
    std::vector<double> v(n * (as_dd ? 4 : 2));
    for (size_t i = 0; i < n; ++i) {
        if (as_dd) {
            v[4 * i] = re[i].hi; v[4 * i + 1] = im[i].hi; v[4 * i + 2] = re[i].lo; v[4 * i + 3] = im[i].lo;
        } else {
            v[2 * i] = re[i].hi; v[2 * i + 1] = im[i].hi;
        }
    }
    return v;
}

void SubFFT::init(int n, bool as_dd, hipStream_t s) {
    std::vector<int> rs;
    int t = n;
    while (t % 4 == 0) { rs.push_back(4); t /= 4; }
    while (t % 2 == 0) { rs.push_back(2); t /= 2; }
    for (int f : {3, 5, 7})
        while (t % f == 0) { rs.push_back(f); t /= f; }
    if (t != 1) throw std::runtime_error("FFT length not 7-smooth");
    if ((int)rs.size() > kMaxStages) throw std::runtime_error("FFT length has too many stages");
    desc.n = n;
    desc.nst = (int)rs.size();
    CisTable ct(n);
    std::vector<dd> re, im;
    int rem = n, stride = 1;
    for (size_t i = 0; i < rs.size(); ++i) {
        const int r = rs[i], m = rem / r;
        desc.st[i] = StageDesc{r, m, stride, (int)re.size()};
        for (int p = 0; p < m; ++p)
            for (int u = 1; u < r; ++u) {
                dd c, sn;
                ct.get(-(long long)stride * p * u, c, sn);
                re.push_back(c);
                im.push_back(sn);
            }
        rem /= r;
        stride *= r;
    }
    const std::vector<double> v = pack_cx(re, im, as_dd);
    tw.upload(v.data(), v.size() * sizeof(double), s);
}

void SubFFT::init_split(hipStream_t s) {
    if (tw16.d() || desc.nst < 2 || desc.st[0].r != 4 || desc.st[1].r != 4) return;
    const int n = desc.n;
    CisTable ct(n);
    std::vector<dd> re, im;
    for (int p = 0; p < n / 16; ++p)
        for (int k = 1; k < 16; ++k) {
            dd c, sn;
            ct.get(-(long long)p * k, c, sn);
            re.push_back(c);
            im.push_back(sn);
        }
    const std::vector<double> v = pack_cx(re, im, true);
    tw16.upload(v.data(), v.size() * sizeof(double), s);
    hip_check(hipStreamSynchronize(s), "split twiddles");   // the host table is freed on return
}

void CFFTPlan::init_split(hipStream_t s) {
    if (blue) {
        inner->init_split(s);
    } else if (four) {
        s1.init_split(s);
        s2.init_split(s);
    } else {
        sd.init_split(s);
    }
}

std::vector<uint64_t> split_matrix_digits(int S) {
    std::vector<int8_t> dig((size_t)S * 32 * 32);
    for (int r = 0; r < 32; ++r)
        for (int cc = 0; cc < 32; ++cc) {
            // row r: Re y[r] (r < 16) or Im y[r - 16]; column cc: Re x[cc / 2] (cc even) or Im x[cc / 2]
            const int kk = r & 15, nn = cc >> 1;
            const bool row_re = r < 16, col_re = (cc & 1) == 0;
            dd c, sn;
            cis((long long)kk * nn % 16, 16, c, sn);   // e^{+i theta}: Fr = c, Fi = -sn
            dd v = row_re == col_re ? c : (row_re ? sn : ddneg(sn));
            v = dd{v.hi / 2, v.lo / 2};
            const double hs = std::ldexp(v.hi, 7 * S), ls = std::ldexp(v.lo, 7 * S);
            const double r1 = std::nearbyint(hs);
            __int128 Z = (__int128)r1 + (__int128)std::nearbyint((hs - r1) + ls);
            for (int j = S - 1; j >= 1; --j) {
                const int d = (int)((Z + 64) & 127) - 64;
                Z = (Z - d) >> 7;
                dig[((size_t)j * 32 + r) * 32 + cc] = (int8_t)d;
            }
            dig[(size_t)r * 32 + cc] = (int8_t)(int)Z;
        }
    std::vector<uint64_t> ad((size_t)S * 2 * 64);
    for (int j = 0; j < S; ++j)
        for (int mt = 0; mt < 2; ++mt)
            for (int l = 0; l < 64; ++l) {
                uint64_t w = 0;
                for (int bb = 0; bb < 8; ++bb)
                    w |= (uint64_t)(uint8_t)dig[((size_t)j * 32 + 16 * mt + l % 16) * 32 + 8 * (l / 16) + bb]
                         << (8 * bb);
                ad[((size_t)j * 2 + mt) * 64 + l] = w;
            }
    return ad;
}

namespace {

constexpr int kThreads = 256;
constexpr size_t kLdsBudget = 48 * 1024;

constexpr size_t kLdsMax = 160 * 1024;

int pick_group(int policy, int n, int nsub) {
    const size_t one = pass_lds_bytes(policy, n, 1);
    int G = (int)std::max<size_t>(1, kLdsBudget / one);
    return std::max(1, std::min({G, nsub, 16}));
}

// Test and tuning switches. Read at every execute: KBSRC_GENERIC_PASSES=1 runs every sub-FFT pass through the
// generic kernel; KBSRC_FUSED=1 folds the split/multiply/fold kernel into the passes next to it (same bits; it saves
// a kernel and ~1.2 GB of HBM traffic per benchmark step but measured no faster, the passes being FP64-bound).
// Read when a plan is made: KBSRC_GLOBAL_PASSES=1 runs the passes through global scratch, as sub-FFTs too long for
// an LDS tile always do (two LDS passes where the stage list splits, =2: one launch per stage). Read at every split
// execute: KBSRC_SPLIT_GLOBAL=1 runs every split-form pass through the global-memory route (same bits).
// KBSRC_LONG_PARTS=0 keeps long column sub-FFTs that have a two-part route in their one-pass kernel (same bits).
bool env_flag(const char* name) {
    const char* v = std::getenv(name);
    return v && v[0] == '1';
}
bool env_off(const char* name) {
    const char* v = std::getenv(name);
    return v && v[0] == '0';
}
bool use_fixed() { return !env_flag("KBSRC_GENERIC_PASSES"); }
bool global_forced() {   // KBSRC_GLOBAL_PASSES=1: split into two LDS passes where possible; =2: one launch per stage
    const char* v = std::getenv("KBSRC_GLOBAL_PASSES");
    return v && (v[0] == '1' || v[0] == '2');
}

// ---- measured kernel shapes. Whatever shape runs a pass, every butterfly computes the same operations, so the
// choice changes time, never bits. Tuning a pass times every registered specialised variant and a range of generic
// (G, T) shapes on the pass's real arguments and keeps the fastest for (policy, length, load, store, sub-FFTs).
struct PassKey {
    int policy, n, ld, st, nsub;
    bool operator<(const PassKey& o) const {
        return std::tie(policy, n, ld, st, nsub) < std::tie(o.policy, o.n, o.ld, o.st, o.nsub);
    }
};
struct PassChoice {
    int variant;   // >= 0: specialised variant; -1: generic kernel with G, T
    int G, T;
};
std::mutex g_tune_mu;
std::map<PassKey, PassChoice> g_tuned;
thread_local bool t_tuning = false;

hipError_t launch_choice(int policy, int ld, int st, PassArgs& a, hipStream_t s, const PassChoice& c) {
    hipError_t e = hipSuccess;
    if (c.variant >= 0) {
        if (!launch_fixed(policy, ld, st, a, s, &e, c.variant)) return hipErrorInvalidValue;
        return e;
    }
    a.G = c.G;
    a.ngrp = (a.nsub + c.G - 1) / c.G;
    return launch_pass(policy, ld, st, a, c.T, s);
}

// Times each candidate shape of this pass (one warm-up launch, then the best of three), runs the fastest last so
// that its output stands, and records it.
hipError_t tune_pass(int policy, int ld, int st, PassArgs& a, hipStream_t s, const PassKey& key) {
    std::vector<PassChoice> cand;
    const int nv = use_fixed() ? fixed_variants(policy, a.sd.n, ld, st) : 0;
    for (int v = 0; v < nv; ++v) cand.push_back({v, 0, 0});
    const size_t one = pass_lds_bytes(policy, a.sd.n, 1);
    const int gmax = (int)std::max<size_t>(1, std::min<size_t>({(size_t)96 * 1024 / one, (size_t)a.nsub, 16}));
    for (int G = 1; G <= gmax; ++G)
        for (int T = 64; T <= kThreads; T += 64)
            if ((long long)G * a.sd.n >= T) cand.push_back({-1, G, T});
    if (cand.empty()) cand.push_back({-1, 1, kThreads});
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hip_check(hipEventCreate(&e0), "tune event");
    hip_check(hipEventCreate(&e1), "tune event");
    float best_ms = 1e30f;
    PassChoice best = cand[0];
    hipError_t err = hipSuccess;
    for (const PassChoice& c : cand) {
        PassArgs t = a;
        if ((err = launch_choice(policy, ld, st, t, s, c)) != hipSuccess) break;
        float ms = 1e30f;
        for (int rep = 0; rep < 3 && err == hipSuccess; ++rep) {
            hip_check(hipEventRecord(e0, s), "tune event");
            err = launch_choice(policy, ld, st, t, s, c);
            hip_check(hipEventRecord(e1, s), "tune event");
            hip_check(hipEventSynchronize(e1), "tune event");
            float m = 0;
            hip_check(hipEventElapsedTime(&m, e0, e1), "tune event");
            ms = std::min(ms, m);
        }
        if (err != hipSuccess) break;
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess) return err;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tuned[key] = best;
    }
    return launch_choice(policy, ld, st, a, s, best);
}

}  // namespace

TuningScope::TuningScope() { t_tuning = true; }
TuningScope::~TuningScope() { t_tuning = false; }
int tuned_shapes() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    return (int)g_tuned.size();
}

namespace {

// launch one sub-FFT pass; errors name the pass and its shape
void run_pass(int policy, int ld, int st, PassArgs a, hipStream_t s, const char* what) {
    hipError_t e = hipSuccess;
    const bool forced = a.gscr0 && ld != LD_FOLD && st != ST_SPLIT && global_forced();
    const bool tunable = !forced && ld != LD_FOLD && st != ST_SPLIT && pass_lds_bytes(policy, a.sd.n, 1) <= kLdsMax;
    const PassKey key{policy, a.sd.n, ld, st, a.nsub};
    bool known = false;
    PassChoice choice{};
    if (a.split > 0 && a.tw16 && policy != POL_F64 && ld != LD_FOLD && st != ST_SPLIT) {   // the split form
        // a specialised split pass where one is built, else inside one generic LDS pass where the sub-FFT fits
        // (KBSRC_SPLIT_GLOBAL=1: always the global-memory route; =2: never the specialised passes)
        const char* sg = std::getenv("KBSRC_SPLIT_GLOBAL");
        const bool global = sg && sg[0] == '1', no_fixed = sg && sg[0] == '2';
        const SplitFixedEntry* fe = policy == POL_DDS && a.pm == 0 && !global && !no_fixed && pass_offsets_fit32(a)
                                        ? find_split_fixed(a.sd.n, ld, st, a.split) : nullptr;
        if (fe) {
            a.G = fe->G;   // the kernel's compile-time group; a partial last group is masked
            a.ngrp = (a.nsub + a.G - 1) / a.G;
            set_prune_window(a, st);
            e = fe->launch(a, s);
        } else if (split_lds_ok(policy, ld, st, a) && !global) {
            // up to 80 KB of LDS per workgroup: the split kernel's registers allow two workgroups per CU, and more
            // vectors per workgroup keep its four waves busy in the matrix stage (measured on 320-point rows:
            // G = 4 beat G = 2 by ~10% at 15 slices)
            a.G = (int)std::max<size_t>(1, std::min<size_t>({(size_t)a.nsub, 16, 80 * 1024 / pass_lds_bytes(policy, a.sd.n, 1)}));
            int threads = kThreads;
            if (const char* v = std::getenv("KBSRC_SPLIT_SHAPE")) {   // tuning knob: "G,T"
                int g = 0, t = 0;
                if (std::sscanf(v, "%d,%d", &g, &t) == 2 && g >= 1 && t >= 64 && t <= 1024 && t % 64 == 0 &&
                    pass_lds_bytes(policy, a.sd.n, g) <= kLdsMax) {
                    a.G = std::min(g, a.nsub);
                    threads = t;
                }
            }
            a.ngrp = (a.nsub + a.G - 1) / a.G;
            e = launch_split_lds(policy, ld, st, a, threads, s);
        } else {
            e = launch_split_pass(policy, ld, st, a, s);
        }
        if (e != hipSuccess)
            throw HipError(std::string(what) + " (split form, n=" + std::to_string(a.sd.n) + "): " +
                           hipGetErrorString(e));
        return;
    }
    if (tunable && !t_tuning) {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        const auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            known = true;
            choice = it->second;
        }
    }
    if (tunable && t_tuning) {
        std::unique_lock<std::mutex> lock(g_tune_mu);
        const auto it = g_tuned.find(key);
        known = it != g_tuned.end();
        if (known) choice = it->second;
    }
    if (tunable && t_tuning && !known) {
        e = tune_pass(policy, ld, st, a, s, key);
    } else if (known) {
        e = launch_choice(policy, ld, st, a, s, choice);
    } else if (!forced && use_fixed() && launch_fixed(policy, ld, st, a, s, &e)) {
    } else if (forced || pass_lds_bytes(policy, a.sd.n, 1) > kLdsMax) {
        if (!a.gscr0) throw std::runtime_error(std::string(what) + ": no global scratch for a long sub-FFT");
        e = launch_global(policy, ld, st, a, s);
    } else {
        e = launch_pass(policy, ld, st, a, kThreads, s);
    }
    if (e != hipSuccess)
        throw HipError(std::string(what) + " (n=" + std::to_string(a.sd.n) + ", G=" + std::to_string(a.G) +
                       ", LDS " + std::to_string(pass_lds_bytes(policy, a.sd.n, a.G)) + " B): " + hipGetErrorString(e));
}

void fill_sub(PassArgs& a, const SubFFT& sf, const DeviceConsts& k, bool as_dd, int policy, int nsub) {
    a.sd = sf.desc;
    a.tw = sf.tw.d();
    a.tw16 = sf.tw16.d();
    a.oddc = as_dd ? k.odd_dd.d() : k.odd_f64.d();
    a.nsub = nsub;
    a.G = pick_group(policy, sf.desc.n, nsub);
    a.ngrp = (nsub + a.G - 1) / a.G;
}

}  // namespace

namespace {

// Long twiddle table of CisTable(N) on the GPU: its two short tables computed there (the host's dd cis, the same
// bits: CisTable's constructor), then every entry composed from them. Queued on s.
void device_cis_table(DevBuf& out, long long N, long long count, int mode, long long N2, bool as_dd, hipStream_t s) {
    const long long B = (long long)std::ceil(std::sqrt((double)N)), J = N / B + 1;   // as CisTable::CisTable
    DevBuf ab((size_t)(J + B) * 32);
    hip_check(launch_cis_short(N, B, J, ab.d(), ab.d() + 4 * J, s), "twiddle short tables");
    out.alloc((size_t)count * (as_dd ? 32 : 16));
    TableArgs t;
    t.a = ab.d(); t.b = ab.d() + 4 * J; t.N = N; t.B = B; t.count = count; t.N2 = N2; t.mode = mode;
    t.as_dd = as_dd; t.out = out.d();
    hip_check(launch_cis_table(t, s), "twiddle table");
    // ab's release: in a PoolScope on s it is ordered after the kernels; otherwise hipFree waits for the device
}

}  // namespace

void DeviceConsts::init(hipStream_t s) {
    std::vector<dd> re(21, mkdd(0.0)), im(21, mkdd(0.0));
    const int rr[3] = {3, 5, 7};
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < rr[i]; ++k) cis(k, rr[i], re[i * 7 + k], im[i * 7 + k]);
    const std::vector<double> a = pack_cx(re, im, true), b = pack_cx(re, im, false);
    odd_dd.upload(a.data(), a.size() * sizeof(double), s);
    odd_f64.upload(b.data(), b.size() * sizeof(double), s);
}

void CFFTPlan::init(int n_, bool as_dd_, const DeviceConsts& k, hipStream_t s) {
    n = n_;
    as_dd = as_dd_;
    blue = !smooth7(n);
    if (blue) {   // Bluestein (src/dsp/fft.hpp:228-240)
        m = (int)next_smooth7(2LL * n - 1);
        inner = std::make_unique<CFFTPlan>();
        inner->init(m, as_dd, k, s);
        // the chirp and the chirp sequence v (zero between n and m - n) on the GPU
        CFFTPlan vp;
        vp.init(m, true, k, s);
        const long long ts = 4LL * vp.work_elems();
        DevBuf src(ts * 8), dst(ts * 8), s0(ts * 8), s1(ts * 8), s2(ts * 8);
        bc.alloc((size_t)n * (as_dd ? 32 : 16));
        hip_check(hipMemsetAsync(src.d(), 0, (size_t)m * 32, s), "chirp clear");
        hip_check(launch_chirp(n, m, as_dd, bc.d(), src.d(), s), "chirp");
        // the chirp sequence's spectrum in the reference's scalar-dd form, divided by m (src/dsp/fft.hpp:236-238)
        RunCtx c;
        c.ts = ts; c.k = &k; c.dd = true; c.lane = false; c.s = s;
        c.proto.ntask = 1;
        DevBuf g0, g1;
        if (vp.needs_global() || global_forced()) {
            g0.alloc(ts * 8);
            g1.alloc(ts * 8);
            c.proto.gscr0 = g0.d();
            c.proto.gscr1 = g1.d();
        }
        double* const scr[3] = {s0.d(), s1.d(), s2.d()};
        double tc = 0, tr = 0;
        run_cfft(vp, c, LD_CPLX, src.d(), ST_CPLX, dst.d(), scr, &tc, &tr);
        vf.alloc((size_t)m * (as_dd ? 32 : 16));
        hip_check(launch_div_m(dst.d(), m, as_dd, vf.d(), s), "chirp spectrum");
        hip_check(hipStreamSynchronize(s), "chirp sync");   // the work buffers are freed on return
        return;
    }
    four = n > 2048;
    if (four) {
        int best = 1;
        for (int d = 1; (long long)d * d <= n; ++d)
            if (n % d == 0) best = d;
        N1 = best;
        N2 = n / best;
        if (N1 < 8) four = false;
    }
    if (!four) {
        sd.init(n, as_dd, s);
        return;
    }
    s1.init(N1, as_dd, s);
    s2.init(N2, as_dd, s);
    device_cis_table(tw4, n, n, TAB_FOUR, N2, as_dd, s);   // w_n^{-k1 n2} at k1*N2 + n2
}

void RFFTPlan::init(int N_, bool as_dd, const DeviceConsts& k, hipStream_t s) {
    N = N_;
    h = N / 2;
    if (N % 2) throw std::runtime_error("RFFT length must be even");
    c.init(h, as_dd, k, s);
    device_cis_table(w, N, (long long)h + 1, TAB_LINEAR, 1, as_dd, s);   // w[q] = e^{-2 pi i q / N}
}

bool CFFTPlan::needs_global() const {
    const int pol = as_dd ? POL_DDS : POL_F64;
    if (blue) return inner->needs_global();
    if (four) return pass_lds_bytes(pol, N1, 1) > kLdsMax || pass_lds_bytes(pol, N2, 1) > kLdsMax;
    return pass_lds_bytes(pol, n, 1) > kLdsMax;
}

namespace {

// brackets one launch with HIP events on the launch stream when timing is asked for; the host does not wait here
// (PassTimes::collect reads the events after the whole execute is queued)
template <class F> void timed(const RunCtx& c, double* acc, F&& fn) {
    if (!c.times) {
        fn();
        return;
    }
    // timing-only events: no system-scope fence when recorded (a fence writes back and invalidates the caches
    // after the kernel, which the interval would then include)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hip_check(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence), "event");
    if (hipEventCreateWithFlags(&e1, hipEventDisableSystemFence) != hipSuccess) {
        (void)hipEventDestroy(e0);
        throw HipError("event");
    }
    c.times->pending.push_back({e0, e1, acc});
    hip_check(hipEventRecord(e0, c.s), "event record");
    fn();
    hip_check(hipEventRecord(e1, c.s), "event record");
    c.times->launches++;
}

}  // namespace

void PassTimes::collect() {
    hipError_t err = hipSuccess;
    for (const Pending& p : pending) {
        float ms = 0;
        if (err == hipSuccess) err = hipEventSynchronize(p.e1);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, p.e0, p.e1);
        if (err == hipSuccess) *p.acc += ms;
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
    pending.clear();
    hip_check(err, "pass timing");
}

PassTimes::~PassTimes() {
    for (const Pending& p : pending) {
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
}

namespace {

// launch one pass, or describe it for the generic kernel when the context collects
void do_pass(const RunCtx& c, double* t, int policy, int ld, int st, PassArgs a, const char* what) {
    if (c.collect) {
        const bool io = (ld == LD_REAL_F64 || ld == LD_CPLX) && (st == ST_CPLX || st == ST_OUT);
        if (!io || pass_lds_bytes(policy, a.sd.n, 1) > kLdsMax || a.split) {
            *c.collect_ok = false;
            return;
        }
        a.G = pick_group(policy, a.sd.n, a.nsub);
        a.ngrp = (a.nsub + a.G - 1) / a.G;
        c.collect->push_back(PassStep{policy, ld, st, a});
        return;
    }
    timed(c, t, [&] { run_pass(policy, ld, st, a, c.s, what); });
}

}  // namespace

void run_cfft(const CFFTPlan& f, const RunCtx& c, int ld, const double* src, int st, double* dst,
              double* const scr[3], double* t_col, double* t_row) {
    const bool lane = c.lane;
    const int pol_four = !c.dd ? POL_F64 : lane ? POL_DDS : POL_DDA;
    const int pol_one = c.dd ? POL_DDA : POL_F64;
    const long long ts = c.ts;
    if (f.blue && c.proto.zkeep >= 0) {   // every output of the inner transforms is needed
        RunCtx all = c;
        all.proto.zkeep = -1;
        run_cfft(f, all, ld, src, st, dst, scr, t_col, t_row);
        return;
    }
    if (f.blue && c.collect) {
        *c.collect_ok = false;
        return;
    }
    if (f.blue) {   // X_k = conj(b_k) sum_j x_j conj(b_j) b_{k-j}, through the inner FFT twice (src/dsp/fft.hpp:268-278)
        BlueArgs b;
        b.n = f.n; b.m = f.m; b.ntask = c.proto.ntask; b.task0 = c.proto.task0; b.nblk = c.proto.nblk;
        b.in_cstride = c.proto.in_cstride; b.in_bstride = c.proto.in_bstride; b.tdesc = c.proto.tdesc;
        b.bc = f.bc.d(); b.vf = f.vf.d();
        b.out = c.proto.out; b.nout = c.proto.nout; b.o = c.proto.o; b.margP = c.proto.margP; b.hopP = c.proto.hopP;
        b.out_dd = c.proto.out_dd;
        const int pol = pol_one;
        b.src = src; b.src_tstride = ld == LD_REAL_DD ? 0 : ts; b.dst = scr[0]; b.dst_tstride = ts;
        timed(c, t_col, [&] { hip_check(launch_blue(pol, BLUE_PRE, ld, b, c.s), "Bluestein pre-multiply"); });
        double* const inner_scr[3] = {scr[1], nullptr, nullptr};
        run_cfft(*f.inner, c, LD_CPLX, scr[0], ST_CPLX, scr[2], inner_scr, t_col, t_row);
        b.dst = scr[2];
        timed(c, t_row, [&] { hip_check(launch_blue(pol, BLUE_MID, 0, b, c.s), "Bluestein chirp product"); });
        run_cfft(*f.inner, c, LD_CPLX, scr[2], ST_CPLX, scr[0], inner_scr, t_col, t_row);
        b.src = scr[0]; b.src_tstride = ts; b.dst = dst; b.dst_tstride = ts;
        timed(c, t_row, [&] { hip_check(launch_blue(pol, BLUE_POST, st, b, c.s), "Bluestein post-multiply"); });
        return;
    }
    const long long src_ts = ld == LD_REAL_DD ? 0 : ts;
    if (f.four) {
        PassArgs a = c.proto;
        fill_sub(a, f.s1, *c.k, c.dd, pol_four, f.N2);
        a.src = src; a.src_tstride = src_ts; a.lde = f.N2; a.ldg = 1;
        a.dst = scr[0]; a.dst_tstride = ts; a.ste = f.N2; a.stg = 1;
        a.tw4 = f.tw4.d(); a.twmode = lane ? 1 : 2; a.renorm = lane && c.dd;
        a.f.M = f.N2;
        a.zkeep = -1;   // the column pass's outputs all feed the rows
        if (const char* v = std::getenv("KBSRC_COL_TGROUP")) a.tgroup = std::max(0, std::atoi(v));   // tuning knob
        // a long column sub-FFT as two specialised passes through the row pass's destination, free until the rows
        // (KBSRC_LONG_PARTS=0: always the one-pass kernel)
        const bool split_stage = a.split > 0 && a.sd.nst >= 2 && a.sd.st[0].r == 4 && a.sd.st[1].r == 4;
        const bool two = !c.collect && !split_stage && dst && long_two_part(pol_four, a.sd.n, ld, ST_CPLX) &&
                         pass_offsets_fit32(a) && (long long)a.sd.n * a.nsub * (c.dd ? 4 : 2) <= ts &&
                         !env_off("KBSRC_LONG_PARTS");
        if (two)
            timed(c, t_col, [&] {
                hip_check(launch_long_two_part(pol_four, ld, ST_CPLX, a, dst, ts, c.s), "column pass (two parts)");
            });
        else
            do_pass(c, t_col, pol_four, ld, ST_CPLX, a, "column pass");
        PassArgs b = c.proto;
        fill_sub(b, f.s2, *c.k, c.dd, pol_four, f.N1);
        b.src = scr[0]; b.src_tstride = ts; b.lde = 1; b.ldg = f.N2;
        b.dst = dst; b.dst_tstride = ts; b.ste = f.N1; b.stg = 1; b.renorm = lane && c.dd;
        b.f.M = f.N1;
        if (const char* v = std::getenv("KBSRC_ROW_TGROUP")) b.tgroup = std::max(0, std::atoi(v));   // tuning knob
        if (c.rows && !c.collect && st == ST_CPLX) (*c.rows)(b);
        else do_pass(c, t_row, pol_four, LD_CPLX, st, b, "row pass");
    } else {
        PassArgs a = c.proto;
        fill_sub(a, f.sd, *c.k, c.dd, pol_one, 1);
        a.src = src; a.src_tstride = src_ts; a.lde = 1; a.ldg = 0;
        a.dst = dst; a.dst_tstride = ts; a.ste = 1; a.stg = 0;
        do_pass(c, t_row, pol_one, ld, st, a, "single pass");
    }
}

namespace {

// tables made once per device and kept for the life of the process (never freed: no teardown order to get wrong)
struct Dft16Tables {
    DeviceConsts k;
    SubFFT sf;
    std::map<int, DevBuf> adig;   // split matrix slices per slice count
};
Dft16Tables& dft16_tables(hipStream_t s) {
    static std::mutex mu;
    static auto* per_device = new std::map<int, Dft16Tables*>();
    int dev = 0;
    hip_check(hipGetDevice(&dev), "device");
    std::lock_guard<std::mutex> lock(mu);
    Dft16Tables*& t = (*per_device)[dev];
    if (!t) {
        auto fresh = std::make_unique<Dft16Tables>();
        fresh->k.init(s);
        fresh->sf.init(16, true, s);
        t = fresh.release();
    }
    return *t;
}

}  // namespace

const long* split_dft16_digits(int S, hipStream_t s) {
    Dft16Tables& t = dft16_tables(s);
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    DevBuf& b = t.adig[S];
    if (!b.d()) {
        const std::vector<uint64_t> ad = split_matrix_digits(S);
        b.upload(ad.data(), ad.size() * sizeof(uint64_t), s);
        hip_check(hipStreamSynchronize(s), "split matrix");
    }
    return reinterpret_cast<const long*>(b.d());
}

void plain_dft16(const double* x, long long nvec, double* y, hipStream_t s) {
    if (nvec < 1) return;
    if (nvec > 0x7fffffffLL / 16) throw std::invalid_argument("bad input buffer");
    const Dft16Tables& t = dft16_tables(s);
    const DeviceConsts& k = t.k;
    const SubFFT& sf = t.sf;
    PassArgs a;
    fill_sub(a, sf, k, true, POL_DDS, (int)nvec);
    a.ntask = 1;
    a.src = x; a.src_tstride = 0; a.lde = 1; a.ldg = 16;
    a.dst = y; a.dst_tstride = 0; a.ste = 1; a.stg = 16;
    a.renorm = 1;
    run_pass(POL_DDS, LD_CPLX, ST_CPLX, a, s, "radix-16 pass");
    hip_check(hipStreamSynchronize(s), "radix-16 sync");
}

void plain_dft16_inwave(const double* x, long long nvec, double* y, hipStream_t s) {
    if (nvec < 1) return;
    if (nvec > 0x7fffffffLL / 16) throw std::invalid_argument("bad input buffer");
    const Dft16Tables& t = dft16_tables(s);
    hip_check(launch_dft16_inwave(x, nvec, y, t.sf.tw.d(), t.k.odd_dd.d(), s), "in-wave radix-16 kernel");
    hip_check(hipStreamSynchronize(s), "radix-16 sync");
}

DevBuf kernel_taps_gpu(long long L, long long a, long long b, double beta, hipStream_t s) {
    TapArgs t;
    t.L = L; t.a = a; t.b = b; t.beta = beta;
    const dd ratio = dddiv1(mkdd((double)a), (double)b), i0 = bessel_i0(mkdd(beta));
    t.ratio_hi = ratio.hi; t.ratio_lo = ratio.lo; t.i0_hi = i0.hi; t.i0_lo = i0.lo;
    DevBuf h((size_t)(L + 1) * 16);
    t.h = h.d();
    hip_check(launch_taps(t, s), "kernel taps");
    return h;
}

DevBuf kernel_spectrum_gpu(const double* taps, long long Lk, int n, double scale, const DeviceConsts& k,
                           hipStream_t s) {
    RFFTPlan R;
    R.init(n, true, k, s);
    const long long ts = 4LL * std::max<long long>(R.c.work_elems(), R.h + 1);
    DevBuf din((size_t)n * 16), z(ts * 8), s0(ts * 8), s1(ts * 8), s2(ts * 8), dH((size_t)(R.h + 1) * 16), g0, g1;
    hip_check(launch_circ(taps, Lk, n, din.d(), s), "circular kernel");
    RunCtx c;
    c.ts = ts; c.k = &k; c.dd = true; c.lane = false; c.s = s;
    c.proto.ntask = 1;
    if (R.c.needs_global() || global_forced()) {
        g0.alloc(ts * 8);
        g1.alloc(ts * 8);
        c.proto.gscr0 = g0.d();
        c.proto.gscr1 = g1.d();
    }
    double* const scr[3] = {s0.d(), s1.d(), s2.d()};
    double tc = 0, tr = 0;
    run_cfft(R.c, c, LD_REAL_DD, din.d(), ST_CPLX, z.d(), scr, &tc, &tr);
    SpecArgs sa;
    sa.z = z.d(); sa.w = R.w.d(); sa.H = dH.d(); sa.h = R.h; sa.scale = scale;
    hip_check(launch_spec_split(sa, s), "spectrum split");
    hip_check(hipStreamSynchronize(s), "spectrum sync");   // the work buffers are freed on return
    return dH;
}

Engine::Engine(const Plan& p, const double* H, long long nH, int Jg, hipStream_t s, long long ntask,
               const long long* tdesc, size_t buffer_budget)
    : p_(p), dd_(!p.fast), Jg_(Jg), tdesc_(tdesc) {
    ntask_ = ntask > 0 ? ntask : (long long)p.ch * p.nblk;
    k_.init(s);
    fw_.init(p.Nb, dd_, k_, s);
    iv_.init(p.Mb, dd_, k_, s);
    // the largest spectrum bin the bin map reads (src/resampler.cpp:53-78): down, bins up to min(Nb/2, Mb - 1);
    // up, bins up to min(Nb/2, Mb/2); general fold, up to min(Nb/2, Jg). The real-FFT split pairs bin i with
    // z[i] and z[h - i], so z[k] for zkeep < k < h - zkeep is never read; -1 when that range is empty.
    {
        const long long h = fw_.h;
        const long long xmax = Jg_ ? std::min<long long>(h, Jg_)
                             : p.up ? std::min<long long>(h, p.Mb / 2) : std::min<long long>(h, p.Mb - 1);
        zkeep_ = xmax + 1 < h - xmax ? xmax : -1;
        if (env_flag("KBSRC_STORE_ALL")) zkeep_ = -1;   // test switch: store every bin (same output)
        // down, plain bin map: X[i] reads z[i] and z[h - i], so the split/fold threads k and k +- |h - Mb| share
        // their z reads; their bin blocks (256 bins each) run next to each other
        if (!p.up && !Jg_) kchain_ = (int)((std::llabs(h - (long long)p.Mb) + 128) / 256);
        if (const char* v = std::getenv("KBSRC_MID_ORDER"); v && std::atoi(v) == 0) kchain_ = 0;   // test switch
        // task groups of 4 (mid_order): a chained block's repeated z reads come back 4 workgroups later (groups of
        // 2, 8, 16 and tasks fastest measured slower on 96 kHz -> 44.1 kHz, exact)
        tgroup_ = kchain_ > 0 ? 4 : 0;
        if (const char* v = std::getenv("KBSRC_MID_TGROUP")) tgroup_ = std::max(0, std::atoi(v));   // test switch
    }
    // H in the element format of the engine: (hi, lo) pairs or FP64
    H_.alloc((size_t)nH * (dd_ ? 16 : 8));
    if (dd_) hip_check(hipMemcpyAsync(H_.d(), H, (size_t)nH * 16, hipMemcpyDeviceToDevice, s), "kernel spectrum copy");
    else hip_check(launch_hi(H, nH, H_.d(), s), "kernel spectrum hi");
    const size_t elt = dd_ ? 32 : 16;
    const long long elems = std::max({(long long)fw_.h + 1, (long long)iv_.h + 1, fw_.c.work_elems(),
                                      iv_.c.work_elems()});   // + 1: the split's X[h]
    const size_t per_task = (size_t)elems * elt;
    ts_ = (long long)(per_task / sizeof(double));
    const bool blue = fw_.c.blue || iv_.c.blue;
    // tasks per launch: each work buffer gets up to 1/12 of free device memory, between 1 and 4 GiB (or the
    // caller's budget, for engines that share the device with others)
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    size_t cap = (size_t)4 << 30;
    if (const char* v = std::getenv("KBSRC_BUFFER_GIB"))   // tuning knob: the per-buffer cap, GiB
        if (std::atoll(v) > 0) cap = (size_t)std::atoll(v) << 30;
    const size_t budget = buffer_budget ? buffer_budget : std::clamp<size_t>(free_b / 12, (size_t)1 << 30, cap);
    tb_ = (int)std::max<long long>(1, std::min<long long>(ntask_, (long long)(budget / per_task)));
    if (const char* v = std::getenv("KBSRC_BATCH_TASKS")) {   // tuning knob: tasks per launch
        const long long want = std::atoll(v);
        if (want > 0) tb_ = (int)std::min<long long>(tb_, want);
    }
    for (int i = 0; i < (blue ? 4 : 2); ++i) w_[i].alloc(per_task * tb_);
    if (fw_.c.needs_global() || iv_.c.needs_global() || global_forced()) {
        g_[0].alloc(per_task * tb_);
        g_[1].alloc(per_task * tb_);
    }
    // the fused row pass's tables in residue-major order, where its kernel reads them so (made once per plan, so
    // that executes allocate nothing)
    if (fw_.c.four && !fw_.c.blue && !Jg_ && rowmid_residue_tables(dd_ ? POL_DDS : POL_F64, fw_.c.N2, fw_.c.N1, iv_.h)) {
        const int t = std::gcd(fw_.c.N1, iv_.h);
        const long long n = std::max({nH, (long long)fw_.h + 1, (long long)iv_.h + 1});
        Lr_ = (int)((n + t - 1) / t);
        const int er = dd_ ? 2 : 1, ec = dd_ ? 4 : 2;   // doubles per real and per complex entry
        Hr_.alloc((size_t)t * Lr_ * er * sizeof(double));
        wfr_.alloc((size_t)t * Lr_ * ec * sizeof(double));
        wir_.alloc((size_t)t * Lr_ * ec * sizeof(double));
        hip_check(launch_residue_major(H_.d(), nH, er, t, Lr_, Hr_.d(), s), "residue-major H");
        hip_check(launch_residue_major(fw_.w.d(), fw_.h + 1, ec, t, Lr_, wfr_.d(), s), "residue-major wf");
        hip_check(launch_residue_major(iv_.w.d(), iv_.h + 1, ec, t, Lr_, wir_.d(), s), "residue-major wi");
    }
    hip_check(hipStreamSynchronize(s), "engine set-up");   // H and the tables are in place; the caller may free H
}

void Engine::execute(const double* xx, double* out, hipStream_t s, PassTimes* times, int split, bool out_dd) {
    if (out_dd && !dd_) throw std::invalid_argument("option out of range");   // pairs: exact-mode plans only
    if (split) {   // tables of the split form, made on first use
        if (!dd_ || !split_slices_ok(split)) throw std::invalid_argument("option out of range");
        fw_.c.init_split(s);
        iv_.c.init_split(s);
        if (!adig_.count(split)) {
            const std::vector<uint64_t> ad = split_matrix_digits(split);
            adig_[split].upload(ad.data(), ad.size() * sizeof(uint64_t), s);
            hip_check(hipStreamSynchronize(s), "split matrix");
        }
        if (!g_[0].d()) {
            g_[0].alloc((size_t)ts_ * sizeof(double) * tb_);
            g_[1].alloc((size_t)ts_ * sizeof(double) * tb_);
        }
    }
    const Work w{{w_[0].d(), w_[1].d(), w_[2].d(), w_[3].d()}, {g_[0].d(), g_[1].d()}, tb_};
    run(xx, out, s, times, split, w, 0, -1, out_dd);
}

void Engine::execute_tasks(const double* xx, double* out, hipStream_t s, long long t0, long long t1) {
    if (t0 < 0 || t1 > ntask_ || t0 > t1) throw std::invalid_argument("task range");
    const Work w{{w_[0].d(), w_[1].d(), w_[2].d(), w_[3].d()}, {g_[0].d(), g_[1].d()}, tb_};
    run(xx, out, s, nullptr, 0, w, t0, t1);
}

namespace {
constexpr size_t kCarveAlign = 256;
size_t carve_bytes(size_t tasks, size_t per_task) { return (tasks * per_task + kCarveAlign - 1) / kCarveAlign * kCarveAlign; }
}  // namespace

size_t Engine::task_bytes() const {
    const bool blue = fw_.c.blue || iv_.c.blue;
    const bool glob = fw_.c.needs_global() || iv_.c.needs_global() || global_forced();
    return (size_t)ts_ * sizeof(double) * ((blue ? 4 : 2) + (glob ? 2 : 0));
}

void Engine::execute_in(const double* xx, double* out, hipStream_t s, double* mem, size_t bytes) {
    const bool blue = fw_.c.blue || iv_.c.blue;
    const bool glob = fw_.c.needs_global() || iv_.c.needs_global() || global_forced();
    const int nw = blue ? 4 : 2, nbuf = nw + (glob ? 2 : 0);
    const size_t per = (size_t)ts_ * sizeof(double);   // one task in one buffer
    if (!mem || reinterpret_cast<uintptr_t>(mem) % kCarveAlign)
        throw std::invalid_argument("scratch memory must be 256-byte aligned");
    size_t tb = bytes / nbuf / per;   // tasks per launch, each buffer's start kept aligned
    while (tb > 0 && nbuf * carve_bytes(tb, per) > bytes) --tb;
    if (tb == 0)
        throw std::invalid_argument("plan too large for the static scratch: one task needs " +
                                    std::to_string(nbuf * carve_bytes(1, per)) + " bytes of work buffers, " +
                                    std::to_string(bytes) + " available");
    tb = std::min<size_t>(tb, (size_t)ntask_);
    Work w{{nullptr, nullptr, nullptr, nullptr}, {nullptr, nullptr}, (int)std::min<size_t>(tb, 0x7fffffff)};
    char* p = reinterpret_cast<char*>(mem);
    for (int i = 0; i < nbuf; ++i, p += carve_bytes(tb, per)) {
        double* b = reinterpret_cast<double*>(p);
        if (i < nw) w.w[i] = b;
        else w.g[i - nw] = b;
    }
    run(xx, out, s, nullptr, 0, w);
}

void Engine::run(const double* xx, double* out, hipStream_t s, PassTimes* times, int split, const Work& w,
                 long long tbeg, long long tend, bool out_dd) {
    if (tend < 0) tend = ntask_;
    PassTimes local;
    const int pol_four = dd_ ? POL_DDS : POL_F64;
    const bool fused = !split && use_fixed() && env_flag("KBSRC_FUSED") && !Jg_ && fw_.c.four && iv_.c.four &&
                       has_fixed(pol_four, fw_.c.N2, LD_CPLX, ST_SPLIT) && has_fixed(pol_four, iv_.c.N1, LD_FOLD, ST_CPLX);
    RunCtx c = context(xx, out, s, split, w);
    c.proto.out_dd = out_dd ? 1 : 0;
    c.times = times ? &local : nullptr;
    PassArgs& a = c.proto;
    double* W0 = w.w[0];
    double* W1 = w.w[1];
    double* W2 = w.w[2];
    double* W3 = w.w[3];
    for (long long t0 = tbeg; t0 < tend; t0 += w.tb) {
        const int nt = (int)std::min<long long>(w.tb, tend - t0);
        a.ntask = nt;
        a.task0 = t0;
        if (fused) {
            // padded input -> (W1) -> X = split of z (W0) -> bin map and pre-split folded into the inverse columns
            double* const sf[3] = {W1, W2, W3};
            run_cfft(fw_.c, c, LD_REAL_F64, xx, ST_SPLIT, W0, sf, &local.fwd_col, &local.fwd_row);
            run_cfft(iv_.c, c, LD_FOLD, W0, ST_OUT, nullptr, sf, &local.inv_col, &local.inv_row);
            continue;
        }
        // split, x H, bin map, inverse pre-split: z (W0) -> z' (W1)
        MidArgs m;
        m.z = W0; m.z_tstride = ts_; m.zi = W1; m.zi_tstride = ts_;
        m.H = H_.d(); m.wf = fw_.w.d(); m.wi = iv_.w.d();
        m.h = fw_.h; m.hinv = iv_.h; m.Nb = p_.Nb; m.Mb = p_.Mb; m.K = p_.Mb / 2 + 1;
        m.jmax = std::min(p_.Nb / 2, p_.Mb - 1); m.up = p_.up; m.ntask = nt; m.Jg = Jg_;
        m.kchain = kchain_;
        m.tgroup = tgroup_;
        m.Hr = Hr_.d(); m.wfr = wfr_.d(); m.wir = wir_.d(); m.Lr = Lr_;
        // forward real FFT: padded input -> z (W0); the last pass stores only the bins the bin map reads. Where the
        // plan allows, the row pass and the split/fold step run as one launch: column output (W1) -> z' (W0)
        double* zi = W1;
        const std::function<void(PassArgs&)> rows = [&](PassArgs& b) {
            if (!pass_offsets_fit32(b)) {   // the row pass and the split/fold kernel, as without the hook
                do_pass(c, &local.fwd_row, pol_four, LD_CPLX, ST_CPLX, b, "row pass");
                timed(c, &local.mid, [&] { hip_check(launch_mid(dd_ ? POL_DDA : POL_F64, m, s), "mid kernel"); });
                return;
            }
            set_prune_window(b, ST_CPLX);
            MidArgs mr = m;
            mr.z = nullptr;
            mr.zi = W0;
            zi = W0;
            timed(c, &local.fwd_row, [&] {
                hip_check(b.split ? launch_rowmid_split(b, mr, s) : launch_rowmid(pol_four, b, mr, s),
                          "row pass with the split/fold");
            });
        };
        const char* sg = std::getenv("KBSRC_SPLIT_GLOBAL");   // a split route forced: the separate kernels
        const bool rowmid = use_fixed() && !global_forced() && !Jg_ && fw_.c.four && !fw_.c.blue &&
                            !env_off("KBSRC_ROWMID") &&
                            (split ? dd_ && !(sg && (sg[0] == '1' || sg[0] == '2')) &&
                                         rowmid_split_available(fw_.c.N2, fw_.c.N1, iv_.h, split)
                                   : rowmid_available(pol_four, fw_.c.N2, fw_.c.N1, iv_.h));
        double* const sf[3] = {W1, W2, W3};
        a.zkeep = zkeep_;
        a.zh = fw_.h;
        c.rows = rowmid ? &rows : nullptr;
        run_cfft(fw_.c, c, LD_REAL_F64, xx, ST_CPLX, W0, sf, &local.fwd_col, &local.fwd_row);
        c.rows = nullptr;
        a.zkeep = -1;
        if (!rowmid)
            timed(c, &local.mid, [&] { hip_check(launch_mid(dd_ ? POL_DDA : POL_F64, m, s), "mid kernel"); });
        // inverse: z' -> output samples
        double* const si[3] = {zi == W1 ? W0 : W1, W2, W3};
        run_cfft(iv_.c, c, LD_CPLX, zi, ST_OUT, nullptr, si, &local.inv_col, &local.inv_row);
    }
    if (times) {
        local.collect();
        times->fwd_col += local.fwd_col;
        times->fwd_row += local.fwd_row;
        times->mid += local.mid;
        times->inv_col += local.inv_col;
        times->inv_row += local.inv_row;
        times->launches += local.launches;
    }
}

RunCtx Engine::context(const double* xx, double* out, hipStream_t s, int split, const Work& w) {
    (void)xx;
    RunCtx c;
    c.ts = ts_; c.k = &k_; c.dd = dd_; c.lane = true; c.s = s;
    PassArgs& a = c.proto;
    a.gscr0 = w.g[0];
    a.gscr1 = w.g[1];
    a.tdesc = tdesc_;
    a.split = split;
    a.adig = split ? reinterpret_cast<const long*>(adig_.at(split).d()) : nullptr;
    a.nblk = p_.nblk;
    a.in_cstride = p_.xstride;
    a.in_bstride = (long long)p_.hop * p_.Q;
    a.out = out;
    a.nout = p_.nout;
    a.o = p_.o;
    a.margP = (long long)p_.marg * p_.P;
    a.hopP = (long long)p_.hop * p_.P;
    FuseArgs& fa = a.f;
    fa.H = H_.d(); fa.wf = fw_.w.d(); fa.wi = iv_.w.d();
    fa.h = fw_.h; fa.hinv = iv_.h; fa.Nb = p_.Nb; fa.Mb = p_.Mb; fa.K = p_.Mb / 2 + 1;
    fa.jmax = std::min(p_.Nb / 2, p_.Mb - 1); fa.up = p_.up; fa.xmax = p_.up ? fw_.h : fa.jmax;
    return c;
}

bool Engine::describe(const double* xx, double* out, EngineSteps& steps) {
    if (ntask_ > tb_) return false;
    bool ok = true;
    const Work w{{w_[0].d(), w_[1].d(), w_[2].d(), w_[3].d()}, {g_[0].d(), g_[1].d()}, tb_};
    RunCtx c = context(xx, out, nullptr, 0, w);
    c.collect_ok = &ok;
    PassArgs& a = c.proto;
    a.ntask = (int)ntask_;
    a.task0 = 0;
    double* W0 = w_[0].d();
    double* W1 = w_[1].d();
    double* W2 = w_[2].d();
    double* W3 = w_[3].d();
    double* const sf[3] = {W1, W2, W3};
    c.collect = &steps.fwd;
    a.zkeep = zkeep_;
    a.zh = fw_.h;
    run_cfft(fw_.c, c, LD_REAL_F64, xx, ST_CPLX, W0, sf, nullptr, nullptr);
    a.zkeep = -1;
    MidArgs& m = steps.mid;
    m.z = W0; m.z_tstride = ts_; m.zi = W1; m.zi_tstride = ts_;
    m.H = H_.d(); m.wf = fw_.w.d(); m.wi = iv_.w.d();
    m.h = fw_.h; m.hinv = iv_.h; m.Nb = p_.Nb; m.Mb = p_.Mb; m.K = p_.Mb / 2 + 1;
    m.jmax = std::min(p_.Nb / 2, p_.Mb - 1); m.up = p_.up; m.ntask = (int)ntask_; m.Jg = Jg_;
    m.kchain = kchain_;
    m.tgroup = tgroup_;
    steps.mid_policy = dd_ ? POL_DDA : POL_F64;
    double* const si[3] = {W0, W2, W3};
    c.collect = &steps.inv;
    run_cfft(iv_.c, c, LD_CPLX, W1, ST_OUT, nullptr, si, nullptr, nullptr);
    return ok && mid_workgroups(m) > 0;
}

}  // namespace kbsrc::gpu
