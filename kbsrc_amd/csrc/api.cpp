// C-ABI entry points (include/kbsrc_gpu.h) and the C++ API (include/kbsrc/kbsrc.hpp) over the GPU engine.
#include <xmmintrin.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <climits>
#include <numeric>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <vector>

#include "copy_crew.hpp"
#include "engine.hpp"
#include "kbsrc/kbsrc.hpp"
#include "kbsrc_gpu.h"

#pragma clang fp contract(off)

static_assert(kbsrc::gpu::kScratchBytes == (size_t)KBSRC_NOALLOC_SCRATCH_BYTES, "static scratch size");

namespace {

using namespace kbsrc::gpu;

thread_local std::string g_error;

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// Runs f with the SSE unit, where the host's double-double arithmetic runs, rounding to nearest when `exact` is set,
// and gives the caller's rounding mode back afterwards, also when f throws (the reference computes Exact in
// round-to-nearest whatever the caller's mode: src/dispatch.cpp:31-36,73). Only the MXCSR rounding field changes:
// exception flags raised meanwhile stay raised, and FTZ/DAZ and the x87 control word are left as they are.
template <class F> void in_nearest_mode(bool exact, F&& f) {
    const unsigned caller = _MM_GET_ROUNDING_MODE();
    if (!exact || caller == _MM_ROUND_NEAREST) {
        f();
        return;
    }
    _MM_SET_ROUNDING_MODE(_MM_ROUND_NEAREST);
    try {
        f();
    } catch (...) {
        _MM_SET_ROUNDING_MODE(caller);
        throw;
    }
    _MM_SET_ROUNDING_MODE(caller);
}

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

// One captured execute (a plan's or a batch's) as an instantiated HIP graph.
struct kbsrc_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::shared_ptr<const void> keep;   // device tables the graph's kernels read (a batch's one-launch step tables)
    kbsrc_graph() = default;
    kbsrc_graph(const kbsrc_graph&) = delete;
    kbsrc_graph& operator=(const kbsrc_graph&) = delete;
    ~kbsrc_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

struct kbsrc_plan {
    Plan p;
    kbsrc_options opt;
    std::unique_ptr<Engine> eng;
    int Jg = 0;
    double design_seconds = 0;
};

// Many jobs (files, channel counts, rate pairs, lengths) in one set of device buffers. Jobs whose plans share a
// shape (rates and block size, hence transform lengths, kernel spectrum and tables) form a group: one engine, one
// launch per pass over all the group's (job, channel, block) tasks through a task table. Groups run concurrently
// on a few streams. Equal-rate jobs are copies.
struct kbsrc_batch {
    struct Job {
        Plan p;
        long long in_off = 0, in_len = 0, out_off = 0, out_len = 0;
        int group = -1;
    };
    struct Group {
        Plan p;
        std::vector<int> jobs;
        long long ntask = 0;
        int stream = 0;   // index into streams
        DevBuf tdesc;
        std::unique_ptr<Engine> eng;
    };
    kbsrc_options opt;
    std::vector<Job> jobs;
    std::vector<Group> groups;
    long long in_total = 0, out_total = 0;
    double design_seconds = 0;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;   // one per stream, plus the fork event
    // One launch per step over every group (for batches of short jobs, whose per-shape launches leave the GPU
    // idle): the generic pass kernel and the split/multiply/fold kernel over descriptor tables, made per pair of
    // device buffers.
    struct StepLaunch {
        bool mid = false;
        int policy = 0, ld = 0, st = 0, n = 0, total = 0;
        size_t lds = 0;
        DevBuf descs, starts;
    };
    struct Fused {
        const double* in = nullptr;
        double* out = nullptr;
        std::vector<StepLaunch> launches;
    };
    bool fuse = false;
    // by buffers, at most four kept here; a captured graph holds its own reference, so an entry evicted here stays
    // alive while a graph that baked in its tables exists
    std::vector<std::shared_ptr<Fused>> fused;
    ~kbsrc_batch() {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (hipStream_t q : streams) (void)hipStreamDestroy(q);
    }
};

namespace {

// The kernel design for a plan on the GPU (src/resampler.cpp:187-206): taps, then the spectrum H as (hi, lo) pairs.
struct Design {
    DevBuf H;
    long long nH = 0;
    int Jg = 0;
};
Design design_kernel(const Plan& p, const kbsrc_options& o, const DeviceConsts& k, hipStream_t s) {
    Design d;
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
        d.Jg = (int)std::min<long long>(Nk / 2, (long long)std::ceil((p.mn / 2.0 + 4 * p.fml) * p.Nb / p.fi));
        d.H = kernel_spectrum_gpu(h2.d(), L2, (int)Nk, (double)p.Nb, k, s);
        d.nH = Nk / 2 + 1;
    } else {
        const DevBuf h = kernel_taps_gpu(p.Lk, p.c2 / p.gc, p.fk / p.gc, o.beta, s);
        const int n = p.up ? p.Mb : p.Nb;
        d.H = kernel_spectrum_gpu(h.d(), p.Lk, n, (double)p.Nb, k, s);
        d.nH = n / 2 + 1;
    }
    return d;
}

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
    const Design d = design_kernel(p, o, k, s);
    pl->Jg = d.Jg;
    pl->eng = std::make_unique<Engine>(p, d.H.d(), d.nH, d.Jg, s);   // keeps its own copy of H
    pl->design_seconds = now() - t0;
    return pl;
}

// The Burg head and tail extensions of every channel (src/lpc.cpp:13-41, src/resampler.cpp:176-185) from the first
// and last edge_frames() interleaved frames on the host, ch x (E + 32) doubles each, zero without LPC. Every
// (channel, end) is its own job; up to 16 worker threads take the jobs from construction on, each in the rounding
// mode of the thread that made the work, so the caller can upload the input meanwhile. wait() joins them.
class EdgeWork {
public:
    EdgeWork(const Plan& p, bool lpc, const double* first, const double* last)
        : ext_(p.E + 32), head_((size_t)p.ch * ext_, 0.0), tail_((size_t)p.ch * ext_, 0.0) {
        if (!lpc || p.nin == 0) return;
        const int jobs = 2 * p.ch;
        const unsigned mode = _MM_GET_ROUNDING_MODE();
        auto worker = [this, &p, first, last, jobs, mode] {
            _MM_SET_ROUNDING_MODE(mode);
            try {
                for (int j; (j = next_.fetch_add(1)) < jobs;) {
                    const int c = j >> 1;
                    const bool at_tail = j & 1;
                    const std::vector<double> e = edge_extension(p, first, last, c, at_tail);
                    std::memcpy((at_tail ? tail_ : head_).data() + (size_t)c * ext_, e.data(), sizeof(double) * ext_);
                }
            } catch (...) {
                std::lock_guard<std::mutex> lock(err_mu_);
                if (!err_) err_ = std::current_exception();
            }
        };
        const int nt = std::min(jobs, 16);
        for (int t = 0; t < nt; ++t) {
            try {
                threads_.emplace_back(worker);
            } catch (...) {   // no thread: the jobs left run in wait()
                break;
            }
        }
        finish_ = worker;
    }
    ~EdgeWork() { join(); }
    EdgeWork(const EdgeWork&) = delete;
    EdgeWork& operator=(const EdgeWork&) = delete;
    void wait() {
        if (finish_) finish_();   // whatever no worker has taken
        join();
        if (err_) std::rethrow_exception(err_);
    }
    long long ext() const { return ext_; }
    const std::vector<double>& head() const { return head_; }
    const std::vector<double>& tail() const { return tail_; }

private:
    void join() {
        for (auto& t : threads_)
            if (t.joinable()) t.join();
        threads_.clear();
    }
    long long ext_;
    std::vector<double> head_, tail_;
    std::atomic<int> next_{0};
    std::vector<std::thread> threads_;
    std::function<void()> finish_;
    std::mutex err_mu_;
    std::exception_ptr err_;
};

// Staging (src/resampler.cpp:168-186) from the interleaved input in device memory (d_in) and the finished edge
// extensions: deinterleave, pad and place on the GPU.
// The edge extensions on the device: every channel's head, then every channel's tail (ext() doubles each). Waits for
// the edge work; the upload is complete on return.
DevBuf upload_edges(EdgeWork& edges, hipStream_t s) {
    edges.wait();
    const std::vector<double>& head = edges.head();
    const std::vector<double>& tail = edges.tail();
    std::vector<double> both(head.size() + tail.size());
    std::copy(head.begin(), head.end(), both.begin());
    std::copy(tail.begin(), tail.end(), both.begin() + (std::ptrdiff_t)head.size());
    DevBuf d;
    d.upload(both.data(), sizeof(double) * both.size(), s);
    return d;
}

// Staging on the device from the interleaved input d_in and the uploaded edges (upload_edges). sync: return once the
// padded input is complete (else only queued on s).
void stage_edges(const Plan& p, const double* d_in, const EdgeWork& edges, const DevBuf& d_edges, double* d_padded,
                 hipStream_t s, bool sync) {
    StageArgs a;
    a.in = d_in; a.head = d_edges.d(); a.tail = d_edges.d() + edges.head().size(); a.xx = d_padded;
    a.nin = p.nin; a.tot = p.tot; a.stride = p.xstride; a.off = p.base + p.pre; a.ext = edges.ext(); a.ch = p.ch;
    hip_check(launch_stage(a, s), "staging kernel");
    if (sync) hip_check(hipStreamSynchronize(s), "staging sync");
}

void stage_dev(const Plan& p, const double* d_in, EdgeWork& edges, double* d_padded, hipStream_t s) {
    const DevBuf d_edges = upload_edges(edges, s);
    stage_edges(p, d_in, edges, d_edges, d_padded, s, true);
}

// Host <-> device copies of caller buffers through pinned staging: a pair of 64 MiB pinned buffers used as a ring of
// 16 MiB slots. Each slot is filled or drained on the host by a crew of copy threads (page faults on a fresh caller
// buffer are taken in parallel) while the DMA of the other slots runs; with small slots the first and the last host
// copy, which nothing overlaps, are short. A pair has its own stream, which first waits (on the device) for the work
// already queued on the caller's stream, so the copies keep the caller's order without any thread blocking on the
// caller's stream. Pairs come from a small pool made on first use and kept for the process; a lease returns its pair
// only once the pair's stream has drained, whatever the exit path. Without pinned memory the copies go straight from
// and to the caller's buffer on the caller's stream. Synchronous: returns when done.
struct ChunkPair {
    static constexpr size_t kBytes = 64u << 20;   // each pinned buffer
    static constexpr size_t kSlot = 16u << 20;
    static constexpr int kSlots = (int)(2 * kBytes / kSlot);
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[kSlots] = {};              // the last DMA issued from / into slot i
    hipEvent_t order = nullptr;              // recorded on the caller's stream
    hipStream_t st = nullptr;
    char* slot(int i) const { return static_cast<char*>(buf[i / (kSlots / 2)]) + (size_t)(i % (kSlots / 2)) * kSlot; }
};

class ChunkPool {
public:
    // a free pair, a new one (at most kMax), or nullptr when pinned memory or the pair's resources cannot be had
    ChunkPair* acquire() {
        std::unique_lock<std::mutex> lock(m_);
        for (;;) {
            if (!free_.empty()) {
                ChunkPair* c = free_.back();
                free_.pop_back();
                return c;
            }
            if (failed_) return nullptr;
            if (made_ < kMax) {
                ++made_;
                lock.unlock();
                ChunkPair* c = make();
                lock.lock();
                if (!c) {
                    --made_;
                    failed_ = true;
                    cv_.notify_all();
                }
                return c;
            }
            cv_.wait(lock);
        }
    }
    void release(ChunkPair* c) {
        {
            std::lock_guard<std::mutex> lock(m_);
            free_.push_back(c);
        }
        cv_.notify_one();
    }

private:
    static constexpr int kMax = 4;
    static ChunkPair* make() {
        auto c = std::make_unique<ChunkPair>();
        bool ok = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&c->order, hipEventDisableTiming) == hipSuccess;
        for (int i = 0; i < 2 && ok; ++i) ok = hipHostMalloc(&c->buf[i], ChunkPair::kBytes, hipHostMallocDefault) == hipSuccess;
        for (int i = 0; i < ChunkPair::kSlots && ok; ++i)
            ok = hipEventCreateWithFlags(&c->ev[i], hipEventDisableTiming) == hipSuccess;
        if (ok) return c.release();
        (void)hipGetLastError();
        for (int i = 0; i < 2; ++i)
            if (c->buf[i]) (void)hipHostFree(c->buf[i]);
        for (int i = 0; i < ChunkPair::kSlots; ++i)
            if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
        if (c->order) (void)hipEventDestroy(c->order);
        if (c->st) (void)hipStreamDestroy(c->st);
        return nullptr;
    }
    std::mutex m_;
    std::condition_variable cv_;
    std::vector<ChunkPair*> free_;
    int made_ = 0;
    bool failed_ = false;
};
ChunkPool& chunk_pool() {
    static ChunkPool* p = new ChunkPool;   // kept until the process ends
    return *p;
}

// A pair on loan: on every exit path its stream is drained (no DMA still reads or fills its chunks) before the pair
// goes back to the pool.
class ChunkLease {
public:
    ChunkLease() : c_(chunk_pool().acquire()) {}
    ~ChunkLease() {
        if (!c_) return;
        (void)hipStreamSynchronize(c_->st);
        chunk_pool().release(c_);
    }
    ChunkLease(const ChunkLease&) = delete;
    ChunkLease& operator=(const ChunkLease&) = delete;
    ChunkPair* get() const { return c_; }

private:
    ChunkPair* c_;
};

void upload_host(void* d_dst, const void* h_src, size_t bytes, hipStream_t s) {
    ChunkLease lease;
    ChunkPair* c = lease.get();
    if (!c) {   // no pinned memory: straight from the caller's buffer
        hip_check(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, s), "upload");
        hip_check(hipStreamSynchronize(s), "upload sync");
        return;
    }
    hip_check(hipEventRecord(c->order, s), "staging order");
    hip_check(hipStreamWaitEvent(c->st, c->order, 0), "staging order");
    const size_t K = ChunkPair::kSlot;
    for (size_t off = 0, k = 0; off < bytes; off += K, ++k) {
        const int b = (int)(k % ChunkPair::kSlots);
        const size_t n = std::min(K, bytes - off);
        if (k >= (size_t)ChunkPair::kSlots) hip_check(hipEventSynchronize(c->ev[b]), "staging wait");   // slot free
        CopyCrew::get().copy(c->slot(b), static_cast<const char*>(h_src) + off, n);
        hip_check(hipMemcpyAsync(static_cast<char*>(d_dst) + off, c->slot(b), n, hipMemcpyHostToDevice, c->st),
                  "upload");
        hip_check(hipEventRecord(c->ev[b], c->st), "staging event");
    }
    hip_check(hipStreamSynchronize(c->st), "upload sync");
}

// Pieces of a device buffer downloaded into host memory in order, each once the work before its event is done: the
// first pieces go down while the work that writes the later ones still runs. Staged like upload_host, in reverse.
struct DownPiece {
    void* host;
    const void* dev;
    size_t bytes;
    hipEvent_t ready;
};
void download_pieces(const std::vector<DownPiece>& pieces) {
    ChunkLease lease;
    ChunkPair* c = lease.get();
    if (!c) {   // no pinned memory: straight into the caller's buffer
        for (const DownPiece& q : pieces) {
            hip_check(hipEventSynchronize(q.ready), "download wait");
            hip_check(hipMemcpy(q.host, q.dev, q.bytes, hipMemcpyDeviceToHost), "download");
        }
        return;
    }
    struct Job {
        size_t piece, off, n;
    };
    std::vector<Job> jobs;
    const size_t K = ChunkPair::kSlot;
    for (size_t i = 0; i < pieces.size(); ++i)
        for (size_t off = 0; off < pieces[i].bytes; off += K) jobs.push_back({i, off, std::min(K, pieces[i].bytes - off)});
    auto issue = [&](size_t k) {
        const Job& j = jobs[k];
        const DownPiece& q = pieces[j.piece];
        if (j.off == 0) hip_check(hipStreamWaitEvent(c->st, q.ready, 0), "download order");
        const int b = (int)(k % ChunkPair::kSlots);
        hip_check(hipMemcpyAsync(c->slot(b), static_cast<const char*>(q.dev) + j.off, j.n, hipMemcpyDeviceToHost, c->st),
                  "download");
        hip_check(hipEventRecord(c->ev[b], c->st), "staging event");
    };
    for (size_t k = 0; k < jobs.size() && k < (size_t)ChunkPair::kSlots; ++k) issue(k);
    for (size_t k = 0; k < jobs.size(); ++k) {
        const int b = (int)(k % ChunkPair::kSlots);
        hip_check(hipEventSynchronize(c->ev[b]), "staging wait");
        const Job& j = jobs[k];
        CopyCrew::get().copy(static_cast<char*>(pieces[j.piece].host) + j.off, c->slot(b), j.n);
        if (k + ChunkPair::kSlots < jobs.size()) issue(k + ChunkPair::kSlots);   // into the slot just emptied
    }
}

// Non-blocking streams on the current device kept for reuse by host-buffer conversions (making one costs ~0.1 ms,
// destroying one ~0.7 ms, measured on MI355X). A lease takes one (or makes one) and gives it back drained.
class StreamLease {
public:
    StreamLease() {
        hip_check(hipGetDevice(&dev_), "device");
        {
            std::lock_guard<std::mutex> lock(mu());
            auto& v = free_list()[dev_];
            if (!v.empty()) {
                s_ = v.back();
                v.pop_back();
            }
        }
        if (!s_) hip_check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "stream");
    }
    ~StreamLease() {
        if (!s_) return;
        if (hipStreamSynchronize(s_) != hipSuccess) {   // a failed stream is not reused
            (void)hipGetLastError();
            (void)hipStreamDestroy(s_);
            return;
        }
        std::lock_guard<std::mutex> lock(mu());
        free_list()[dev_].push_back(s_);
    }
    StreamLease(const StreamLease&) = delete;
    StreamLease& operator=(const StreamLease&) = delete;
    hipStream_t get() const { return s_; }

private:
    static std::mutex& mu() {
        static std::mutex* m = new std::mutex;
        return *m;
    }
    static std::map<int, std::vector<hipStream_t>>& free_list() {
        static auto* f = new std::map<int, std::vector<hipStream_t>>;   // kept until the process ends
        return *f;
    }
    int dev_ = 0;
    hipStream_t s_ = nullptr;
};

// Staging from host interleaved input: the edge extensions on worker threads while the input goes up, then as above.
void stage(const Plan& p, bool lpc, const double* in, double* d_padded, hipStream_t s) {
    const long long m = edge_frames(p);
    EdgeWork edges(p, lpc, in, in + (p.nin - m) * p.ch);
    DevBuf din;
    if (p.nin > 0) {
        din.alloc(sizeof(double) * (size_t)(p.nin * p.ch));
        upload_host(din.d(), in, sizeof(double) * (size_t)(p.nin * p.ch), s);
    } else {
        din.alloc(sizeof(double));
    }
    stage_dev(p, din.d(), edges, d_padded, s);
}

// Staging from device interleaved input: the edge frames come down to the host for the Burg recursion.
void stage_device(const Plan& p, bool lpc, const double* d_in, double* d_padded, hipStream_t s) {
    const long long m = edge_frames(p);
    std::vector<double> first((size_t)std::max(1LL, m * p.ch)), last((size_t)std::max(1LL, m * p.ch));
    if (m > 0 && lpc) {
        const size_t bytes = sizeof(double) * (size_t)(m * p.ch);
        hip_check(hipMemcpyAsync(first.data(), d_in, bytes, hipMemcpyDeviceToHost, s), "edge download");
        hip_check(hipMemcpyAsync(last.data(), d_in + (p.nin - m) * p.ch, bytes, hipMemcpyDeviceToHost, s),
                  "edge download");
        hip_check(hipStreamSynchronize(s), "edge sync");
    }
    EdgeWork edges(p, lpc, first.data(), last.data());
    stage_dev(p, d_in, edges, d_padded, s);
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
    in_nearest_mode(o.precision != KBSRC_FAST, [&] {
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
        // the plan's integers first (host, microseconds): the reference's rejections in its order, and the sizes
        check_options(o);
        const Plan P = make_plan(frames, channels, rate_in, rate_out, o.precision == KBSRC_FAST, o.length, o.beta,
                                 o.block);
        const StreamLease sl;
        const hipStream_t s = sl.get();
        // every device buffer of this call, the plan's included, is allocated and freed in stream order (on s, the
        // design's on its own stream), so that no allocation or free waits for the copies and kernels in flight;
        // the plan is released before the scope ends, after the last work on s
        const PoolScope scope(s);
        // the kernel design runs on a thread and a stream of its own and the Burg edges on worker threads, all in
        // this thread's rounding mode, while this thread uploads the input; the design thread then uploads the edges
        const long long m = edge_frames(P);
        EdgeWork edges(P, o.lpc_edges, in, in + (P.nin - m) * P.ch);
        std::unique_ptr<kbsrc_plan> pl;
        DevBuf dedge;
        std::exception_ptr design_err;
        const unsigned mode = _MM_GET_ROUNDING_MODE();
        std::thread design([&] {
            _MM_SET_ROUNDING_MODE(mode);
            try {
                const StreamLease ds;
                const PoolScope dscope(ds.get());
                pl = make(frames, channels, rate_in, rate_out, o, ds.get());
                hip_check(hipStreamSynchronize(ds.get()), "design sync");
                dedge = upload_edges(edges, ds.get());
            } catch (...) {
                design_err = std::current_exception();
            }
        });
        struct Joiner {
            std::thread& t;
            ~Joiner() {
                if (t.joinable()) t.join();
            }
        } joiner{design};
        DevBuf din(sizeof(double) * (size_t)std::max(1LL, P.nin * P.ch));
        if (P.nin > 0) upload_host(din.d(), in, sizeof(double) * (size_t)(P.nin * P.ch), s);
        DevBuf xx((size_t)P.ch * P.xstride * sizeof(double));
        DevBuf dout((size_t)std::max(1LL, P.ch * P.nout) * sizeof(double));
        design.join();
        if (design_err) std::rethrow_exception(design_err);
        const Plan& p = pl->p;
        stage_edges(p, din.d(), edges, dedge, xx.d(), s, false);
        din.release();
        pl->design_seconds = now() - t0;   // planning, design and staging, the input's upload overlapping them
        const double t2 = now();
        // the engine in up to four parts of the channel-major task order: a part's output samples (planar, the
        // same offsets on both sides) go down while the next part computes
        Engine& eng = *pl->eng;
        const long long nt = eng.tasks(), hopP = (long long)p.hop * p.P;
        const int parts = (int)std::min<long long>(nt, 4);
        struct Events {
            std::vector<hipEvent_t> e;
            ~Events() {
                for (hipEvent_t x : e) (void)hipEventDestroy(x);
            }
        } ev;
        std::vector<DownPiece> pieces;
        // first output sample of block b (blocks tile [0, nout) in order, src/resampler.cpp:80-84)
        auto start = [&](long long b) { return b >= p.nblk ? p.nout : std::clamp<long long>(b * hopP - p.o, 0, p.nout); };
        for (int k = 0; k < parts; ++k) {
            const long long t0 = nt * k / parts, t1 = nt * (k + 1) / parts;
            eng.execute_tasks(xx.d(), dout.d(), s, t0, t1);
            hipEvent_t e = nullptr;
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
            ev.e.push_back(e);
            hip_check(hipEventRecord(e, s), "event");
            for (long long t = t0; t < t1;) {   // per channel, the samples of blocks [b0, b1)
                const long long c = t / p.nblk, b0 = t - c * p.nblk, b1 = std::min(p.nblk, b0 + (t1 - t));
                const long long lo = start(b0), hi = b1 >= p.nblk ? p.nout : start(b1);
                if (hi > lo)
                    pieces.push_back({out + c * p.nout + lo, dout.d() + c * p.nout + lo, sizeof(double) * (hi - lo), e});
                t += b1 - b0;
            }
        }
        if (!pieces.empty()) download_pieces(pieces);
        hip_check(hipStreamSynchronize(s), "engine sync");
        fill_stats(*pl, now() - t2, st);
    });
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
    return guarded(
        [&] { resample_host(in, frames, channels, rate_in, rate_out, opt ? *opt : defaults(), out, stats); });
}

int kbsrc_plan_create(long long frames, int channels, int rate_in, int rate_out, const kbsrc_options* opt,
                      kbsrc_plan** plan) {
    return guarded([&] {
        if (!plan) throw std::invalid_argument("bad input buffer");
        const kbsrc_options o = opt ? *opt : defaults();
        in_nearest_mode(o.precision != KBSRC_FAST, [&] {
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
        in_nearest_mode(pl->opt.precision != KBSRC_FAST, [&] {
            stage_device(pl->p, pl->opt.lpc_edges, d_interleaved, d_padded, static_cast<hipStream_t>(stream));
        });
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
        in_nearest_mode(pl->opt.precision != KBSRC_FAST, [&] {
            stage(pl->p, pl->opt.lpc_edges, in, d_padded, static_cast<hipStream_t>(stream));
        });
    });
}

int kbsrc_plan_execute(kbsrc_plan* pl, const double* d_padded, double* d_out, void* stream) {
    return guarded([&] {
        if (!pl || !pl->eng || !d_padded || !d_out) throw std::invalid_argument("bad input buffer");
        pl->eng->execute(d_padded, d_out, static_cast<hipStream_t>(stream));
    });
}

int kbsrc_plan_execute_noalloc(kbsrc_plan* pl, const double* d_padded, double* d_out, void* stream) {
    return guarded([&] {
        if (!pl || !pl->eng || !d_padded || !d_out) throw std::invalid_argument("bad input buffer");
        // the work buffers come from the library's static device array (its address is looked up once per device)
        pl->eng->execute_in(d_padded, d_out, static_cast<hipStream_t>(stream), scratch_memory(), kScratchBytes);
    });
}

int kbsrc_plan_execute_split(kbsrc_plan* pl, const double* d_padded, double* d_out, int slices, void* stream) {
    return guarded([&] {
        if (!pl || !pl->eng || !d_padded || !d_out) throw std::invalid_argument("bad input buffer");
        // the split tables are made in host dd arithmetic on first use
        in_nearest_mode(true, [&] {
            pl->eng->execute(d_padded, d_out, static_cast<hipStream_t>(stream), nullptr, slices);
        });
    });
}

int kbsrc_plan_execute_dd(kbsrc_plan* pl, const double* d_padded, double* d_out_pairs, int slices, void* stream) {
    return guarded([&] {
        if (!pl || !pl->eng || !d_padded || !d_out_pairs) throw std::invalid_argument("bad input buffer");
        if (slices < 0) throw std::invalid_argument("option out of range");
        in_nearest_mode(true, [&] {
            pl->eng->execute(d_padded, d_out_pairs, static_cast<hipStream_t>(stream), nullptr, slices, true);
        });
    });
}

int kbsrc_plan_execute_timed(kbsrc_plan* pl, const double* d_padded, double* d_out, void* stream, double* ms) {
    return guarded([&] {
        if (!pl || !pl->eng || !d_padded || !d_out || !ms) throw std::invalid_argument("bad input buffer");
        PassTimes t;
        pl->eng->execute(d_padded, d_out, static_cast<hipStream_t>(stream), &t);
        hip_check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "timed execute sync");
        ms[0] = t.fwd_col; ms[1] = t.fwd_row; ms[2] = t.mid; ms[3] = t.inv_col; ms[4] = t.inv_row;
    });
}

void kbsrc_plan_destroy(kbsrc_plan* pl) { delete pl; }

int kbsrc_batch_create(int njobs, const kbsrc_job* jobs, const kbsrc_options* opt, kbsrc_batch** batch) {
    return guarded([&] {
        if (njobs < 1 || !jobs || !batch) throw std::invalid_argument("bad input buffer");
        const kbsrc_options o = opt ? *opt : defaults();
        check_options(o);
        in_nearest_mode(o.precision != KBSRC_FAST, [&] {
            const double t0 = now();
            auto b = std::make_unique<kbsrc_batch>();
            b->opt = o;
            b->jobs.resize(njobs);
            auto align = [](long long v) { return (v + 31) / 32 * 32; };
            for (int j = 0; j < njobs; ++j) {
                const kbsrc_job& J = jobs[j];
                if (J.channels < 1 || J.frames < 0) throw std::invalid_argument("bad input buffer");
                kbsrc_batch::Job& job = b->jobs[j];
                job.p = make_plan(J.frames, J.channels, J.rate_in, J.rate_out, o.precision == KBSRC_FAST, o.length,
                                  o.beta, o.block);
                const Plan& p = job.p;
                job.in_off = b->in_total;
                job.in_len = p.equal ? p.nin * p.ch : p.ch * p.xstride;
                b->in_total += align(job.in_len);
                job.out_off = b->out_total;
                job.out_len = p.ch * p.nout;
                b->out_total += align(job.out_len);
                if (p.equal) continue;
                int g = 0;   // jobs with the same rates and block size share a group
                for (; g < (int)b->groups.size(); ++g) {
                    const Plan& q = b->groups[g].p;
                    if (q.fi == p.fi && q.fo == p.fo && q.nb == p.nb) break;
                }
                if (g == (int)b->groups.size()) {
                    b->groups.emplace_back();
                    b->groups.back().p = p;
                }
                job.group = g;
                b->groups[g].jobs.push_back(j);
                b->groups[g].ntask += (long long)p.ch * p.nblk;
            }
            // shapes run on up to four streams (HIP's default of hardware queues per process is four)
            int want = 4;
            if (const char* v = std::getenv("KBSRC_BATCH_STREAMS"))   // tuning knob: streams per batch
                want = std::clamp(std::atoi(v), 1, 32);
            const int nstreams = (int)std::min<size_t>((size_t)want, std::max<size_t>(1, b->groups.size()));
            {   // groups to streams, largest first onto the least-loaded stream (work ~ tasks x n log n of both FFTs)
                std::vector<std::pair<double, int>> w;
                for (int g = 0; g < (int)b->groups.size(); ++g) {
                    const Plan& q = b->groups[g].p;
                    const double nb = q.Nb, mb = q.Mb;
                    w.push_back({(double)b->groups[g].ntask * (nb * std::log2(nb + 2) + mb * std::log2(mb + 2)), g});
                }
                std::sort(w.begin(), w.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
                std::vector<double> load(nstreams, 0.0);
                for (const auto& [wt, g] : w) {
                    const int i = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                    b->groups[g].stream = i;
                    load[i] += wt;
                }
            }
            for (int i = 0; i < nstreams; ++i) {
                hipStream_t q = nullptr;
                hip_check(hipStreamCreateWithFlags(&q, hipStreamNonBlocking), "stream");
                b->streams.push_back(q);
            }
            for (int i = 0; i <= nstreams; ++i) {
                hipEvent_t e = nullptr;
                hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
                b->events.push_back(e);
            }
            hipStream_t s = b->streams[0];
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
            const size_t per_group = std::clamp<size_t>(free_b / (16 * std::max<size_t>(1, b->groups.size())),
                                                        (size_t)256 << 20, (size_t)4 << 30);
            DeviceConsts k;
            k.init(s);
            for (kbsrc_batch::Group& G : b->groups) {
                // task table: (input offset of the block, output offset of the channel, block, channel length)
                std::vector<long long> t;
                t.reserve(4 * G.ntask);
                for (int j : G.jobs) {
                    const kbsrc_batch::Job& job = b->jobs[j];
                    const Plan& p = job.p;
                    for (int c = 0; c < p.ch; ++c)
                        for (long long blk = 0; blk < p.nblk; ++blk) {
                            t.push_back(job.in_off + c * p.xstride + blk * (long long)p.hop * p.Q);
                            t.push_back(job.out_off + c * p.nout);
                            t.push_back(blk);
                            t.push_back(p.nout);
                        }
                }
                G.tdesc.upload(t.data(), t.size() * sizeof(long long), s);
                const Design d = design_kernel(G.p, o, k, s);
                G.eng = std::make_unique<Engine>(G.p, d.H.d(), d.nH, d.Jg, s, G.ntask,
                                                 reinterpret_cast<const long long*>(G.tdesc.d()), per_group);
            }
            // one launch per step over every group: opt-in (KBSRC_BATCH_FUSE=1). Measured slower than the per-group
            // specialised launches on four streams for every batch tried (256 jobs of 0.5-2 s: 14.06 against 9.62 ms;
            // 16 such jobs: 1.01 against 0.85 ms; 32 jobs of 10-30 s: 19.17 against 13.25 ms): the generic kernel's
            // runtime stage list costs more than the launches it saves.
            {
                const char* v = std::getenv("KBSRC_BATCH_FUSE");
                b->fuse = v && v[0] == '1';
                for (kbsrc_batch::Group& G : b->groups) {   // every group must fit the generic-kernel form
                    EngineSteps probe;
                    if (b->fuse && !G.eng->describe(nullptr, nullptr, probe)) b->fuse = false;
                }
            }
            b->design_seconds = now() - t0;
            *batch = b.release();
        });
    });
}

int kbsrc_batch_query(const kbsrc_batch* b, int job, long long* v) {
    return guarded([&] {
        if (!b || !v || job < -1 || job >= (int)b->jobs.size()) throw std::invalid_argument("bad input buffer");
        if (job < 0) {
            v[0] = b->fuse; v[1] = b->in_total; v[2] = 0; v[3] = b->out_total; v[4] = (long long)b->jobs.size();
            v[5] = (long long)b->groups.size();
            return;
        }
        const kbsrc_batch::Job& J = b->jobs[job];
        v[0] = J.in_off; v[1] = J.in_len; v[2] = J.out_off; v[3] = J.out_len; v[4] = J.p.nout; v[5] = J.group;
    });
}

int kbsrc_batch_stage(kbsrc_batch* b, int job, const double* interleaved, int on_device, double* d_in, void* stream) {
    return guarded([&] {
        if (!b || !d_in || job < 0 || job >= (int)b->jobs.size()) throw std::invalid_argument("bad input buffer");
        const kbsrc_batch::Job& J = b->jobs[job];
        const Plan& p = J.p;
        if (p.nin > 0 && !interleaved) throw std::invalid_argument("bad input buffer");
        hipStream_t s = static_cast<hipStream_t>(stream);
        in_nearest_mode(b->opt.precision != KBSRC_FAST, [&] {
            if (p.equal) {   // a planar copy: channel c at in_off + c * frames
                DevBuf tmp;
                const double* src = interleaved;
                if (!on_device && p.nin > 0) {
                    tmp.upload(interleaved, sizeof(double) * (size_t)(p.nin * p.ch), s);
                    src = tmp.d();
                }
                if (p.nin > 0) {
                    StageArgs a;
                    a.in = src; a.head = src; a.tail = src; a.xx = d_in + J.in_off;
                    a.nin = p.nin; a.tot = p.nin; a.stride = p.nin; a.off = 0; a.ext = 0; a.ch = p.ch;
                    hip_check(launch_stage(a, s), "staging kernel");
                    hip_check(hipStreamSynchronize(s), "staging sync");
                }
                return;
            }
            if (on_device) stage_device(p, b->opt.lpc_edges, interleaved, d_in + J.in_off, s);
            else stage(p, b->opt.lpc_edges, interleaved, d_in + J.in_off, s);
        });
    });
}

}  // extern "C"

namespace {

// The batch's steps as one launch each over every group (nullptr when a group does not fit that form)
std::shared_ptr<kbsrc_batch::Fused> fused_steps(kbsrc_batch* b, const double* d_in, double* d_out, hipStream_t s) {
    for (const auto& f : b->fused)
        if (f->in == d_in && f->out == d_out) return f;
    std::vector<EngineSteps> all(b->groups.size());
    for (size_t g = 0; g < b->groups.size(); ++g)
        if (!b->groups[g].eng->describe(d_in, d_out, all[g])) return nullptr;
    auto F = std::make_shared<kbsrc_batch::Fused>();
    F->in = d_in;
    F->out = d_out;
    // passes at position k of every group, one launch per (policy, load, store)
    auto passes = [&](bool forward, size_t k) {
        std::vector<std::tuple<int, int, int>> keys;
        for (const EngineSteps& e : all) {
            const std::vector<PassStep>& v = forward ? e.fwd : e.inv;
            if (k < v.size()) {
                const auto key = std::make_tuple(v[k].policy, v[k].ld, v[k].st);
                if (std::find(keys.begin(), keys.end(), key) == keys.end()) keys.push_back(key);
            }
        }
        for (const auto& key : keys) {
            std::vector<PassArgs> d;
            std::vector<int> starts;
            kbsrc_batch::StepLaunch L;
            std::tie(L.policy, L.ld, L.st) = key;
            for (const EngineSteps& e : all) {
                const std::vector<PassStep>& v = forward ? e.fwd : e.inv;
                if (k >= v.size() || std::make_tuple(v[k].policy, v[k].ld, v[k].st) != key) continue;
                starts.push_back(L.total);
                d.push_back(v[k].a);
                L.total += v[k].a.ngrp * v[k].a.ntask;
                L.lds = std::max(L.lds, pass_lds_bytes(L.policy, v[k].a.sd.n, v[k].a.G));
            }
            L.n = (int)d.size();
            L.descs.upload(d.data(), d.size() * sizeof(PassArgs), s);
            L.starts.upload(starts.data(), starts.size() * sizeof(int), s);
            F->launches.push_back(std::move(L));
        }
    };
    passes(true, 0);
    passes(true, 1);
    for (int pol : {POL_DDA, POL_F64}) {
        std::vector<MidArgs> d;
        std::vector<int> starts;
        kbsrc_batch::StepLaunch L;
        L.mid = true;
        L.policy = pol;
        for (const EngineSteps& e : all) {
            if (e.mid_policy != pol) continue;
            starts.push_back(L.total);
            d.push_back(e.mid);
            L.total += mid_workgroups(e.mid);
        }
        if (d.empty()) continue;
        L.n = (int)d.size();
        L.descs.upload(d.data(), d.size() * sizeof(MidArgs), s);
        L.starts.upload(starts.data(), starts.size() * sizeof(int), s);
        F->launches.push_back(std::move(L));
    }
    passes(false, 0);
    passes(false, 1);
    hip_check(hipStreamSynchronize(s), "batch step tables");
    if (b->fused.size() >= 4) b->fused.erase(b->fused.begin());
    b->fused.push_back(F);
    return F;
}

// every group's engine on the batch's streams, forked from and joined back into s
void batch_run(kbsrc_batch* b, const double* d_in, double* d_out, hipStream_t s) {
    if (b->fuse) {
        if (const std::shared_ptr<kbsrc_batch::Fused> F = fused_steps(b, d_in, d_out, s)) {
            for (const kbsrc_batch::StepLaunch& L : F->launches) {
                const hipError_t e = L.mid
                    ? launch_mid_multi(L.policy, reinterpret_cast<const MidArgs*>(L.descs.d()),
                                       reinterpret_cast<const int*>(L.starts.d()), L.n, L.total, s)
                    : launch_pass_multi(L.policy, L.ld, L.st, reinterpret_cast<const PassArgs*>(L.descs.d()),
                                        reinterpret_cast<const int*>(L.starts.d()), L.n, L.total, L.lds, s);
                hip_check(e, "batch step");
            }
            for (const kbsrc_batch::Job& J : b->jobs)
                if (J.p.equal && J.out_len > 0)
                    hip_check(hipMemcpyAsync(d_out + J.out_off, d_in + J.in_off, sizeof(double) * J.out_len,
                                             hipMemcpyDeviceToDevice, s), "copy");
            return;
        }
    }
    const int ns = (int)b->streams.size();
    hip_check(hipEventRecord(b->events[ns], s), "batch fork");
    for (int i = 0; i < ns; ++i) hip_check(hipStreamWaitEvent(b->streams[i], b->events[ns], 0), "batch fork");
    for (size_t g = 0; g < b->groups.size(); ++g)   // each group on the stream its weight was given
        b->groups[g].eng->execute(d_in, d_out, b->streams[b->groups[g].stream % ns]);
    for (const kbsrc_batch::Job& J : b->jobs)
        if (J.p.equal && J.out_len > 0)
            hip_check(hipMemcpyAsync(d_out + J.out_off, d_in + J.in_off, sizeof(double) * J.out_len,
                                     hipMemcpyDeviceToDevice, b->streams[0]), "copy");
    for (int i = 0; i < ns; ++i) {
        hip_check(hipEventRecord(b->events[i], b->streams[i]), "batch join");
        hip_check(hipStreamWaitEvent(s, b->events[i], 0), "batch join");
    }
}

// the work `run` enqueues on a private stream, captured into an instantiated HIP graph
kbsrc_graph* capture(const std::function<void(hipStream_t)>& run) {
    hipStream_t cs = nullptr;
    hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "capture stream");
    auto g = std::make_unique<kbsrc_graph>();
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        try {
            run(cs);
        } catch (...) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(cs, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            (void)hipStreamDestroy(cs);
            throw;
        }
        e = hipStreamEndCapture(cs, &g->graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    hip_check(e, "graph capture");
    return g.release();
}

}  // namespace

extern "C" {

int kbsrc_batch_execute(kbsrc_batch* b, const double* d_in, double* d_out, void* stream) {
    return guarded([&] {
        if (!b || !d_in || !d_out) throw std::invalid_argument("bad input buffer");
        batch_run(b, d_in, d_out, static_cast<hipStream_t>(stream));
    });
}

int kbsrc_plan_capture(kbsrc_plan* pl, const double* d_padded, double* d_out, kbsrc_graph** graph) {
    return guarded([&] {
        if (!pl || !pl->eng || !d_padded || !d_out || !graph) throw std::invalid_argument("bad input buffer");
        *graph = capture([&](hipStream_t cs) { pl->eng->execute(d_padded, d_out, cs); });
    });
}

int kbsrc_batch_capture(kbsrc_batch* b, const double* d_in, double* d_out, kbsrc_graph** graph) {
    return guarded([&] {
        if (!b || !d_in || !d_out || !graph) throw std::invalid_argument("bad input buffer");
        std::shared_ptr<kbsrc_batch::Fused> F;
        if (b->fuse) F = fused_steps(b, d_in, d_out, b->streams[0]);   // tables made outside the capture
        *graph = capture([&](hipStream_t cs) { batch_run(b, d_in, d_out, cs); });
        (*graph)->keep = F;
    });
}

int kbsrc_graph_launch(kbsrc_graph* g, void* stream) {
    return guarded([&] {
        if (!g || !g->exec) throw std::invalid_argument("bad input buffer");
        hip_check(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)), "graph launch");
    });
}

void kbsrc_graph_destroy(kbsrc_graph* g) { delete g; }

void kbsrc_batch_destroy(kbsrc_batch* b) { delete b; }

int kbsrc_plan_tune(kbsrc_plan* pl, const double* d_padded, double* d_out, void* stream) {
    return guarded([&] {
        if (!pl || !pl->eng || !d_padded || !d_out) throw std::invalid_argument("bad input buffer");
        TuningScope tune;
        pl->eng->execute(d_padded, d_out, static_cast<hipStream_t>(stream));
        hip_check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "tuning sync");
    });
}

int kbsrc_batch_tune(kbsrc_batch* b, const double* d_in, double* d_out, void* stream) {
    return guarded([&] {
        if (!b || !d_in || !d_out) throw std::invalid_argument("bad input buffer");
        hipStream_t s = static_cast<hipStream_t>(stream);
        TuningScope tune;
        for (kbsrc_batch::Group& G : b->groups) G.eng->execute(d_in, d_out, s);   // one after another: clean timings
        for (const kbsrc_batch::Job& J : b->jobs)
            if (J.p.equal && J.out_len > 0)
                hip_check(hipMemcpyAsync(d_out + J.out_off, d_in + J.in_off, sizeof(double) * J.out_len,
                                         hipMemcpyDeviceToDevice, s), "copy");
        hip_check(hipStreamSynchronize(s), "tuning sync");
    });
}

int kbsrc_tuned_shapes(void) { return tuned_shapes(); }

int kbsrc_plain_dft16(const double* d_x, long long nvec, double* d_y, void* stream) {
    return guarded([&] {
        if (nvec < 0 || (nvec > 0 && (!d_x || !d_y))) throw std::invalid_argument("bad input buffer");
        // the stage tables are made in host dd arithmetic on first use
        in_nearest_mode(true, [&] {
            plain_dft16(d_x, nvec, d_y, static_cast<hipStream_t>(stream));
        });
    });
}

int kbsrc_plain_dft16_inwave(const double* d_x, long long nvec, double* d_y, void* stream) {
    return guarded([&] {
        if (nvec < 0 || (nvec > 0 && (!d_x || !d_y))) throw std::invalid_argument("bad input buffer");
        in_nearest_mode(true, [&] {
            plain_dft16_inwave(d_x, nvec, d_y, static_cast<hipStream_t>(stream));
        });
    });
}

int kbsrc_split_dft16(const double* d_x, long long nvec, int slices, double* d_y, void* stream) {
    return guarded([&] {
        if (nvec < 0 || (nvec > 0 && (!d_x || !d_y))) throw std::invalid_argument("bad input buffer");
        const int S = slices;
        if (!split_slices_ok(S)) throw std::invalid_argument("option out of range");
        in_nearest_mode(true, [&] {
            hipStream_t s = static_cast<hipStream_t>(stream);
            hip_check(launch_split_dft16(d_x, nvec, S, split_dft16_digits(S, s), d_y, s), "split radix-16");
            hip_check(hipStreamSynchronize(s), "split radix-16 sync");
        });
    });
}

int kbsrc_kernel_taps(int L, long long a, long long b, double beta, double* out) {
    return guarded([&] {
        if (L < 1 || a < 1 || b < 1 || !(beta > 0) || !out) throw std::invalid_argument("option out of range");
        in_nearest_mode(true, [&] {
            const DevBuf h = kernel_taps_gpu(L, a, b, beta, nullptr);
            hip_check(hipMemcpy(out, h.d(), sizeof(double) * 2 * ((size_t)L + 1), hipMemcpyDeviceToHost),
                      "taps download");
        });
    });
}

int kbsrc_kernel_spectrum(const double* taps, int Lk, int n, double scale, double* out) {
    return guarded([&] {
        if (!taps || !out || Lk < 0 || n < 2 || (n & 1) || Lk >= n || !(scale != 0))
            throw std::invalid_argument("option out of range");
        in_nearest_mode(true, [&] {
            DevBuf h;
            h.upload(taps, sizeof(double) * 2 * ((size_t)Lk + 1), nullptr);
            DeviceConsts k;
            k.init(nullptr);
            const DevBuf H = kernel_spectrum_gpu(h.d(), Lk, n, scale, k, nullptr);
            hip_check(hipMemcpy(out, H.d(), sizeof(double) * 2 * ((size_t)n / 2 + 1), hipMemcpyDeviceToHost),
                      "spectrum download");
        });
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
