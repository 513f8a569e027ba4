// GPU engine: FFT plans with device-resident twiddle tables, the kernel spectrum, and the overlap-save block
// engine over all (channel, block) tasks. Host-side orchestration only; the kernels are in kernels*.hip.
#pragma once
#include <functional>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_design.hpp"
#include "kernels.hpp"

namespace kbsrc::gpu {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
void hip_check(hipError_t e, const char* what);

// Stream-ordered allocation scope of the calling thread. While one is active, a DevBuf released on this thread goes
// to the library's block cache instead of hipFree, reusable once the work queued on the scope's stream so far is
// done, and a DevBuf allocated takes a cached block when one fits (its stream waiting for that block's last work)
// before it calls hipMalloc. A release then neither waits for the device nor pays hipFree: for a conversion that
// allocates and frees while its own copies and kernels run, and for the next conversion, which finds its buffers
// cached. The owner of a scope keeps every use of such memory ordered before the release (on the scope's stream,
// or synchronised before it). Scopes nest; the innermost one is current.
class PoolScope {
public:
    explicit PoolScope(hipStream_t s);
    ~PoolScope();
    PoolScope(const PoolScope&) = delete;
    PoolScope& operator=(const PoolScope&) = delete;
    static const PoolScope* current();
    hipStream_t stream() const { return s_; }

private:
    hipStream_t s_;
    const PoolScope* prev_;
};

// Device allocation owned by RAII.
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), pooled_(o.pooled_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_; pooled_ = o.pooled_;
            o.p_ = nullptr; o.n_ = 0;
        }
        return *this;
    }
    void alloc(size_t bytes);
    void release();
    double* d() const { return static_cast<double*>(p_); }
    size_t bytes() const { return n_; }
    void upload(const void* host, size_t bytes, hipStream_t s);

private:
    void* p_ = nullptr;
    size_t n_ = 0;
    bool pooled_ = false;   // allocated in a PoolScope: released to the block cache
};

// Element format of a policy's device tables: dd = {re.hi, im.hi, re.lo, im.lo}, FP64 = {re, im}.
std::vector<double> pack_cx(const std::vector<dd>& re, const std::vector<dd>& im, bool as_dd);

// Shared device constants (odd-radix butterfly constants in both element formats).
struct DeviceConsts {
    DevBuf odd_dd, odd_f64;
    void init(hipStream_t s);
};

// One Stockham sub-FFT length with its stage list (radix order 4s, 2, 3s, 5s, 7s: src/dsp/fft.hpp:91-107).
struct SubFFT {
    SubDesc desc;
    DevBuf tw;   // stage twiddles in the plan's element format
    DevBuf tw16;   // split form: the radix-16 stage's twiddles (dd elements), built by init_split
    void init(int n, bool as_dd, hipStream_t s);
    // the split form's radix-16 stage twiddles w_n^{p k} at p * 15 + k - 1 (p < n / 16, k = 1..15), for lengths
    // whose first two stages are radix 4; host dd arithmetic, round to nearest
    void init_split(hipStream_t s);
};

// Complex FFT of length n as the reference computes it (src/dsp/fft.hpp:226-265): four-step N1 x N2 above
// 2048 points (N1 the largest divisor <= sqrt n, at least 8), otherwise one Stockham; lengths with a prime factor
// above 7 by Bluestein around an inner 7-smooth FFT of length m = next 7-smooth >= 2n - 1.
struct CFFTPlan {
    int n = 0, N1 = 0, N2 = 0;
    bool four = false, as_dd = true;
    SubFFT s1, s2, sd;
    DevBuf tw4;   // four-step twiddles w_n^{-k1 n2} at k1*N2 + n2
    bool blue = false;
    int m = 0;
    std::unique_ptr<CFFTPlan> inner;
    DevBuf bc, vf;   // chirp (n) and the inner FFT of the chirp sequence divided by m (m)
    void init(int n, bool as_dd, const DeviceConsts& k, hipStream_t s);
    void init_split(hipStream_t s);   // SubFFT::init_split on every sub-FFT length of the plan
    long long work_elems() const { return blue ? m : n; }   // elements per task a run needs in each buffer
    bool needs_global() const;   // some sub-FFT too long for an LDS tile: passes go through global scratch
};

// Real FFT of even length N through a complex FFT of N/2 (src/dsp/fft.hpp:352-358).
struct RFFTPlan {
    int N = 0, h = 0;
    CFFTPlan c;
    DevBuf w;   // w[k] = e^{-2 pi i k / N}, k = 0..h
    void init(int N, bool as_dd, const DeviceConsts& k, hipStream_t s);
};

// Milliseconds per kernel kind, summed over one execute(). Each launch is bracketed by two HIP events on its stream
// and the launches are queued back to back (no host wait between them), so an interval is the kernel's own time
// plus the queue's few-microsecond dispatch gap; the events are read once the execute has been queued.
struct PassTimes {
    double fwd_col = 0, fwd_row = 0, mid = 0, inv_col = 0, inv_row = 0;
    int launches = 0;
    struct Pending {
        hipEvent_t e0, e1;
        double* acc;
    };
    std::vector<Pending> pending;
    void collect();   // waits for the pending events and adds their intervals (destroys the events)
    PassTimes() = default;
    PassTimes(const PassTimes&) = delete;
    PassTimes& operator=(const PassTimes&) = delete;
    ~PassTimes();     // events still pending (an execute that threw) are destroyed unread
};

// How a complex FFT runs: lane = the reference's engine form (sub-FFTs on its 4-lane types: one-TwoSum dd adds,
// renormalised stores, four-step twiddle on every k1 != 0), otherwise its scalar form (accurate dd adds, no twiddle
// on column 0), which the kernel design and the Bluestein chirp spectrum use.
// A sub-FFT pass described for the generic kernel (G and ngrp set for 256 threads) instead of launched.
struct PassStep {
    int policy, ld, st;
    PassArgs a;
};
struct RunCtx {
    PassArgs proto;        // tasks, block addressing, output window, fused-pass arguments
    long long ts = 0;      // doubles per task in every work buffer
    const DeviceConsts* k = nullptr;
    bool dd = true, lane = true;
    hipStream_t s = nullptr;
    PassTimes* times = nullptr;
    std::vector<PassStep>* collect = nullptr;   // set: passes are described here, not launched
    bool* collect_ok = nullptr;                 // cleared when a pass cannot run on the generic kernel
    // set: a four-step transform's row pass is handed to it (its arguments complete) instead of launched
    const std::function<void(PassArgs&)>* rows = nullptr;
};
// An engine's whole execute as generic-kernel steps: forward passes, the split/multiply/fold, inverse passes.
struct EngineSteps {
    std::vector<PassStep> fwd, inv;
    MidArgs mid;
    int mid_policy = 0;
};

// Runs f over proto.ntask tasks: load kind ld from src, store kind st to dst (work buffers, stride ctx.ts).
// scr: three work buffers the run may use (the four-step's middle buffer; Bluestein's).
void run_cfft(const CFFTPlan& f, const RunCtx& c, int ld, const double* src, int st, double* dst,
              double* const scr[3], double* t_col, double* t_row);

// While a TuningScope lives on this thread, every sub-FFT pass whose shape (policy, length, load, store, sub-FFTs)
// has not been measured in this process times its candidate kernel shapes and keeps the fastest for all later runs.
struct TuningScope {
    TuningScope();
    ~TuningScope();
};
int tuned_shapes();   // pass shapes measured so far in this process

// nvec 16-point complex DFTs of dd vectors (plan element format, device) by the engine's dd Stockham pass: the
// plain baseline of the split-precision form. Synchronous.
void plain_dft16(const double* x, long long nvec, double* y, hipStream_t s);
// the same stage with its exchange inside the wave instead of LDS (same bits). Synchronous.
void plain_dft16_inwave(const double* x, long long nvec, double* y, hipStream_t s);
// the split matrix slices (split_matrix_digits) on the current device, made on first use and kept for the process
const long* split_dft16_digits(int S, hipStream_t s);

// The Kaiser-sinc taps h[0..L] (src/design.cpp:17-37) on the GPU, as (hi, lo) pairs in device memory.
DevBuf kernel_taps_gpu(long long L, long long a, long long b, double beta, hipStream_t s);

// The kernel spectrum H (src/design.cpp:40-49) on the GPU: device taps h[0..Lk] (pairs) placed circularly in
// length n, through the passes in the reference's scalar-dd form, then a split that keeps Re X / scale. Returns
// the n/2 + 1 values of H as (hi, lo) pairs in device memory.
DevBuf kernel_spectrum_gpu(const double* taps, long long Lk, int n, double scale, const DeviceConsts& k,
                           hipStream_t s);

// The overlap-save engine for one plan (src/resampler.cpp:34-88). H: device, nH (hi, lo) pairs. Jg > 0 selects
// the general fold.
class Engine {
public:
    // ntask > 0 with tdesc (device, 4 values per task, see task_io): the engine runs that many tasks of several
    // jobs that share p's shape (a batch) instead of p's own channels x blocks. buffer_budget > 0: bytes per work
    // buffer (tasks per launch follow from it).
    Engine(const Plan& p, const double* H, long long nH, int Jg, hipStream_t s, long long ntask = 0,
           const long long* tdesc = nullptr, size_t buffer_budget = 0);
    // xx: device, ch x plan.xstride padded planar input. out: device, ch x plan.nout. times: optional per-pass timing.
    // split > 0: the split-precision form (exact-mode plans only; not bit-exact): sub-FFTs whose first two stages
    // are radix 4 run them as one int8-MFMA radix-16 stage with `split` base-128 slices per operand.
    // out_dd (exact-mode plans): each output sample as its dd value before the final rounding, a (hi, lo) pair;
    // out holds ch x 2 nout doubles.
    void execute(const double* xx, double* out, hipStream_t s, PassTimes* times = nullptr, int split = 0,
                 bool out_dd = false);
    // The plain execute of tasks [t0, t1) only (task t: channel t / nblk, block t % nblk), which write the output
    // samples of their blocks; the other tasks' output is left as it is.
    void execute_tasks(const double* xx, double* out, hipStream_t s, long long t0, long long t1);
    // The plain execute with every work buffer carved from the caller's device memory [mem, mem + bytes) instead of
    // the engine's own: tasks run in as many launches as that memory needs, with the same bits. Allocates nothing.
    // Throws std::invalid_argument when one task's buffers do not fit.
    void execute_in(const double* xx, double* out, hipStream_t s, double* mem, size_t bytes);
    size_t task_bytes() const;   // bytes one task needs over all its work buffers
    // The same execute described as generic-kernel steps on these buffers (for one launch per step over several
    // engines); false when it does not fit that form (Bluestein lengths, sub-FFTs longer than an LDS tile, more
    // tasks than one batch of the work buffers).
    bool describe(const double* xx, double* out, EngineSteps& steps);
    long long tasks() const { return ntask_; }
    int batch() const { return tb_; }
    const Plan& plan() const { return p_; }

private:
    struct Work {
        double* w[4];   // the work buffers (two, four with Bluestein)
        double* g[2];   // global-pass scratch (sub-FFTs longer than an LDS tile), or null
        int tb;         // tasks per launch
    };
    void run(const double* xx, double* out, hipStream_t s, PassTimes* times, int split, const Work& w,
             long long t0 = 0, long long t1 = -1, bool out_dd = false);
    RunCtx context(const double* xx, double* out, hipStream_t s, int split, const Work& w);
    Plan p_;
    bool dd_;
    int Jg_ = 0;
    long long ntask_ = 0, ts_ = 0;   // ts_: doubles per task in the work buffers
    long long zkeep_ = -1;           // forward spectrum bins above this (and their split partners) go unstored
    int kchain_ = 0;                 // MidArgs::kchain
    int tgroup_ = 0;                 // MidArgs::tgroup
    const long long* tdesc_ = nullptr;
    int tb_ = 1;
    DeviceConsts k_;
    RFFTPlan fw_, iv_;
    DevBuf H_, w_[4], g_[2];
    DevBuf Hr_, wfr_, wir_;   // residue-major H, wf and wi for the fused row pass (MidArgs::Hr, wfr, wir)
    int Lr_ = 0;
    std::map<int, DevBuf> adig_;   // split form: the DFT matrix's slices per slice count
};

// Bytes of the device array kept for executes that must not allocate (scratch.hip), and its address on the
// current device (taken once per device from the symbol; the array lives until the process ends).
constexpr size_t kScratchBytes = (size_t)128 << 20;
double* scratch_memory();

// The split form's 16-point DFT matrix as S base-128 slices in the MFMA A layout: M / 2, M the real form of
// F = e^{-2 pi i k n / 16} with rows (Re y[0..15], Im y[0..15]) and columns (Re x[0], Im x[0], Re x[1], ...),
// S - 1 balanced low digits and the leading remainder, most significant first; 64-bit word (slice j, M-tile mt,
// lane l) holds bytes b = slice j of M[16 mt + l % 16][8 (l / 16) + b]. Round to nearest.
std::vector<uint64_t> split_matrix_digits(int S);

}  // namespace kbsrc::gpu
