// GPU engine: FFT plans with device-resident twiddle tables, the kernel spectrum, and the overlap-save block
// engine over all (channel, block) tasks. Host-side orchestration only; the kernels are in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

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

// Device allocation owned by RAII.
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
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
};

// Element format of a policy's device tables: dd = {re.hi, im.hi, re.lo, im.lo}, FP64 = {re, im}.
std::vector<double> pack_cx(const std::vector<dd>& re, const std::vector<dd>& im, bool as_dd);

// One Stockham sub-FFT length with its stage list (radix order 4s, 2, 3s, 5s, 7s: src/dsp/fft.hpp:91-107).
struct SubFFT {
    SubDesc desc;
    DevBuf tw;   // stage twiddles in the plan's element format
    void init(int n, bool as_dd, hipStream_t s);
};

// Complex FFT of length n as the reference computes it (src/dsp/fft.hpp:226-265): four-step N1 x N2 above
// 2048 points (N1 the largest divisor <= sqrt n, at least 8), otherwise one Stockham.
struct CFFTPlan {
    int n = 0, N1 = 0, N2 = 0;
    bool four = false, as_dd = true;
    SubFFT s1, s2, sd;
    DevBuf tw4;   // four-step twiddles w_n^{-k1 n2} at k1*N2 + n2
    void init(int n, bool as_dd, hipStream_t s);
};

// Real FFT of even length N through a complex FFT of N/2 (src/dsp/fft.hpp:352-358).
struct RFFTPlan {
    int N = 0, h = 0;
    CFFTPlan c;
    DevBuf w;   // w[k] = e^{-2 pi i k / N}, k = 0..h
    void init(int N, bool as_dd, hipStream_t s);
};

struct PassTimes {   // milliseconds per kernel kind, summed over one execute()
    double fwd_col = 0, fwd_row = 0, mid = 0, inv_col = 0, inv_row = 0;
    int launches = 0;
};

// Shared device constants (odd-radix butterfly constants in both element formats).
struct DeviceConsts {
    DevBuf odd_dd, odd_f64;
    void init(hipStream_t s);
};

// The kernel spectrum H (src/design.cpp:40-49) on the GPU: the circularly placed dd kernel goes through the
// same sub-FFT passes in the reference's scalar-dd form, then a split that keeps Re X / scale.
std::vector<dd> kernel_spectrum_gpu(const std::vector<dd>& h, int n, double scale, const DeviceConsts& k,
                                    hipStream_t s);

// The overlap-save engine for one plan (src/resampler.cpp:34-88).
class Engine {
public:
    Engine(const Plan& p, const std::vector<dd>& H, hipStream_t s);
    // xx: device, ch x plan.tot padded planar input. out: device, ch x plan.nout. times: optional per-pass timing.
    void execute(const double* xx, double* out, hipStream_t s, PassTimes* times = nullptr);
    long long tasks() const { return ntask_; }
    int batch() const { return tb_; }
    const Plan& plan() const { return p_; }

private:
    void run_cfft(const CFFTPlan& f, int ld, const double* src, int st, double* mid, double* dst, long long task0,
                  int nt, const double* xx, double* out, hipStream_t s, double* t_col, double* t_row,
                  PassTimes* times);
    Plan p_;
    bool dd_;
    long long ntask_ = 0, ts_ = 0;   // ts_: doubles per task in the work buffers
    int tb_ = 1;
    RFFTPlan fw_, iv_;
    DeviceConsts k_;
    DevBuf H_, w1_, w2_;
};

}  // namespace kbsrc::gpu
