// GPU engine orchestration: FFT plans, device tables, and the pass sequence of the overlap-save engine.
#include "engine.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

#pragma clang fp contract(off)

namespace kbsrc::gpu {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p_, bytes);
    if (e != hipSuccess) {
        p_ = nullptr;
        (void)hipGetLastError();
        throw std::bad_alloc();
    }
    n_ = bytes;
}
void DevBuf::release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
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

void CFFTPlan::init(int n_, bool as_dd_, hipStream_t s) {
    n = n_;
    as_dd = as_dd_;
    if (!smooth7(n)) throw std::runtime_error("transform lengths with prime factors above 7 are not supported yet");
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
    CisTable ct(n);
    std::vector<dd> re((size_t)n), im((size_t)n);
    for (int k1 = 0; k1 < N1; ++k1)
        for (int n2 = 0; n2 < N2; ++n2) ct.get(-(long long)k1 * n2, re[(size_t)k1 * N2 + n2], im[(size_t)k1 * N2 + n2]);
    const std::vector<double> v = pack_cx(re, im, as_dd);
    tw4.upload(v.data(), v.size() * sizeof(double), s);
}

void RFFTPlan::init(int N_, bool as_dd, hipStream_t s) {
    N = N_;
    h = N / 2;
    if (N % 2) throw std::runtime_error("RFFT length must be even");
    c.init(h, as_dd, s);
    CisTable ct(N);
    std::vector<dd> re(h + 1), im(h + 1);
    for (int k = 0; k <= h; ++k) ct.get(-k, re[k], im[k]);
    const std::vector<double> v = pack_cx(re, im, as_dd);
    w.upload(v.data(), v.size() * sizeof(double), s);
}

void DeviceConsts::init(hipStream_t s) {
    std::vector<dd> re(21, mkdd(0.0)), im(21, mkdd(0.0));
    const int rr[3] = {3, 5, 7};
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < rr[i]; ++k) cis(k, rr[i], re[i * 7 + k], im[i * 7 + k]);
    const std::vector<double> a = pack_cx(re, im, true), b = pack_cx(re, im, false);
    odd_dd.upload(a.data(), a.size() * sizeof(double), s);
    odd_f64.upload(b.data(), b.size() * sizeof(double), s);
}

namespace {

constexpr int kThreads = 256;
constexpr size_t kLdsBudget = 48 * 1024;

int pick_group(int policy, int n, int nsub) {
    const size_t one = pass_lds_bytes(policy, n, 1);
    if (one > 160 * 1024) throw std::runtime_error("sub-FFT too long for the LDS kernel");
    int G = (int)std::max<size_t>(1, kLdsBudget / one);
    return std::max(1, std::min({G, nsub, 16}));
}

// launch one sub-FFT pass; errors name the pass and its shape
// Test and tuning switches, read at every execute: KBSRC_GENERIC_PASSES=1 runs every sub-FFT pass through the
// generic kernel; KBSRC_FUSED=1 folds the split/multiply/fold kernel into the passes next to it (same bits; it
// saves a kernel and ~1.2 GB of HBM traffic per benchmark step but measured no faster, the passes being
// FP64-bound).
bool env_flag(const char* name) {
    const char* v = std::getenv(name);
    return v && v[0] == '1';
}
bool use_fixed() { return !env_flag("KBSRC_GENERIC_PASSES"); }

void run_pass(int policy, int ld, int st, PassArgs a, hipStream_t s, const char* what) {
    hipError_t e = hipSuccess;
    if (!(use_fixed() && launch_fixed(policy, ld, st, a, s, &e))) e = launch_pass(policy, ld, st, a, kThreads, s);
    if (e != hipSuccess)
        throw HipError(std::string(what) + " (n=" + std::to_string(a.sd.n) + ", G=" + std::to_string(a.G) +
                       ", LDS " + std::to_string(pass_lds_bytes(policy, a.sd.n, a.G)) + " B): " + hipGetErrorString(e));
}

void fill_sub(PassArgs& a, const SubFFT& sf, const DeviceConsts& k, bool as_dd, int policy, int nsub) {
    a.sd = sf.desc;
    a.tw = sf.tw.d();
    a.oddc = as_dd ? k.odd_dd.d() : k.odd_f64.d();
    a.nsub = nsub;
    a.G = pick_group(policy, sf.desc.n, nsub);
    a.ngrp = (nsub + a.G - 1) / a.G;
}

}  // namespace

std::vector<dd> kernel_spectrum_gpu(const std::vector<dd>& h, int n, double scale, const DeviceConsts& k,
                                    hipStream_t s) {
    RFFTPlan R;
    R.init(n, true, s);
    const std::vector<dd> r = circular_kernel(h, n);
    std::vector<double> rv(2 * (size_t)n);
    for (int i = 0; i < n; ++i) { rv[2 * i] = r[i].hi; rv[2 * i + 1] = r[i].lo; }
    DevBuf din, w1((size_t)R.h * 32), w2((size_t)R.h * 32), dH((size_t)(R.h + 1) * 16);
    din.upload(rv.data(), rv.size() * sizeof(double), s);
    const CFFTPlan& c = R.c;
    if (c.four) {   // scalar-dd four-step: accurate adds, no twiddle on column 0, values stored as computed
        PassArgs a;
        fill_sub(a, c.s1, k, true, POL_DDA, c.N2);
        a.src = din.d(); a.lde = c.N2; a.ldg = 1;
        a.dst = w1.d(); a.ste = c.N2; a.stg = 1; a.tw4 = c.tw4.d(); a.twmode = 2;
        run_pass(POL_DDA, LD_REAL_DD, ST_CPLX, a, s, "spectrum column pass");
        PassArgs b;
        fill_sub(b, c.s2, k, true, POL_DDA, c.N1);
        b.src = w1.d(); b.lde = 1; b.ldg = c.N2;
        b.dst = w2.d(); b.ste = c.N1; b.stg = 1;
        run_pass(POL_DDA, LD_CPLX, ST_CPLX, b, s, "spectrum row pass");
    } else {
        PassArgs a;
        fill_sub(a, c.sd, k, true, POL_DDA, 1);
        a.src = din.d(); a.lde = 1; a.ldg = 0;
        a.dst = w2.d(); a.ste = 1; a.stg = 0;
        run_pass(POL_DDA, LD_REAL_DD, ST_CPLX, a, s, "spectrum pass");
    }
    SpecArgs sa;
    sa.z = w2.d(); sa.w = R.w.d(); sa.H = dH.d(); sa.h = R.h; sa.scale = scale;
    hip_check(launch_spec_split(sa, s), "spectrum split");
    std::vector<double> hv(2 * (size_t)(R.h + 1));
    hip_check(hipMemcpyAsync(hv.data(), dH.d(), hv.size() * sizeof(double), hipMemcpyDeviceToHost, s), "spectrum copy");
    hip_check(hipStreamSynchronize(s), "spectrum sync");
    std::vector<dd> H(R.h + 1);
    for (int i = 0; i <= R.h; ++i) H[i] = dd{hv[2 * i], hv[2 * i + 1]};
    return H;
}

Engine::Engine(const Plan& p, const std::vector<dd>& H, hipStream_t s) : p_(p), dd_(!p.fast) {
    if (p.gen) throw std::runtime_error("near-unity rate ratios (general fold) are not supported yet");
    ntask_ = (long long)p.ch * p.nblk;
    k_.init(s);
    fw_.init(p.Nb, dd_, s);
    iv_.init(p.Mb, dd_, s);
    // H in the element format of the engine: (hi, lo) pairs or FP64
    std::vector<double> hv(H.size() * (dd_ ? 2 : 1));
    for (size_t i = 0; i < H.size(); ++i) {
        if (dd_) { hv[2 * i] = H[i].hi; hv[2 * i + 1] = H[i].lo; }
        else hv[i] = H[i].hi;
    }
    H_.upload(hv.data(), hv.size() * sizeof(double), s);
    const size_t elt = dd_ ? 32 : 16;
    const size_t per_task = (size_t)(std::max(fw_.h, iv_.h) + 1) * elt;   // + 1: the split's X[h]
    ts_ = (long long)(per_task / sizeof(double));
    const size_t budget = (size_t)1 << 30;
    tb_ = (int)std::max<long long>(1, std::min<long long>(ntask_, (long long)(budget / per_task)));
    if (const char* v = std::getenv("KBSRC_BATCH_TASKS")) {   // tuning knob: tasks per launch
        const long long want = std::atoll(v);
        if (want > 0) tb_ = (int)std::min<long long>(tb_, want);
    }
    w1_.alloc(per_task * tb_);
    w2_.alloc(per_task * tb_);
}

void Engine::run_cfft(const CFFTPlan& f, int ld, const double* src, int st, double* mid, double* dst,
                      long long task0, int nt, const double* xx, double* out, hipStream_t s, double* t_col,
                      double* t_row, PassTimes* times) {
    const int pol_four = dd_ ? POL_DDS : POL_F64, pol_one = dd_ ? POL_DDA : POL_F64;
    const long long mid_ts = ts_, src_tstride = ts_, dst_tstride = ts_;
    auto common = [&](PassArgs& a) {
        a.ntask = nt;
        a.task0 = task0;
        a.nblk = p_.nblk;
        a.in_cstride = p_.xstride;
        a.in_bstride = (long long)p_.hop * p_.Q;
        a.src_tstride = src_tstride;
        a.out = out;
        a.nout = p_.nout;
        a.o = p_.o;
        a.margP = (long long)p_.marg * p_.P;
        a.hopP = (long long)p_.hop * p_.P;
        FuseArgs& fa = a.f;
        fa.H = H_.d(); fa.wf = fw_.w.d(); fa.wi = iv_.w.d();
        fa.h = fw_.h; fa.hinv = iv_.h; fa.Nb = p_.Nb; fa.Mb = p_.Mb; fa.K = p_.Mb / 2 + 1;
        fa.jmax = std::min(p_.Nb / 2, p_.Mb - 1); fa.up = p_.up; fa.xmax = p_.up ? fw_.h : fa.jmax;
    };
    auto timed = [&](double* acc, auto&& fn) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (times) {
            hip_check(hipEventCreate(&e0), "event");
            hip_check(hipEventCreate(&e1), "event");
            hip_check(hipEventRecord(e0, s), "event record");
        }
        fn();
        if (times) {
            hip_check(hipEventRecord(e1, s), "event record");
            hip_check(hipEventSynchronize(e1), "event sync");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
            *acc += ms;
            times->launches++;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
    };
    const double* in_src = ld == LD_REAL_F64 ? xx : src;
    if (f.four) {
        PassArgs a;
        fill_sub(a, f.s1, k_, dd_, pol_four, f.N2);
        common(a);
        a.src = in_src; a.lde = f.N2; a.ldg = 1;
        a.dst = mid; a.dst_tstride = mid_ts; a.ste = f.N2; a.stg = 1;
        a.tw4 = f.tw4.d(); a.twmode = 1; a.renorm = dd_;
        a.f.M = f.N2;
        timed(t_col, [&] { run_pass(pol_four, ld, ST_CPLX, a, s, "column pass"); });
        PassArgs b;
        fill_sub(b, f.s2, k_, dd_, pol_four, f.N1);
        common(b);
        b.src = mid; b.src_tstride = mid_ts; b.lde = 1; b.ldg = f.N2;
        b.dst = dst; b.dst_tstride = dst_tstride; b.ste = f.N1; b.stg = 1; b.renorm = dd_;
        b.f.M = f.N1;
        timed(t_row, [&] { run_pass(pol_four, LD_CPLX, st, b, s, "row pass"); });
    } else {
        PassArgs a;
        fill_sub(a, f.sd, k_, dd_, pol_one, 1);
        common(a);
        a.src = in_src; a.lde = 1; a.ldg = 0;
        a.dst = dst; a.dst_tstride = dst_tstride; a.ste = 1; a.stg = 0;
        timed(t_row, [&] { run_pass(pol_one, ld, st, a, s, "single pass"); });
    }
}

void Engine::execute(const double* xx, double* out, hipStream_t s, PassTimes* times) {
    const long long ts = ts_;
    PassTimes local;
    PassTimes* tp = times ? &local : nullptr;
    const int pol_four = dd_ ? POL_DDS : POL_F64;
    const bool fused = use_fixed() && env_flag("KBSRC_FUSED") && fw_.c.four && iv_.c.four &&
                       has_fixed(pol_four, fw_.c.N2, LD_CPLX, ST_SPLIT) && has_fixed(pol_four, iv_.c.N1, LD_FOLD, ST_CPLX);
    for (long long t0 = 0; t0 < ntask_; t0 += tb_) {
        const int nt = (int)std::min<long long>(tb_, ntask_ - t0);
        if (fused) {
            // padded input -> (w1) -> X = split of z (w2) -> bin map and pre-split folded into the inverse columns
            run_cfft(fw_.c, LD_REAL_F64, nullptr, ST_SPLIT, w1_.d(), w2_.d(), t0, nt, xx, out, s, &local.fwd_col,
                     &local.fwd_row, tp);
            run_cfft(iv_.c, LD_FOLD, w2_.d(), ST_OUT, w1_.d(), nullptr, t0, nt, xx, out, s, &local.inv_col,
                     &local.inv_row, tp);
            continue;
        }
        // forward real FFT: padded input -> z (w2)
        run_cfft(fw_.c, LD_REAL_F64, nullptr, ST_CPLX, w1_.d(), w2_.d(), t0, nt, xx, out, s, &local.fwd_col,
                 &local.fwd_row, tp);
        // split, x H, bin map, inverse pre-split: z (w2) -> z' (w1)
        MidArgs m;
        m.z = w2_.d(); m.z_tstride = ts; m.zi = w1_.d(); m.zi_tstride = ts;
        m.H = H_.d(); m.wf = fw_.w.d(); m.wi = iv_.w.d();
        m.h = fw_.h; m.hinv = iv_.h; m.Nb = p_.Nb; m.Mb = p_.Mb; m.K = p_.Mb / 2 + 1;
        m.jmax = std::min(p_.Nb / 2, p_.Mb - 1); m.up = p_.up; m.ntask = nt;
        {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (tp) {
                hip_check(hipEventCreate(&e0), "event"); hip_check(hipEventCreate(&e1), "event");
                hip_check(hipEventRecord(e0, s), "event record");
            }
            hip_check(launch_mid(dd_ ? POL_DDA : POL_F64, m, s), "mid kernel");
            if (tp) {
                hip_check(hipEventRecord(e1, s), "event record");
                hip_check(hipEventSynchronize(e1), "event sync");
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
                local.mid += ms; local.launches++;
                (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
            }
        }
        // inverse: z' (w1) -> (w2) -> output samples
        run_cfft(iv_.c, LD_CPLX, w1_.d(), ST_OUT, w2_.d(), nullptr, t0, nt, xx, out, s, &local.inv_col,
                 &local.inv_row, tp);
    }
    if (times) *times = local;
}

}  // namespace kbsrc::gpu
