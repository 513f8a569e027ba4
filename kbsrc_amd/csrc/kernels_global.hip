// Sub-FFT passes whose length does not fit LDS (e.g. the 3200-point rows of a 10 MHz -> 48 kHz plan): gathered
// into global scratch (sub-FFT j of task t at scratch[(t * nsub + j) * n .. + n)), transformed there, scattered
// to the pass's destination. The transform runs as two LDS passes of the generic kernel when the stage list splits
// into two parts that each fit a tile, else one launch per Stockham stage. Arithmetic identical to the other passes.
#include <algorithm>
#include <cstdlib>

#include "bfly.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {

namespace {

template <class A, int LD> __global__ __launch_bounds__(256) void k_gather(PassArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = a.sd.n, total = (long long)a.ntask * a.nsub * n;
    if (idx >= total) return;
    const long long e = idx % n, r = idx / n, j = r % a.nsub, task = r / a.nsub;
    const long long i = e * a.lde + j * a.ldg;
    Cx<A> v;
    if constexpr (LD == LD_REAL_F64) {
        const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
        const double* s = a.src + io.in_off;
        v = {A::lift(s[2 * i]), A::lift(s[2 * i + 1])};
    } else if constexpr (LD == LD_REAL_DD) {
        const double* s = a.src + task * a.src_tstride;   // real dd {hi, lo} pairs (FP64: plain reals)
        if constexpr (A::kDD) {
            const double2 x = reinterpret_cast<const double2*>(s)[2 * i], y = reinterpret_cast<const double2*>(s)[2 * i + 1];
            v = {dd{x.x, x.y}, dd{y.x, y.y}};
        } else {
            v = {s[2 * i], s[2 * i + 1]};
        }
    } else {
        v = gld<A>(a.src + task * a.src_tstride, i);
    }
    gst<A>(a.gscr0, idx, v);
}

template <class A, int R>
__device__ __forceinline__ void gstage_r(const PassArgs& a, const StageDesc& st, const double* in, double* out,
                                         long long base, int b) {
    const int n = a.sd.n, sm = n / R, s = st.s, p = b / s, q = b - p * s, dbase = s * R * p + q;
    Cx<A> x[R], y[R];
#pragma unroll
    for (int t = 0; t < R; ++t) x[t] = gld<A>(in, base + b + t * sm);
    butterfly<A, R>(x, y, p != 0, a.tw + (long long)(st.tw + p * (R - 1)) * kElt<A>, a.oddc);
#pragma unroll
    for (int u = 0; u < R; ++u) gst<A>(out, base + dbase + u * s, y[u]);
}

template <class A> __global__ __launch_bounds__(256) void k_gstage(PassArgs a, int si, const double* in, double* out) {
    const StageDesc st = a.sd.st[si];
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long sm = a.sd.n / st.r, total = (long long)a.ntask * a.nsub * sm;
    if (idx >= total) return;
    const int b = (int)(idx % sm);
    const long long base = (idx / sm) * a.sd.n;
    switch (st.r) {
        case 4: gstage_r<A, 4>(a, st, in, out, base, b); break;
        case 2: gstage_r<A, 2>(a, st, in, out, base, b); break;
        case 3: gstage_r<A, 3>(a, st, in, out, base, b); break;
        case 5: gstage_r<A, 5>(a, st, in, out, base, b); break;
        default: gstage_r<A, 7>(a, st, in, out, base, b); break;
    }
}

template <class A, int ST> __global__ __launch_bounds__(256) void k_scatter(PassArgs a, const double* buf) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = a.sd.n, total = (long long)a.ntask * a.nsub * n;
    if (idx >= total) return;
    const long long e = idx % n, r = idx / n, j = r % a.nsub, task = r / a.nsub;
    const long long i = e * a.ste + j * a.stg;
    if (ST == ST_CPLX && z_unused(a, i)) return;
    Cx<A> v = gld<A>(buf, idx);
    if (a.twmode && e != 0 && (a.twmode == 1 || j != 0)) v = c_mul<A>(v, gld<A>(a.tw4, i));
    if constexpr (A::kDD) {
        if (a.renorm) v = {quick2(v.re.hi, v.re.lo), quick2(v.im.hi, v.im.lo)};
    }
    if constexpr (ST == ST_CPLX) {
        gst<A>(a.dst + task * a.dst_tstride, i, v);
    } else {
        const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
        // the value rounded to double, or (out_dd) the dd value as its (hi, lo) pair
        auto put = [&](long long k, Val<A> y) {
            if constexpr (A::kDD) {
                if (a.out_dd) {
                    a.out[2 * k] = y.hi;
                    a.out[2 * k + 1] = y.lo;
                } else {
                    a.out[k] = y.hi + y.lo;
                }
            } else {
                a.out[k] = y;
            }
        };
        const long long x0 = 2 * i - a.margP, x1 = x0 + 1;
        if (x0 >= 0 && x0 < a.hopP) {
            const long long d = io.blk * a.hopP + x0 - a.o;
            if (d >= 0 && d < io.nout) put(io.out_off + d, v.re);
        }
        if (x1 >= 0 && x1 < a.hopP) {
            const long long d = io.blk * a.hopP + x1 - a.o;
            if (d >= 0 && d < io.nout) put(io.out_off + d, A::neg(v.im));
        }
    }
}

constexpr size_t kTile = 160 * 1024;   // LDS of one workgroup

// Split point of a long sub-FFT: its first k stages, on the n / R columns c + (n / R) l of R = s_k points each, and
// its last stages, on the R classes q + R j of n / R points each, are sub-FFTs that fit an LDS tile (the stages
// before k never mix columns, those from k on never mix classes). k balances the two lengths; 0 if none fits.
int split_point(int policy, const SubDesc& sd) {
    int best = 0, best_len = 1 << 30, R = 1;
    for (int k = 1; k < sd.nst; ++k) {
        R *= sd.st[k - 1].r;
        const int rest = sd.n / R, len = std::max(R, rest);
        if (pass_lds_bytes(policy, R, 1) > kTile || pass_lds_bytes(policy, rest, 1) > kTile) continue;
        if (len < best_len) {
            best_len = len;
            best = k;
        }
    }
    return best;
}

void set_group(PassArgs& p, int policy) {   // sub-FFTs per workgroup within 48 KB of LDS, at most 16
    const int G = (int)std::max<size_t>(1, (size_t)48 * 1024 / pass_lds_bytes(policy, p.sd.n, 1));
    p.G = std::max(1, std::min({G, p.nsub, 16}));
    p.ngrp = (p.nsub + p.G - 1) / p.G;
}

dim3 grid_of(long long items) { return dim3((unsigned)((items + 255) / 256)); }

template <class A> hipError_t gather(int ld, const PassArgs& a, hipStream_t s) {
    const long long elems = (long long)a.ntask * a.nsub * a.sd.n;
    if (ld == LD_REAL_F64) hipLaunchKernelGGL((k_gather<A, LD_REAL_F64>), grid_of(elems), dim3(256), 0, s, a);
    else if (ld == LD_REAL_DD) hipLaunchKernelGGL((k_gather<A, LD_REAL_DD>), grid_of(elems), dim3(256), 0, s, a);
    else if (ld == LD_CPLX) hipLaunchKernelGGL((k_gather<A, LD_CPLX>), grid_of(elems), dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <class A> hipError_t scatter(int st, const PassArgs& a, const double* in, hipStream_t s) {
    const long long elems = (long long)a.ntask * a.nsub * a.sd.n;
    if (st == ST_CPLX) hipLaunchKernelGGL((k_scatter<A, ST_CPLX>), grid_of(elems), dim3(256), 0, s, a, in);
    else if (st == ST_OUT) hipLaunchKernelGGL((k_scatter<A, ST_OUT>), grid_of(elems), dim3(256), 0, s, a, in);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// the stages k.. of a's sub-FFT (R = the stride of stage k) as one LDS pass over its R classes q + R j of n / R
// points, every (task, sub-FFT) transform stored contiguously in src and dst
PassArgs tail_pass(const PassArgs& a, int k, int elt, const double* src, double* dst) {
    const int R = a.sd.st[k].s, rest = a.sd.n / R;
    const long long tdoubles = (long long)a.sd.n * elt;
    PassArgs pb = a;
    pb.sd.n = rest;
    pb.sd.nst = a.sd.nst - k;
    for (int i = k; i < a.sd.nst; ++i) pb.sd.st[i - k] = {a.sd.st[i].r, a.sd.st[i].m, a.sd.st[i].s / R, a.sd.st[i].tw};
    pb.src = src; pb.src_tstride = tdoubles; pb.dst = dst; pb.dst_tstride = tdoubles;
    pb.ntask = (int)((long long)a.ntask * a.nsub); pb.task0 = 0; pb.tdesc = nullptr;
    pb.nsub = R; pb.lde = R; pb.ldg = 1; pb.ste = R; pb.stg = 1;
    pb.twmode = 0; pb.renorm = 0; pb.pm = 0; pb.split = 0; pb.zkeep = -1;
    return pb;
}

template <class A> hipError_t run_global(int policy, int ld, int st, const PassArgs& a, hipStream_t s) {
    const long long n = a.sd.n;
    const dim3 block(256);
    auto grid = grid_of;
    if (hipError_t e = gather<A>(ld, a, s); e != hipSuccess) return e;
    double* in = a.gscr0;
    double* out = a.gscr1;
    const char* env = std::getenv("KBSRC_GLOBAL_PASSES");
    const int k = env && env[0] == '2' ? 0 : split_point(policy, a.sd);
    if (k > 0) {   // two LDS passes through the scratch buffers
        const int R = a.sd.st[k].s, rest = (int)n / R;
        const long long tdoubles = n * kElt<A>;   // one long transform in the scratch buffers
        PassArgs pa = a;
        pa.sd.n = R;
        pa.sd.nst = k;
        for (int i = 0; i < k; ++i) pa.sd.st[i] = {a.sd.st[i].r, R / (a.sd.st[i].s * a.sd.st[i].r), a.sd.st[i].s,
                                                   a.sd.st[i].tw};
        pa.src = a.gscr0; pa.src_tstride = tdoubles; pa.dst = a.gscr1; pa.dst_tstride = tdoubles;
        pa.ntask = (int)((long long)a.ntask * a.nsub); pa.task0 = 0; pa.tdesc = nullptr;
        pa.nsub = rest; pa.lde = rest; pa.ldg = 1; pa.ste = 1; pa.stg = R;
        pa.twmode = 0; pa.renorm = 0; pa.pm = rest; pa.zkeep = -1;
        set_group(pa, policy);
        hipError_t e = launch_pass(policy, LD_CPLX, ST_CPLX, pa, 256, s);
        if (e != hipSuccess) return e;
        PassArgs pb = tail_pass(a, k, kElt<A>, a.gscr1, a.gscr0);
        set_group(pb, policy);
        e = launch_pass(policy, LD_CPLX, ST_CPLX, pb, 256, s);
        if (e != hipSuccess) return e;
    } else {   // one launch per stage
        for (int si = 0; si < a.sd.nst; ++si) {
            const long long bf = (long long)a.ntask * a.nsub * (n / a.sd.st[si].r);
            hipLaunchKernelGGL(k_gstage<A>, grid(bf), block, 0, s, a, si, (const double*)in, out);
            std::swap(in, out);
        }
    }
    return scatter<A>(st, a, in, s);
}

}  // namespace

hipError_t launch_split_pass(int policy, int ld, int st, const PassArgs& a, hipStream_t s) {
    if (policy == POL_F64 || !a.gscr0 || !a.gscr1 || !a.tw16 || !a.adig || a.sd.nst < 2 || a.sd.st[0].r != 4 ||
        a.sd.st[1].r != 4)
        return hipErrorInvalidValue;
    const bool dds = policy == POL_DDS;
    hipError_t e = dds ? gather<DDS>(ld, a, s) : gather<DDA>(ld, a, s);
    if (e != hipSuccess) return e;
    const long long nsub = (long long)a.ntask * a.nsub;
    e = launch_split16_stage(a.gscr0, a.gscr1, a.sd.n, nsub, a.split, a.tw16, a.adig, s);
    if (e != hipSuccess) return e;
    const double* res = a.gscr1;
    if (a.sd.nst > 2) {
        PassArgs pb = tail_pass(a, 2, 4, a.gscr1, a.gscr0);
        set_group(pb, policy);
        e = launch_pass(policy, LD_CPLX, ST_CPLX, pb, 256, s);
        if (e != hipSuccess) return e;
        res = a.gscr0;
    }
    return dds ? scatter<DDS>(st, a, res, s) : scatter<DDA>(st, a, res, s);
}

hipError_t launch_global(int policy, int ld, int st, const PassArgs& a, hipStream_t s) {
    if (!a.gscr0 || !a.gscr1) return hipErrorInvalidValue;
    if (policy == POL_F64) return run_global<F64>(policy, ld, st, a, s);
    if (policy == POL_DDA) return run_global<DDA>(policy, ld, st, a, s);
    return run_global<DDS>(policy, ld, st, a, s);
}

}  // namespace kbsrc::gpu
