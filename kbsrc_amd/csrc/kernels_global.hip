// Sub-FFT passes whose length does not fit LDS (e.g. the 3200-point rows of a 10 MHz -> 48 kHz plan): the same
// Stockham stages as the LDS kernels, one launch per stage, ping-ponging through two global scratch buffers.
// Sub-FFT j of task t occupies scratch[(t * nsub + j) * n .. + n). Arithmetic identical to the other passes.
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
    Cx<A> v = gld<A>(buf, idx);
    if (a.twmode && e != 0 && (a.twmode == 1 || j != 0)) v = c_mul<A>(v, gld<A>(a.tw4, i));
    if constexpr (A::kDD) {
        if (a.renorm) v = {quick2(v.re.hi, v.re.lo), quick2(v.im.hi, v.im.lo)};
    }
    if constexpr (ST == ST_CPLX) {
        gst<A>(a.dst + task * a.dst_tstride, i, v);
    } else {
        const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
        double* o = a.out + io.out_off;
        const long long x0 = 2 * i - a.margP, x1 = x0 + 1;
        if (x0 >= 0 && x0 < a.hopP) {
            const long long d = io.blk * a.hopP + x0 - a.o;
            if (d >= 0 && d < io.nout) {
                if constexpr (A::kDD) o[d] = v.re.hi + v.re.lo;
                else o[d] = v.re;
            }
        }
        if (x1 >= 0 && x1 < a.hopP) {
            const long long d = io.blk * a.hopP + x1 - a.o;
            if (d >= 0 && d < io.nout) {
                const Val<A> ni = A::neg(v.im);
                if constexpr (A::kDD) o[d] = ni.hi + ni.lo;
                else o[d] = ni;
            }
        }
    }
}

template <class A> hipError_t run_global(int ld, int st, const PassArgs& a, hipStream_t s) {
    const long long n = a.sd.n, elems = (long long)a.ntask * a.nsub * n;
    const dim3 block(256);
    auto grid = [](long long items) { return dim3((unsigned)((items + 255) / 256)); };
    if (ld == LD_REAL_F64) hipLaunchKernelGGL((k_gather<A, LD_REAL_F64>), grid(elems), block, 0, s, a);
    else if (ld == LD_REAL_DD) hipLaunchKernelGGL((k_gather<A, LD_REAL_DD>), grid(elems), block, 0, s, a);
    else if (ld == LD_CPLX) hipLaunchKernelGGL((k_gather<A, LD_CPLX>), grid(elems), block, 0, s, a);
    else return hipErrorInvalidValue;
    double* in = a.gscr0;
    double* out = a.gscr1;
    for (int si = 0; si < a.sd.nst; ++si) {
        const long long bf = (long long)a.ntask * a.nsub * (n / a.sd.st[si].r);
        hipLaunchKernelGGL(k_gstage<A>, grid(bf), block, 0, s, a, si, (const double*)in, out);
        std::swap(in, out);
    }
    if (st == ST_CPLX) hipLaunchKernelGGL((k_scatter<A, ST_CPLX>), grid(elems), block, 0, s, a, (const double*)in);
    else if (st == ST_OUT) hipLaunchKernelGGL((k_scatter<A, ST_OUT>), grid(elems), block, 0, s, a, (const double*)in);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace

hipError_t launch_global(int policy, int ld, int st, const PassArgs& a, hipStream_t s) {
    if (!a.gscr0 || !a.gscr1) return hipErrorInvalidValue;
    if (policy == POL_F64) return run_global<F64>(ld, st, a, s);
    if (policy == POL_DDA) return run_global<DDA>(ld, st, a, s);
    return run_global<DDS>(ld, st, a, s);
}

}  // namespace kbsrc::gpu
