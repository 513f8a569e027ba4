// gfx950 kernels of the overlap-save engine.
//
// The arithmetic of every kernel is the reference's, operation for operation (see dd.hpp), so the output is
// bit-identical; what is MI355X-specific is how the work is laid out:
//   - sub-FFT passes: a workgroup stages G sub-FFTs in LDS as split planes (re.hi, re.lo, im.hi, im.lo), runs
//     all Stockham stages there, and reads/writes global memory once, with the lane mapping chosen so that
//     consecutive lanes touch consecutive addresses;
//   - global complex dd elements are 32 bytes {re.hi, im.hi, re.lo, im.lo}: a 16-byte "hi" complex followed by
//     a 16-byte "lo" complex, two dwordx4 accesses per element; FP64 elements are the 16-byte hi half alone;
//   - every (channel, block) task of a batch is one launch dimension, so a launch has thousands of workgroups.
#include "kernels.hpp"
#include "mid_impl.hpp"
#include "split_core.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {

namespace {

template <class A> using Val = typename A::T;
template <class A> using Cx = cx<typename A::T>;

// global complex element i of a buffer in the policy's element format
template <class A> __device__ __forceinline__ Cx<A> gload(const double* base, long long i) {
    if constexpr (A::kDD) {
        const double2* p = reinterpret_cast<const double2*>(base) + 2 * i;
        const double2 h = p[0], l = p[1];
        return {dd{h.x, l.x}, dd{h.y, l.y}};
    } else {
        const double2 v = reinterpret_cast<const double2*>(base)[i];
        return {v.x, v.y};
    }
}
template <class A> __device__ __forceinline__ void gstore(double* base, long long i, Cx<A> v) {
    if constexpr (A::kDD) {
        double2* p = reinterpret_cast<double2*>(base) + 2 * i;
        p[0] = make_double2(v.re.hi, v.im.hi);
        p[1] = make_double2(v.re.lo, v.im.lo);
    } else {
        reinterpret_cast<double2*>(base)[i] = make_double2(v.re, v.im);
    }
}
// The split/fold kernel's output, stored non-temporally: it does not take L2 space from the z bins that the next
// bin blocks read again. (The same stores in the pass kernels, whose element halves are written by separate
// instructions, made the forward row pass 1.45x slower; non-temporal loads of the passes' sources were slower too.)
__device__ __forceinline__ void st_stream2(double* p, double a, double b) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(d2{a, b}, reinterpret_cast<d2*>(p));
}
template <class A> __device__ __forceinline__ void gstore_stream(double* base, long long i, Cx<A> v) {
    if constexpr (A::kDD) {
        st_stream2(base + 4 * i, v.re.hi, v.im.hi);
        st_stream2(base + 4 * i + 2, v.re.lo, v.im.lo);
    } else {
        st_stream2(base + 2 * i, v.re, v.im);
    }
}
template <class A> __device__ __forceinline__ Val<A> gload_real(const double* base, long long i) {
    if constexpr (A::kDD) {
        const double2 v = reinterpret_cast<const double2*>(base)[i];
        return dd{v.x, v.y};
    } else {
        return base[i];
    }
}

// LDS split planes
template <class A> struct Planes {
    double* b;
    int cap;
    __device__ __forceinline__ Cx<A> get(int i) const {
        if constexpr (A::kDD) return {dd{b[i], b[cap + i]}, dd{b[2 * cap + i], b[3 * cap + i]}};
        else return {b[i], b[cap + i]};
    }
    __device__ __forceinline__ void put(int i, Cx<A> v) const {
        if constexpr (A::kDD) {
            b[i] = v.re.hi; b[cap + i] = v.re.lo; b[2 * cap + i] = v.im.hi; b[3 * cap + i] = v.im.lo;
        } else {
            b[i] = v.re; b[cap + i] = v.im;
        }
    }
};
template <class A> constexpr int kPlanes = A::kDD ? 4 : 2;

template <class A> __device__ __forceinline__ Cx<A> renorm(Cx<A> v) {
    if constexpr (A::kDD) return {quick2(v.re.hi, v.re.lo), quick2(v.im.hi, v.im.lo)};
    else return v;
}
template <class A> __device__ __forceinline__ double to_f64(Val<A> v) {
    if constexpr (A::kDD) return v.hi + v.lo;
    else return v;
}
// output sample k of a buffer: the value rounded to double, or with pairs (PassArgs::out_dd, dd policies) the dd
// value as the (hi, lo) pair at 2 k
template <class A> __device__ __forceinline__ void put_sample(double* out, int pairs, long long k, Val<A> v) {
    if constexpr (A::kDD) {
        if (pairs) {
            out[2 * k] = v.hi;
            out[2 * k + 1] = v.lo;
            return;
        }
    }
    out[k] = to_f64<A>(v);
}

// Odd-radix butterfly by symmetric sums (src/dsp/fft.hpp:62-84)
template <class A, int R>
__device__ __forceinline__ void bfly_odd(const Planes<A>& cur, const Planes<A>& nxt, int Gs, int g, int sbase,
                                         int sm, int dbase, int s, bool twid, const double* tw,
                                         const double* oddc) {
    constexpr int H = (R - 1) / 2;
    constexpr int ri = R == 3 ? 0 : R == 5 ? 1 : 2;
    const Cx<A> a0 = cur.get(sbase * Gs + g);
    Cx<A> S[H], D[H];
#pragma unroll
    for (int t = 1; t <= H; ++t) {
        const Cx<A> x = cur.get((sbase + t * sm) * Gs + g), y = cur.get((sbase + (R - t) * sm) * Gs + g);
        S[t - 1] = c_add<A>(x, y);
        D[t - 1] = c_sub<A>(x, y);
    }
    Cx<A> sum = a0;
#pragma unroll
    for (int t = 0; t < H; ++t) sum = c_add<A>(sum, S[t]);
    nxt.put(dbase * Gs + g, sum);
#pragma unroll
    for (int u = 1; u <= H; ++u) {
        Cx<A> ac = a0, bc;
#pragma unroll
        for (int t = 1; t <= H; ++t) {
            const Cx<A> cs = gload<A>(oddc, ri * 7 + (t * u) % R);   // (cos, sin) of 2 pi m / R
            ac = {A::add(ac.re, A::mul(S[t - 1].re, cs.re)), A::add(ac.im, A::mul(S[t - 1].im, cs.re))};
            if (t == 1)   // the sum starts at zero: zero plus the first product (A::add0)
                bc = {A::add0(A::mul(D[0].re, cs.im)), A::add0(A::mul(D[0].im, cs.im))};
            else
                bc = {A::add(bc.re, A::mul(D[t - 1].re, cs.im)), A::add(bc.im, A::mul(D[t - 1].im, cs.im))};
        }
        Cx<A> xu = {A::add(ac.re, bc.im), A::sub(ac.im, bc.re)};
        Cx<A> xr = {A::sub(ac.re, bc.im), A::add(ac.im, bc.re)};
        if (twid) {
            xu = c_mul<A>(xu, gload<A>(tw, u - 1));
            xr = c_mul<A>(xr, gload<A>(tw, R - u - 1));
        }
        nxt.put((dbase + u * s) * Gs + g, xu);
        nxt.put((dbase + (R - u) * s) * Gs + g, xr);
    }
}

// SPL > 0 (split form, dd policies, PassArgs::pm == 0): the first two radix-4 stages run as one int8-MFMA
// radix-16 stage with SPL slices (split_core.hpp); y[16 p + k] = w_n^{p k} DFT16(x[p + (n / 16) m])[k], the
// twiddle in dd with the accurate add, as the global-memory route's stage computes it.
template <class A, int LD, int ST, int SPL>
__device__ __forceinline__ void subfft_body(const PassArgs& a, int bid) {
    extern __shared__ double lds[];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int task = bid / a.ngrp, grp = bid % a.ngrp;
    const int Gs = a.G, j0 = grp * Gs;
    const int G = min(Gs, a.nsub - j0);
    const int n = a.sd.n;
    const int cap = n * Gs;
    Planes<A> cur{lds, cap}, nxt{lds + kPlanes<A> * cap, cap};
    const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);

    // ---- load G sub-FFTs into LDS
    {
        const bool gfast = a.ldg == 1 && a.lde != 1;
        for (int w = tid; w < G * n; w += nthr) {
            const int g = gfast ? w % G : w / n, e = gfast ? w / G : w % n;
            const long long i = e * a.lde + (long long)(j0 + g) * a.ldg;
            Cx<A> v;
            if constexpr (LD == LD_REAL_F64) {
                const double* s = a.src + io.in_off;
                v = {A::lift(s[2 * i]), A::lift(s[2 * i + 1])};
            } else if constexpr (LD == LD_REAL_DD) {
                const double* s = a.src + task * a.src_tstride;
                v = {gload_real<A>(s, 2 * i), gload_real<A>(s, 2 * i + 1)};
            } else {
                v = gload<A>(a.src + task * a.src_tstride, i);
            }
            cur.put(e * Gs + g, v);
        }
    }
    __syncthreads();

    int si0 = 0;
    if constexpr (SPL > 0) {
        const int nv = n >> 4, NV = G * nv;
        const int lane = tid & 63, q = lane >> 4, col = lane & 15;
        for (int base = (tid >> 6) * 16; base < NV; base += (nthr >> 6) * 16) {   // uniform over the wave
            const int v = base + col;
            const bool live = v < NV;
            const int g = live ? v % G : 0, p = live ? v / G : 0;
            double hi[8], lo[8];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (live) {
                    const Cx<A> c = cur.get((p + nv * (4 * q + i)) * Gs + g);
                    hi[2 * i] = c.re.hi;
                    hi[2 * i + 1] = c.im.hi;
                    lo[2 * i] = c.re.lo;
                    lo[2 * i + 1] = c.im.lo;
                } else {
                    hi[2 * i] = hi[2 * i + 1] = lo[2 * i] = lo[2 * i + 1] = 0.0;
                }
            }
            dd re[4], im[4];
            split::split_core<SPL>(hi, lo, a.adig, lane, re, im);
            if (live) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int k = 4 * q + r;
                    cx<dd> z{re[r], im[r]};
                    if (p != 0 && k != 0) z = c_mul<DDA>(z, gload<DDA>(a.tw16, (long long)p * 15 + k - 1));
                    nxt.put((16 * p + k) * Gs + g, z);
                }
            }
        }
        __syncthreads();
        const Planes<A> t = cur;
        cur = nxt;
        nxt = t;
        si0 = 2;
    }

    // ---- Stockham stages, ping-pong between the two LDS halves (src/dsp/fft.hpp:115-150)
    for (int si = si0; si < a.sd.nst; ++si) {
        const StageDesc st = a.sd.st[si];
        const int r = st.r, s = st.s, sm = n / r;
        const int nbf = sm;   // butterflies per sub-FFT
        for (int w = tid; w < G * nbf; w += nthr) {
            const int g = w % G, b = w / G;
            const int p = b / s, q = b - p * s;
            const int dbase = s * r * p + q;
            // a leading part of a longer transform (see PassArgs::pm) twiddles by the butterfly's index there
            const long long pg = a.pm ? (long long)(j0 + g) + (long long)a.pm * p : p;
            const bool twid = pg != 0;
            const double* tw = a.tw + (st.tw + pg * (r - 1)) * (A::kDD ? 4 : 2);
            if (r == 4) {
                const Cx<A> a0 = cur.get(b * Gs + g), a1 = cur.get((b + sm) * Gs + g);
                const Cx<A> a2 = cur.get((b + 2 * sm) * Gs + g), a3 = cur.get((b + 3 * sm) * Gs + g);
                const Cx<A> t0 = c_add<A>(a0, a2), t1 = c_sub<A>(a0, a2), t2 = c_add<A>(a1, a3);
                const Cx<A> t3 = c_mul_mi<A>(c_sub<A>(a1, a3));
                nxt.put(dbase * Gs + g, c_add<A>(t0, t2));
                Cx<A> o1 = c_add<A>(t1, t3), o2 = c_sub<A>(t0, t2), o3 = c_sub<A>(t1, t3);
                if (twid) {
                    o1 = c_mul<A>(o1, gload<A>(tw, 0));
                    o2 = c_mul<A>(o2, gload<A>(tw, 1));
                    o3 = c_mul<A>(o3, gload<A>(tw, 2));
                }
                nxt.put((dbase + s) * Gs + g, o1);
                nxt.put((dbase + 2 * s) * Gs + g, o2);
                nxt.put((dbase + 3 * s) * Gs + g, o3);
            } else if (r == 2) {
                const Cx<A> a0 = cur.get(b * Gs + g), a1 = cur.get((b + sm) * Gs + g);
                nxt.put(dbase * Gs + g, c_add<A>(a0, a1));
                Cx<A> d = c_sub<A>(a0, a1);
                if (twid) d = c_mul<A>(d, gload<A>(tw, 0));
                nxt.put((dbase + s) * Gs + g, d);
            } else if (r == 3) {
                bfly_odd<A, 3>(cur, nxt, Gs, g, b, sm, dbase, s, twid, tw, a.oddc);
            } else if (r == 5) {
                bfly_odd<A, 5>(cur, nxt, Gs, g, b, sm, dbase, s, twid, tw, a.oddc);
            } else {
                bfly_odd<A, 7>(cur, nxt, Gs, g, b, sm, dbase, s, twid, tw, a.oddc);
            }
        }
        __syncthreads();
        const Planes<A> t = cur;
        cur = nxt;
        nxt = t;
    }

    // ---- store: four-step twiddle, renormalisation, then either a complex buffer or the output samples
    {
        const bool gfast = a.stg == 1 && a.ste != 1;
        for (int w = tid; w < G * n; w += nthr) {
            const int g = gfast ? w % G : w / n, e = gfast ? w / G : w % n;
            const long long jg = j0 + g;
            const long long i = e * a.ste + jg * a.stg;
            if (ST == ST_CPLX && z_unused(a, i)) continue;
            Cx<A> v = cur.get(e * Gs + g);
            if (a.twmode && e != 0 && (a.twmode == 1 || jg != 0)) v = c_mul<A>(v, gload<A>(a.tw4, i));
            if (a.renorm) v = renorm<A>(v);
            if constexpr (ST == ST_CPLX) {
                gstore<A>(a.dst + task * a.dst_tstride, i, v);
            } else {
                // r[2i] = re, r[2i+1] = -im (src/dsp/fft.hpp:460); keep the block's hop (src/resampler.cpp:80-84)
                const long long x0 = 2 * i - a.margP;
                if (x0 >= 0 && x0 < a.hopP) {
                    const long long d = io.blk * a.hopP + x0 - a.o;
                    if (d >= 0 && d < io.nout) put_sample<A>(a.out, a.out_dd, io.out_off + d, v.re);
                }
                const long long x1 = x0 + 1;
                if (x1 >= 0 && x1 < a.hopP) {
                    const long long d = io.blk * a.hopP + x1 - a.o;
                    if (d >= 0 && d < io.nout) put_sample<A>(a.out, a.out_dd, io.out_off + d, A::neg(v.im));
                }
            }
        }
    }
}

template <class A, int LD, int ST, int SPL = 0>
__global__ __launch_bounds__(256) void k_subfft(PassArgs a) {
    subfft_body<A, LD, ST, SPL>(a, blockIdx.x);
}

// One launch over the passes of several plan shapes (a batch of short jobs): workgroups [starts[s], starts[s + 1])
// run shape s's pass d[s] as k_subfft would.
template <class A, int LD, int ST>
__global__ __launch_bounds__(256) void k_subfft_multi(const PassArgs* __restrict__ d, const int* __restrict__ starts,
                                                      int n) {
    const int bid = blockIdx.x;
    int s = 0;
    while (s + 1 < n && bid >= starts[s + 1]) ++s;
    subfft_body<A, LD, ST, 0>(d[s], bid - starts[s]);
}

// forward split X[i] of a global array of the complex FFT's bins (the spectrum design's k_spec_split)
template <class A> __device__ __forceinline__ Cx<A> fwd_split(const double* z, const double* w, int h, int i) {
    return fwd_split_at<A>([z](int k) { return gload<A>(z, k); }, MidTabs{nullptr, w, nullptr}, h, i);
}

// Bin block of the q-th block in the launch order. With c = kchain > 0 the nk blocks run as the columns of a
// matrix with rows of c blocks (b, b + c, b + 2c, ..., then b + 1, ...): a thread's eight z reads are shared
// with the threads |h - Mb| bins away, so both reads of a z bin fall close together in time and the second one
// finds it in cache. A bijection of [0, nk).
__device__ __forceinline__ int mid_block(int q, int nk, int c) {
    if (c <= 0 || c >= nk) return q;
    const int R = (nk + c - 1) / c, full = nk - (R - 1) * c;   // columns 0..full-1 hold R blocks, the rest R - 1
    int j, i;
    if (q < full * R) {
        j = q / R; i = q - j * R;
    } else {
        const int r = q - full * R;
        j = full + r / (R - 1); i = r - (j - full) * (R - 1);
    }
    return j + i * c;
}

// Task and launch-order bin block of workgroup L. tgroup = 0: tasks fastest (every task of a bin block, then the next
// block). tgroup > 0: tasks in groups of tgroup; a group runs all its bin blocks, its tasks fastest, before the next
// group. With the chained block order, a block's z reads that the next block repeats then come back tgroup
// workgroups later, and with the launch's workgroups dealt to the XCDs in contiguous ranges (k_mid), on the same
// XCD: from its L2, not from HBM. H and the split twiddles are read once per group instead of once per launch.
__device__ __forceinline__ void mid_order(const MidArgs& a, int nk, unsigned L, int& task, int& q) {
    if (a.tgroup <= 0) {
        task = (int)(L % (unsigned)a.ntask);
        q = (int)(L / (unsigned)a.ntask);
        return;
    }
    const unsigned per = (unsigned)a.tgroup * (unsigned)nk;   // workgroups of a full group
    const int grp = (int)(L / per), t0 = grp * a.tgroup, gs = min(a.tgroup, a.ntask - t0);
    const unsigned r = L - (unsigned)grp * per;
    q = (int)(r / (unsigned)gs);
    task = t0 + (int)(r - (unsigned)q * (unsigned)gs);
}

template <class A> __device__ __forceinline__ void mid_body(const MidArgs& a, unsigned bid) {
    const int nk = (a.hinv / 2 + 1 + (int)blockDim.x - 1) / (int)blockDim.x;
    int task, q;
    mid_order(a, nk, bid, task, q);
    const int k = mid_block(q, nk, a.kchain) * blockDim.x + threadIdx.x;
    if (k > a.hinv / 2) return;
    const double* z = a.z + task * a.z_tstride;
    double* zi = a.zi + task * a.zi_tstride;
    mid_pair<A>(a, [z](int i) { return gload<A>(z, i); }, MidTabs{a.H, a.wf, a.wi}, k,
                [zi](int i, Cx<A> v) { gstore_stream<A>(zi, i, v); });
}

// Workgroups are dealt to the 8 XCDs round-robin; with tgroup > 0 each XCD takes a contiguous range of the launch
// order instead, so that a task group's blocks share one L2.
template <class A> __global__ __launch_bounds__(256) void k_mid(MidArgs a) {
    unsigned L = blockIdx.x;
    if (a.tgroup > 0) {
        const unsigned n = gridDim.x, full = n & ~7u;
        if (L < full) L = (L & 7u) * (full >> 3) + (L >> 3);
    }
    mid_body<A>(a, L);
}

template <class A>
__global__ __launch_bounds__(256) void k_mid_multi(const MidArgs* __restrict__ d, const int* __restrict__ starts, int n) {
    const int bid = blockIdx.x;
    int s = 0;
    while (s + 1 < n && bid >= starts[s + 1]) ++s;
    mid_body<A>(d[s], (unsigned)(bid - starts[s]));
}

__global__ __launch_bounds__(256) void k_stage(StageArgs a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= a.tot) return;
    const long long pos = i - a.off;   // sample index in the signal
    double v = 0.0;
    if (pos >= 0 && pos < a.nin) v = a.in[pos * a.ch + c];
    else if (pos >= a.nin && pos < a.nin + a.ext) v = a.tail[c * a.ext + (pos - a.nin)];
    else if (pos < 0 && pos >= -a.ext) v = a.head[c * a.ext + (-1 - pos)];
    a.xx[c * a.stride + i] = v;
}

__global__ __launch_bounds__(256) void k_spec_split(SpecArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > a.h) return;
    const Cx<DDA> x = fwd_split<DDA>(a.z, a.w, a.h, i);
    const dd v = dddiv1(x.re, a.scale);
    a.H[2 * i] = v.hi;
    a.H[2 * i + 1] = v.lo;
}

// ---- Bluestein steps (src/dsp/fft.hpp:268-278); arithmetic of the storage type (scalar dd or FP64)
template <class A, int LD> __global__ __launch_bounds__(256) void k_blue_pre(BlueArgs a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, task = blockIdx.y;
    if (j >= a.m) return;
    Cx<A> u = {A::lift(0.0), A::lift(0.0)};
    if (j < a.n) {
        Cx<A> x;
        if constexpr (LD == LD_REAL_F64) {
            const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
            const double* s = a.src + io.in_off;
            x = {A::lift(s[2 * j]), A::lift(s[2 * j + 1])};
        } else if constexpr (LD == LD_REAL_DD) {
            const double* s = a.src + task * a.src_tstride;
            x = {gload_real<A>(s, 2 * j), gload_real<A>(s, 2 * j + 1)};
        } else {
            x = gload<A>(a.src + task * a.src_tstride, j);
        }
        u = c_mul<A>(x, c_conj<A>(gload<A>(a.bc, j)));
    }
    gstore<A>(a.dst + task * a.dst_tstride, j, u);
}

template <class A> __global__ __launch_bounds__(256) void k_blue_mid(BlueArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, task = blockIdx.y;
    if (k >= a.m) return;
    double* U = a.dst + task * a.dst_tstride;
    gstore<A>(U, k, c_conj<A>(c_mul<A>(gload<A>(U, k), gload<A>(a.vf, k))));
}

template <class A, int ST> __global__ __launch_bounds__(256) void k_blue_post(BlueArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, task = blockIdx.y;
    if (k >= a.n) return;
    const Cx<A> y = c_mul<A>(c_conj<A>(gload<A>(a.src + task * a.src_tstride, k)), c_conj<A>(gload<A>(a.bc, k)));
    if constexpr (ST == ST_CPLX) {
        gstore<A>(a.dst + task * a.dst_tstride, k, y);
    } else {
        const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
        const long long x0 = 2LL * k - a.margP, x1 = x0 + 1;
        if (x0 >= 0 && x0 < a.hopP) {
            const long long d = io.blk * a.hopP + x0 - a.o;
            if (d >= 0 && d < io.nout) put_sample<A>(a.out, a.out_dd, io.out_off + d, y.re);
        }
        if (x1 >= 0 && x1 < a.hopP) {
            const long long d = io.blk * a.hopP + x1 - a.o;
            if (d >= 0 && d < io.nout) put_sample<A>(a.out, a.out_dd, io.out_off + d, A::neg(y.im));
        }
    }
}

template <class A, int LD, int ST> hipError_t launch_one(const PassArgs& a, int threads, hipStream_t s) {
    const size_t lds = (size_t)2 * kPlanes<A> * a.sd.n * a.G * sizeof(double);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_subfft<A, LD, ST>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((k_subfft<A, LD, ST>), dim3(a.ngrp * a.ntask), dim3(threads), lds, s, a);
    return hipGetLastError();
}

template <int LD, int ST, int SPL> hipError_t launch_one_split(const PassArgs& a, int threads, hipStream_t s) {
    const size_t lds = (size_t)2 * kPlanes<DDS> * a.sd.n * a.G * sizeof(double);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_subfft<DDS, LD, ST, SPL>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((k_subfft<DDS, LD, ST, SPL>), dim3(a.ngrp * a.ntask), dim3(threads), lds, s, a);
    return hipGetLastError();
}

template <int SPL> hipError_t launch_split_slices(int ld, int st, const PassArgs& a, int threads, hipStream_t s) {
    if (ld == LD_REAL_F64 && st == ST_CPLX) return launch_one_split<LD_REAL_F64, ST_CPLX, SPL>(a, threads, s);
    if (ld == LD_CPLX && st == ST_CPLX) return launch_one_split<LD_CPLX, ST_CPLX, SPL>(a, threads, s);
    if (ld == LD_CPLX && st == ST_OUT) return launch_one_split<LD_CPLX, ST_OUT, SPL>(a, threads, s);
    return hipErrorInvalidValue;
}

}  // namespace

bool split_lds_ok(int policy, int ld, int st, const PassArgs& a) {
    const bool slices = a.split == 6 || a.split == 7 || a.split == 8 || a.split == 15;
    const bool io = (ld == LD_REAL_F64 && st == ST_CPLX) || (ld == LD_CPLX && (st == ST_CPLX || st == ST_OUT));
    return policy == POL_DDS && slices && io && a.pm == 0 && a.tw16 && a.adig && a.sd.nst >= 2 &&
           a.sd.st[0].r == 4 && a.sd.st[1].r == 4 && pass_lds_bytes(policy, a.sd.n, 1) <= 160 * 1024;
}

hipError_t launch_split_lds(int policy, int ld, int st, const PassArgs& a, int threads, hipStream_t s) {
    if (!split_lds_ok(policy, ld, st, a) || threads % 64 || pass_lds_bytes(policy, a.sd.n, a.G) > 160 * 1024)
        return hipErrorInvalidValue;
    switch (a.split) {
        case 6: return launch_split_slices<6>(ld, st, a, threads, s);
        case 7: return launch_split_slices<7>(ld, st, a, threads, s);
        case 8: return launch_split_slices<8>(ld, st, a, threads, s);
        case 15: return launch_split_slices<15>(ld, st, a, threads, s);
        default: return hipErrorInvalidValue;
    }
}

size_t pass_lds_bytes(int policy, int n, int G) {
    return (size_t)2 * (policy == POL_F64 ? 2 : 4) * n * G * sizeof(double);
}

hipError_t launch_pass(int policy, int ld, int st, const PassArgs& a, int threads, hipStream_t s) {
#define KB_CASE(P, POL, L, S) \
    if (policy == P && ld == L && st == S) return launch_one<POL, L, S>(a, threads, s);
    KB_CASE(POL_DDS, DDS, LD_REAL_F64, ST_CPLX)
    KB_CASE(POL_DDS, DDS, LD_CPLX, ST_CPLX)
    KB_CASE(POL_DDS, DDS, LD_CPLX, ST_OUT)
    KB_CASE(POL_DDA, DDA, LD_REAL_F64, ST_CPLX)
    KB_CASE(POL_DDA, DDA, LD_CPLX, ST_CPLX)
    KB_CASE(POL_DDA, DDA, LD_CPLX, ST_OUT)
    KB_CASE(POL_DDA, DDA, LD_REAL_DD, ST_CPLX)
    KB_CASE(POL_F64, F64, LD_REAL_F64, ST_CPLX)
    KB_CASE(POL_F64, F64, LD_CPLX, ST_CPLX)
    KB_CASE(POL_F64, F64, LD_CPLX, ST_OUT)
#undef KB_CASE
    return hipErrorInvalidValue;
}

int mid_workgroups(const MidArgs& a) {
    const unsigned long long n = (unsigned long long)((a.hinv / 2 + 1 + 255) / 256) * (unsigned long long)a.ntask;
    return a.ntask < 1 || n > 0x7fffffffULL ? -1 : (int)n;
}

hipError_t launch_mid_multi(int policy, const MidArgs* d, const int* starts, int n, int total, hipStream_t s) {
    if (n < 1 || total < 1) return hipErrorInvalidValue;
    if (policy == POL_F64) hipLaunchKernelGGL(k_mid_multi<F64>, dim3(total), dim3(256), 0, s, d, starts, n);
    else hipLaunchKernelGGL(k_mid_multi<DDA>, dim3(total), dim3(256), 0, s, d, starts, n);
    return hipGetLastError();
}

hipError_t launch_pass_multi(int policy, int ld, int st, const PassArgs* d, const int* starts, int n, int total,
                             size_t lds, hipStream_t s) {
    if (n < 1 || total < 1 || lds > 160 * 1024) return hipErrorInvalidValue;
#define KB_MULTI(P, POL, L, S)                                                                                     \
    if (policy == P && ld == L && st == S) {                                                                       \
        static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_subfft_multi<POL, L, S>), \
                                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024); \
        if (attr != hipSuccess) return attr;                                                                       \
        hipLaunchKernelGGL((k_subfft_multi<POL, L, S>), dim3(total), dim3(256), lds, s, d, starts, n);              \
        return hipGetLastError();                                                                                   \
    }
    KB_MULTI(POL_DDS, DDS, LD_REAL_F64, ST_CPLX)
    KB_MULTI(POL_DDS, DDS, LD_CPLX, ST_CPLX)
    KB_MULTI(POL_DDS, DDS, LD_CPLX, ST_OUT)
    KB_MULTI(POL_DDA, DDA, LD_REAL_F64, ST_CPLX)
    KB_MULTI(POL_DDA, DDA, LD_CPLX, ST_CPLX)
    KB_MULTI(POL_DDA, DDA, LD_CPLX, ST_OUT)
    KB_MULTI(POL_F64, F64, LD_REAL_F64, ST_CPLX)
    KB_MULTI(POL_F64, F64, LD_CPLX, ST_CPLX)
    KB_MULTI(POL_F64, F64, LD_CPLX, ST_OUT)
#undef KB_MULTI
    return hipErrorInvalidValue;
}

hipError_t launch_mid(int policy, const MidArgs& a, hipStream_t s) {
    const unsigned long long n = (unsigned long long)((a.hinv / 2 + 1 + 255) / 256) * (unsigned long long)a.ntask;
    if (a.ntask < 1 || n > 0x7fffffffULL) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)n);
    if (policy == POL_F64) hipLaunchKernelGGL(k_mid<F64>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_mid<DDA>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_blue(int policy, int phase, int kind, const BlueArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)(((phase == BLUE_POST ? a.n : a.m) + 255) / 256), a.ntask), block(256);
#define KB_BLUE(POL)                                                                                             \
    if (phase == BLUE_PRE && kind == LD_REAL_F64) hipLaunchKernelGGL((k_blue_pre<POL, LD_REAL_F64>), grid, block, 0, s, a); \
    else if (phase == BLUE_PRE && kind == LD_REAL_DD) hipLaunchKernelGGL((k_blue_pre<POL, LD_REAL_DD>), grid, block, 0, s, a); \
    else if (phase == BLUE_PRE) hipLaunchKernelGGL((k_blue_pre<POL, LD_CPLX>), grid, block, 0, s, a);                \
    else if (phase == BLUE_MID) hipLaunchKernelGGL((k_blue_mid<POL>), grid, block, 0, s, a);                        \
    else if (kind == ST_OUT) hipLaunchKernelGGL((k_blue_post<POL, ST_OUT>), grid, block, 0, s, a);                  \
    else hipLaunchKernelGGL((k_blue_post<POL, ST_CPLX>), grid, block, 0, s, a);
    if (policy == POL_F64) {
        if (kind == LD_REAL_DD && phase == BLUE_PRE) return hipErrorInvalidValue;
        KB_BLUE(F64)
    } else {
        KB_BLUE(DDA)
    }
#undef KB_BLUE
    return hipGetLastError();
}

hipError_t launch_stage(const StageArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_stage, dim3((unsigned)((a.tot + 255) / 256), a.ch), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_spec_split(const SpecArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_spec_split, dim3((a.h + 1 + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace kbsrc::gpu
