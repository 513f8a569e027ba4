// Sub-FFT passes specialised at compile time for given transform lengths (included by the kernel tables).
//
// Same arithmetic as the generic pass in kernels.hip (and so as the reference: src/dsp/fft.hpp:86-151), but the
// radix sequence, the number of sub-FFTs per workgroup G and the thread count T are template constants. That
// turns every index division into shifts or multiplies, unrolls the butterflies, and lets the stages run in a
// single LDS buffer: each thread computes all of its butterflies of a stage into registers, the workgroup
// synchronises, and the results are written back in place. Half the LDS of the ping-pong form, so twice the
// workgroups per CU.
#pragma once
#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <utility>

#include "bfly.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {

namespace {

template <class A, int CAP> struct Lds {
    double* b;
    __device__ __forceinline__ Cx<A> get(int i) const {
        if constexpr (A::kDD) return {dd{b[i], b[CAP + i]}, dd{b[2 * CAP + i], b[3 * CAP + i]}};
        else return {b[i], b[CAP + i]};
    }
    __device__ __forceinline__ void put(int i, Cx<A> v) const {
        if constexpr (A::kDD) {
            b[i] = v.re.hi; b[CAP + i] = v.re.lo; b[2 * CAP + i] = v.im.hi; b[3 * CAP + i] = v.im.lo;
        } else {
            b[i] = v.re; b[CAP + i] = v.im;
        }
    }
};

// Compile-time stage list in the reference's radix order.
template <int... R> struct Radices {
    static constexpr int nst = sizeof...(R);
    static constexpr int r[nst] = {R...};
    static constexpr int N = (R * ...);
    static constexpr int stride(int i) {   // product of the earlier radices
        int p = 1;
        for (int k = 0; k < i; ++k) p *= r[k];
        return p;
    }
    static constexpr int twoff(int i) {    // offset of stage i's twiddles: sum of m_k (r_k - 1) before it
        int o = 0;
        for (int k = 0; k < i; ++k) o += (N / (stride(k) * r[k])) * (r[k] - 1);
        return o;
    }
    // complex values one thread holds in the widest stage, for G sub-FFTs on T threads
    static constexpr int hold(int G, int T) {
        int h = 0;
        for (int k = 0; k < nst; ++k) {
            const int W = G * (N / r[k]), it = (W + T - 1) / T;
            h = it * r[k] > h ? it * r[k] : h;
        }
        return h;
    }
};

// The sub-FFT outputs some later reader needs (PassArgs::wstart, wlen, from zkeep): e in the cyclic window
// [wstart, wstart + wlen) mod N; wlen < 0: all of them. A Stockham stage of stride s needs exactly its butterflies
// (p, q) with q in that window reduced mod s: butterfly (p, q) writes elements s r p + q + u s, which the next stage
// (stride s r) reads only where their residue q + u s mod s r lies in the window reduced mod s r, and the last
// stage's outputs are the sub-FFT's. Stages whose stride exceeds the window run only those butterflies; the
// skipped ones' results would never be read, so no stored value changes.
struct Prune {
    int wstart = 0, wlen = -1;
};

// Where a stage's twiddles come from: the stage's table and the index there of butterfly p of local sub-FFT g
// (index 0: no twiddle).
struct OwnTw {   // the sub-FFT's own tables: stage I's at RL::twoff(I), butterfly p
    const double* tw;
    template <class A, class RL, int I> __device__ __forceinline__ const double* table() const {
        return tw + RL::twoff(I) * kElt<A>;
    }
    __device__ __forceinline__ int index(int, int p) const { return p; }
};
// Stages of a longer transform (the two parts of a long sub-FFT): stage I's table at st[I].tw, the long
// transform's; butterfly p of local sub-FFT g is the long transform's butterfly c0 + cg g + pm p (a leading part
// on column c0 + cg g of the long transform takes pm = its column count; a trailing part c0 = cg = 0, pm = 1).
struct LongTw {
    const double* tw;
    const StageDesc* st;
    int c0, cg, pm;
    template <class A, class RL, int I> __device__ __forceinline__ const double* table() const {
        return tw + st[I].tw * kElt<A>;
    }
    __device__ __forceinline__ int index(int g, int p) const { return c0 + cg * g + pm * p; }
};

// A stage whose stride s exceeds the window: the butterflies (p, q), q = (qs + k) mod s for k < len, every p.
template <class A, class RL, int G, int T, int I, class TW>
__device__ __forceinline__ void stage_window(const Lds<A, RL::N * G>& L, int tid, const TW& tw, const double* tws,
                                             const double* oddc, int qs, int len) {
    constexpr int N = RL::N, r = RL::r[I], s = RL::stride(I), nbf = N / r, W = G * nbf, IT = (W + T - 1) / T;
    const int Wn = G * (nbf / s) * len;
    auto at = [&](int w, int& g, int& p, int& q) {
        g = w % G;
        const int k = w / G;
        p = k / len;
        q = qs + (k - p * len);
        if (q >= s) q -= s;
    };
    Cx<A> held[IT][r];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = tid + it * T;
        if (w < Wn) {
            int g, p, q;
            at(w, g, p, q);
            Cx<A> x[r];
#pragma unroll
            for (int t = 0; t < r; ++t) x[t] = L.get((p * s + q + t * nbf) * G + g);
            const int pg = tw.index(g, p);
            butterfly<A, r>(x, held[it], pg != 0, tws, oddc, (unsigned)(pg * (r - 1)));
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = tid + it * T;
        if (w < Wn) {
            int g, p, q;
            at(w, g, p, q);
#pragma unroll
            for (int u = 0; u < r; ++u) L.put((s * r * p + q + u * s) * G + g, held[it][u]);
        }
    }
    __syncthreads();
}

// One Stockham stage over G sub-FFTs in place: compute all of this thread's butterflies, sync, write back; only
// the butterflies some reader needs when the window is narrower than the stride (stage_window).
template <class A, class RL, int G, int T, int I, class TW>
__device__ __forceinline__ void stage(const Lds<A, RL::N * G>& L, int tid, const TW& tw, const double* oddc,
                                      const Prune& pr) {
    constexpr int N = RL::N, r = RL::r[I], s = RL::stride(I), nbf = N / r, W = G * nbf, IT = (W + T - 1) / T;
    const double* tws = tw.template table<A, RL, I>();
    if constexpr (I > 0) {
        if (pr.wlen >= 0 && pr.wlen < s) {
            stage_window<A, RL, G, T, I>(L, tid, tw, tws, oddc, pr.wstart % s, pr.wlen);
            return;
        }
    }
    Cx<A> held[IT][r];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = tid + it * T;
        if (W % T == 0 || w < W) {
            const int g = w % G, b = w / G, p = b / s;
            Cx<A> x[r];
#pragma unroll
            for (int t = 0; t < r; ++t) x[t] = L.get((b + t * nbf) * G + g);
            const int pg = tw.index(g, p);
            butterfly<A, r>(x, held[it], pg != 0, tws, oddc, (unsigned)(pg * (r - 1)));
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = tid + it * T;
        if (W % T == 0 || w < W) {
            const int g = w % G, b = w / G, p = b / s, q = b - p * s;
            const int dbase = s * r * p + q;
#pragma unroll
            for (int u = 0; u < r; ++u) L.put((dbase + u * s) * G + g, held[it][u]);
        }
    }
    __syncthreads();
}

template <class A, class RL, int G, int T, class TW, int... I>
__device__ __forceinline__ void all_stages(const Lds<A, RL::N * G>& L, int tid, const TW& tw, const double* oddc,
                                           const Prune& pr, std::integer_sequence<int, I...>) {
    (stage<A, RL, G, T, I>(L, tid, tw, oddc, pr), ...);
}

// Stages 1 .. nst-2 (the first and last run with their global-memory side, below).
template <class A, class RL, int G, int T, class TW, int... I>
__device__ __forceinline__ void inner_stages(const Lds<A, RL::N * G>& L, int tid, const TW& tw, const double* oddc,
                                             const Prune& pr, std::integer_sequence<int, I...>) {
    (stage<A, RL, G, T, I + 1>(L, tid, tw, oddc, pr), ...);
}

// Stage 0 with its inputs read from global memory: load(e, j) gives element e of sub-FFT j. It has stride 1, so
// butterfly b reads elements b + t (N / r) and writes r b + u; nothing in LDS is read, so it writes its outputs
// straight away, and one sync follows.
template <class A, class RL, int G, int T, class TW, class Jof, class Load>
__device__ __forceinline__ void first_stage(const Lds<A, RL::N * G>& L, int tid, const TW& tw, const double* oddc,
                                            Jof jof, Load load) {
    constexpr int N = RL::N, r = RL::r[0], nbf = N / r, W = G * nbf, IT = (W + T - 1) / T;
    const double* tws = tw.template table<A, RL, 0>();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = tid + it * T;
        if (W % T == 0 || w < W) {
            const int g = w % G, b = w / G, j = jof(g);
            if (j < 0) continue;
            Cx<A> x[r], y[r];
#pragma unroll
            for (int t = 0; t < r; ++t) x[t] = load(b + t * nbf, j);
            const int pg = tw.index(g, b);
            butterfly<A, r>(x, y, pg != 0, tws, oddc, (unsigned)(pg * (r - 1)));
#pragma unroll
            for (int u = 0; u < r; ++u) L.put((r * b + u) * G + g, y[u]);
        }
    }
    __syncthreads();
}

// The last stage with its outputs written to global memory: store(e, j, v) takes element e of sub-FFT j. Its
// stride is N / r, so butterfly b writes elements b + u (N / r); its p is 0, so it has no twiddles unless it is the
// leading part of a longer transform (LongTw). With a window narrower than the stride only the butterflies b in it
// run, as in stage().
template <class A, class RL, int G, int T, class TW, class Jof, class Store>
__device__ __forceinline__ void last_stage(const Lds<A, RL::N * G>& L, int tid, const TW& tw, const double* oddc,
                                           const Prune& pr, Jof jof, Store store) {
    constexpr int I = RL::nst - 1, N = RL::N, r = RL::r[I], s = RL::stride(I), nbf = N / r, W = G * nbf,
                  IT = (W + T - 1) / T;
    static_assert(s == nbf, "the last stage has stride N / r");
    const double* tws = tw.template table<A, RL, I>();
    const bool win = pr.wlen >= 0 && pr.wlen < s;
    const int Wn = win ? G * pr.wlen : W, qs = win ? pr.wstart % s : 0;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = tid + it * T;
        if (win ? w < Wn : (W % T == 0 || w < W)) {
            const int g = w % G, k = w / G, j = jof(g);
            if (j < 0) continue;
            const int b = win ? (qs + k < s ? qs + k : qs + k - s) : k;
            Cx<A> x[r], y[r];
#pragma unroll
            for (int t = 0; t < r; ++t) x[t] = L.get((b + t * nbf) * G + g);
            const int pg = tw.index(g, 0);
            butterfly<A, r>(x, y, pg != 0, tws, oddc, (unsigned)(pg * (r - 1)));
#pragma unroll
            for (int u = 0; u < r; ++u) store(b + u * s, j, y[u]);
        }
    }
}

// Workgroup wg's task and sub-FFT group in a launch of ntask x ngrp: tasks slowest (tgroup <= 0), or in groups of
// tgroup tasks, each group's sub-FFT groups in turn with its tasks fastest.
__device__ __forceinline__ void task_group_order(int wg, int ntask, int ngrp, int tgroup, int& task, int& grp) {
    if (tgroup <= 0) {
        task = wg / ngrp;
        grp = wg - task * ngrp;
        return;
    }
    const int per = tgroup * ngrp, blk = wg / per, t0 = blk * tgroup, gs = min(tgroup, ntask - t0);
    const int r = wg - blk * per;
    grp = r / gs;
    task = t0 + (r - grp * gs);
}

// Workgroup index remapped so that each of the 8 XCDs (which take workgroups round-robin) works on a contiguous
// range of the grid: neighbouring sub-FFT groups share cache lines in that XCD's L2.
__device__ __forceinline__ int xcd_wg() {
    const int n = gridDim.x, i = blockIdx.x, full = n & ~7;
    if (i >= full) return i;
    return (i & 7) * (full >> 3) + (i >> 3);
}

// Arithmetic of the splits and the bin map: the reference's scalar dd (accurate add) or FP64.
template <class A> using SplitPol = std::conditional_t<A::kDD, DDA, F64>;

template <class A> __device__ __forceinline__ Val<A> ldH(const double* H, long long i) {
    if constexpr (A::kDD) {
        const double2 v = reinterpret_cast<const double2*>(H)[i];
        return dd{v.x, v.y};
    } else {
        return H[i];
    }
}

// Slot pairing of fused passes: local sub-FFT g is member (g & 1) of slot grp*G/2 + g/2; member 0 is the slot's
// own index s, member 1 its mirror M - s. Returns the sub-FFT index, or -1 for an empty place.
template <int G> __device__ __forceinline__ int slot_index(int g, int grp, int M) {
    const int slot = grp * (G / 2) + g / 2, mem = g & 1;
    if (2 * slot > M) return -1;
    if (mem && (slot == 0 || 2 * slot == M)) return -1;
    return mem ? M - slot : slot;
}
template <int G> __device__ __forceinline__ int slot_local(int j, int grp, int M) {
    const int mem = (j != 0 && 2 * j > M) ? 1 : 0;
    const int slot = mem ? M - j : j;
    return 2 * (slot - grp * (G / 2)) + mem;
}

// Y[idx] = X[idx] H[idx] with the bin map (src/resampler.cpp:64-78): down folds bins above the output Nyquist onto
// Mb - j, up extends X periodically.
template <class S> __device__ __forceinline__ Cx<S> ybin(const FuseArgs& f, const double* X, int idx) {
    if (!f.up) {
        const Cx<S> x = gld<S>(X, idx);
        const Val<S> hk = ldH<S>(f.H, idx);
        if (idx == f.hinv) return {S::mul(S::mul(S::lift(2.0), x.re), hk), S::lift(0.0)};
        Cx<S> y = {S::mul(x.re, hk), S::mul(x.im, hk)};
        const int j = f.Mb - idx;
        if (j >= f.K && j <= f.jmax) {
            const Cx<S> xj = gld<S>(X, j);
            const Val<S> hj = ldH<S>(f.H, j);
            y = c_add<S>(y, c_conj<S>(Cx<S>{S::mul(xj.re, hj), S::mul(xj.im, hj)}));
        }
        return y;
    }
    const int kk = idx % f.Nb;
    const Cx<S> v = kk <= f.Nb / 2 ? gld<S>(X, kk) : c_conj<S>(gld<S>(X, f.Nb - kk));
    const Val<S> hk = ldH<S>(f.H, idx);
    Cx<S> y = {S::mul(v.re, hk), S::mul(v.im, hk)};
    if (idx == f.hinv) y.im = S::lift(0.0);
    return y;
}

// inverse pre-split of bin k from Y[k] and its partner Y[hinv - k] (src/dsp/fft.hpp:412-458)
template <class S> __device__ __forceinline__ Cx<S> presplit(const FuseArgs& f, int k, Cx<S> Y, Cx<S> Yp) {
    if (!S::kDD || 2 * k <= f.hinv) {   // v-form: A = Y[k], B = conj(Y[hinv-k])
        const Cx<S> B = c_conj<S>(Yp);
        const Cx<S> E = c_add<S>(Y, B), D = c_mul<S>(c_sub<S>(Y, B), c_conj<S>(gld<S>(f.wi, k)));
        return c_conj<S>(Cx<S>{S::sub(E.re, D.im), S::add(E.im, D.re)});
    }
    // dd, upper half: the partner's E and split (A = Y[hinv-k], B = conj(Y[k])) with this bin's twiddle
    const Cx<S> B = c_conj<S>(Y);
    const Cx<S> E = c_add<S>(Yp, B), sp = c_sub<S>(Yp, B);
    const Cx<S> Ej = c_conj<S>(E);
    const Cx<S> Dj = c_mul<S>(Cx<S>{S::neg(sp.re), sp.im}, c_conj<S>(gld<S>(f.wi, k)));
    return c_conj<S>(Cx<S>{S::sub(Ej.re, Dj.im), S::add(Ej.im, Dj.re)});
}

// forward split X[i] from z[i] and z[h - i] (src/dsp/fft.hpp:365-407), 0 < i < h
template <class S> __device__ __forceinline__ Cx<S> fsplit(const FuseArgs& f, int i, Cx<S> zi, Cx<S> zp) {
    const Val<S> half = S::lift(0.5);
    Cx<S> v;
    if (!S::kDD || 2 * i <= f.h) {   // v-form: A = z[i], B = conj(z[h-i])
        const Cx<S> B = c_conj<S>(zp);
        const Cx<S> E = c_add<S>(zi, B), O = c_mul_mi<S>(c_sub<S>(zi, B));
        v = c_add<S>(E, c_mul<S>(O, gld<S>(f.wf, i)));
    } else {                          // dd, upper half: from the partner's E and O (A = z[h-i], B = conj(z[i]))
        const Cx<S> B = c_conj<S>(zi);
        const Cx<S> E = c_add<S>(zp, B), O = c_mul_mi<S>(c_sub<S>(zp, B));
        v = c_add<S>(c_conj<S>(E), c_mul<S>(c_conj<S>(O), gld<S>(f.wf, i)));
    }
    return {S::mul(v.re, half), S::mul(v.im, half)};
}

template <class A> __device__ __forceinline__ Cx<A> renorm_dd(Cx<A> v) {
    if constexpr (A::kDD) return {quick2(v.re.hi, v.re.lo), quick2(v.im.hi, v.im.lo)};
    else return v;
}

// One task's addresses in a pass, formed once per workgroup (the same for every lane, so they stay in scalar
// registers): the task's source and destination, the element strides, and for ST_OUT the channel's output samples.
// Each lane then reaches element e of sub-FFT j by one 32-bit offset from a base: a task's element indices stay
// below 2^27 (launch_fixed checks), so the byte offsets fit 32 bits.
struct PassView {
    const double* src;
    double* dst;
    double* out;          // ST_OUT: sample x of this block's kept range [xlo, xhi) is out[x + xoff]
    int lde, ldg, ste, stg;
    int margP, xlo, xhi, xoff;
    int pairs;            // ST_OUT: (hi, lo) pairs (PassArgs::out_dd), out[x + xoff] being the pair's index
    int zkeep, zh;        // zkeep >= 0: elements zkeep < i < zh - zkeep are not stored (z_unused)
};
template <int LD>
__device__ __forceinline__ PassView pass_view(const PassArgs& a, const TaskIO& io, int task) {
    PassView v;
    v.src = LD == LD_REAL_F64 ? a.src + io.in_off : a.src + task * a.src_tstride;
    v.dst = a.dst + task * a.dst_tstride;
    v.lde = (int)a.lde; v.ldg = (int)a.ldg; v.ste = (int)a.ste; v.stg = (int)a.stg;
    v.zkeep = (int)a.zkeep; v.zh = (int)a.zh;
    // unpacked sample x (x = 2i - margP and x + 1) is kept when 0 <= x < hopP and its output sample
    // d = blk * hopP + x - o lies in [0, nout) (src/resampler.cpp:80-84): x in [xlo, xhi), written at
    // out + out_off + base + (x - (base - shift)) with shift = blk * hopP - o and base = clamp(shift, 0, nout)
    const long long shift = io.blk * a.hopP - a.o;
    const long long lo = std::max(0LL, std::min(a.hopP, -shift)), hi = std::max(0LL, std::min(a.hopP, io.nout - shift));
    const long long base = std::max(0LL, std::min(io.nout, shift));
    v.pairs = a.out_dd;
    v.out = a.out ? a.out + (a.out_dd ? 2 : 1) * (io.out_off + base) : nullptr;
    v.margP = (int)a.margP; v.xlo = (int)lo; v.xhi = (int)hi; v.xoff = (int)(shift - base);
    return v;
}

// Element e of sub-FFT j from a pass's source (real pairs of the padded input, or complex elements).
template <class A, int LD> __device__ __forceinline__ Cx<A> pass_load(const PassView& v, int e, int j) {
    const unsigned i = (unsigned)(e * v.lde + j * v.ldg);
    if constexpr (LD == LD_REAL_F64) {
        const double2 pr = *at32<F64>(v.src, i);   // a real pair: 16 bytes, the size of an FP64 complex element
        return {A::lift(pr.x), A::lift(pr.y)};
    } else {
        return gld32<A>(v.src, i);
    }
}

// Element e of sub-FFT j's result to a pass's destination: four-step twiddle, renormalisation, then the complex
// store (bins no reader needs are skipped) or the overlap discard (samples 2i and 2i + 1 of the block's output,
// src/dsp/fft.hpp:460 and src/resampler.cpp:80-84).
template <class A, int ST>
__device__ __forceinline__ void pass_store(const PassArgs& a, const PassView& w, int e, int j, Cx<A> v) {
    const int i = e * w.ste + j * w.stg;
    if (ST == ST_CPLX && w.zkeep >= 0 && i > w.zkeep && i < w.zh - w.zkeep) return;
    if (a.twmode && e != 0 && (a.twmode == 1 || j != 0)) v = c_mul<A>(v, gld32<A>(a.tw4, (unsigned)i));
    if constexpr (A::kDD) {
        if (a.renorm) v = renorm_dd<A>(v);
    }
    if constexpr (ST == ST_CPLX) {
        gst32<A>(w.dst, (unsigned)i, v);
    } else if constexpr (ST == ST_OUT) {
        const unsigned esz = w.pairs ? 16u : 8u;
        auto at = [&](int x) -> double* {
            return reinterpret_cast<double*>(reinterpret_cast<char*>(w.out) + (size_t)((unsigned)(x + w.xoff) * esz));
        };
        // the sample's dd value rounded to double, or (out_dd) the pair itself
        auto put = [&](int x, Val<A> y) {
            if constexpr (A::kDD) {
                if (w.pairs) {
                    double* p = at(x);
                    p[0] = y.hi;
                    p[1] = y.lo;
                } else {
                    *at(x) = y.hi + y.lo;
                }
            } else {
                *at(x) = y;
            }
        };
        const int x0 = 2 * i - w.margP;
        if (x0 >= w.xlo && x0 < w.xhi) put(x0, v.re);
        const int x1 = x0 + 1;
        if (x1 >= w.xlo && x1 < w.xhi) put(x1, A::neg(v.im));
    }
}

// __launch_bounds__'s second value is the minimum number of waves per SIMD the register allocation must allow:
// 4 caps a thread at 128 VGPRs, so that LDS (not registers) limits occupancy.
// IO: the first stage reads global memory and the last stage writes it (first_stage, last_stage), instead of a
// load into LDS, a sync, the first stage, ..., the last stage, a sync and a store from LDS; same operations.
template <class A, class RL, int G, int T, int LD, int ST, bool IO = false>
__global__ __launch_bounds__(T, 4) void k_fixed(PassArgs a) {
    constexpr int N = RL::N, CAP = N * G;
    constexpr bool kPaired = LD == LD_FOLD || ST == ST_SPLIT;
    static_assert(!IO || (!kPaired && RL::nst >= 2), "global-memory stages need an unpaired pass of two stages");
    using S = SplitPol<A>;
    static_assert(std::is_same_v<Val<S>, Val<A>>);
    extern __shared__ double lds[];   // kPlanes * CAP doubles
    const Lds<A, CAP> L{lds};
    const int tid = threadIdx.x;
    const int wg = xcd_wg();
    int task, grp;
    task_group_order(wg, a.ntask, a.ngrp, a.tgroup, task, grp);
    const int j0 = grp * G;
    const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
    // sub-FFT index of local g, or -1
    auto jof = [&](int g) -> int {
        if constexpr (kPaired) return slot_index<G>(g, grp, a.f.M);
        else return j0 + g < a.nsub ? j0 + g : -1;
    };
    const PassView view = pass_view<LD>(a, io, task);
    auto load_at = [&](int e, int j) -> Cx<A> { return pass_load<A, LD>(view, e, j); };
    auto store_at = [&](int e, int j, Cx<A> v) { pass_store<A, ST>(a, view, e, j, v); };

    if constexpr (LD == LD_FOLD) {
        // Y = X H with the bin map for every element, then the pre-split of each with its mirror bin
        const double* X = a.src + task * a.src_tstride;
        constexpr int IT = (G * N + T - 1) / T;
        Cx<A> yh = {A::lift(0.0), A::lift(0.0)};
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int w = tid + it * T;
            if (w < G * N) {
                const int g = w % G, e = w / G, j = jof(g);
                if (j >= 0) {
                    const int k = e * a.f.M + j;
                    L.put(e * G + g, ybin<S>(a.f, X, k));
                    if (k == 0) yh = ybin<S>(a.f, X, a.f.hinv);
                }
            }
        }
        __syncthreads();
        Cx<A> zs[IT];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int w = tid + it * T;
            if (w < G * N) {
                const int g = w % G, e = w / G, j = jof(g);
                if (j >= 0) {
                    const int k = e * a.f.M + j;
                    Cx<A> yp;
                    if (k == 0) {
                        yp = yh;
                    } else {
                        const int kp = a.f.hinv - k, ep = kp / a.f.M, jp = kp - ep * a.f.M;
                        yp = L.get(ep * G + slot_local<G>(jp, grp, a.f.M));
                    }
                    zs[it] = presplit<S>(a.f, k, L.get(e * G + g), yp);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int w = tid + it * T;
            if (w < G * N) {
                const int g = w % G, e = w / G;
                if (jof(g) >= 0) L.put(e * G + g, zs[it]);
            }
        }
    } else if constexpr (!IO) {
        // load, lanes sub-FFT-fastest: a wave writes one contiguous LDS range (no bank conflicts) and reads 64/G
        // consecutive elements of each of its G sub-FFTs (or G neighbouring columns of 64/G rows)
        for (int w = tid; w < G * N; w += T) {
            const int g = w % G, e = w / G;
            const int j = jof(g);
            if (j < 0) continue;
            L.put(e * G + g, load_at(e, j));
        }
    }
    Prune pr;
    if constexpr (ST == ST_CPLX && !kPaired) pr = Prune{a.wstart, a.wlen};
    const OwnTw tw{a.tw};
    if constexpr (IO) {
        first_stage<A, RL, G, T>(L, tid, tw, a.oddc, jof, load_at);
        inner_stages<A, RL, G, T>(L, tid, tw, a.oddc, pr, std::make_integer_sequence<int, RL::nst - 2>{});
        last_stage<A, RL, G, T>(L, tid, tw, a.oddc, pr, jof, store_at);
        return;
    }
    __syncthreads();
    all_stages<A, RL, G, T>(L, tid, tw, a.oddc, pr, std::make_integer_sequence<int, RL::nst>{});

    if constexpr (ST == ST_SPLIT) {
        // the sub-FFT outputs are z[M*e + j] (natural order); X[i] needs z[i] and z[h - i]
        double* X = a.dst + task * a.dst_tstride;
        for (int w = tid; w < G * N; w += T) {
            const int g = w % G, e = w / G, j = jof(g);
            if (j < 0) continue;
            const int i = e * a.f.M + j;
            const Cx<S> zi = renorm_dd<A>(L.get(e * G + g));
            if (i == 0) {
                gst<S>(X, 0, Cx<S>{S::add(zi.re, zi.im), S::lift(0.0)});
                if (a.f.h <= a.f.xmax) gst<S>(X, a.f.h, Cx<S>{S::sub(zi.re, zi.im), S::lift(0.0)});
                continue;
            }
            if (i > a.f.xmax) continue;
            const int ip = a.f.h - i, ep = ip / a.f.M, jp = ip - ep * a.f.M;
            const Cx<S> zp = renorm_dd<A>(L.get(ep * G + slot_local<G>(jp, grp, a.f.M)));
            gst<S>(X, i, fsplit<S>(a.f, i, zi, zp));
        }
    } else {
        for (int w = tid; w < G * N; w += T) {   // lanes sub-FFT-fastest, as the load
            const int g = w % G, e = w / G;
            const int j = jof(g);
            if (j < 0) continue;
            store_at(e, j, L.get(e * G + g));
        }
    }
}


// Unpaired passes of two or more stages run their first and last stages on global memory (k_fixed's IO).
template <class A, class RL, int G, int T, int LD, int ST,
          bool IO = LD != LD_FOLD && ST != ST_SPLIT && RL::nst >= 2>
hipError_t launch_fixed_one(const PassArgs& a, hipStream_t s) {
    static_assert(RL::hold(G, T) <= (A::kDD ? 8 : 16), "too many values held per thread");
    static_assert(LD != LD_FOLD || (G * RL::N + T - 1) / T <= (A::kDD ? 8 : 16), "too many values held per thread");
    static_assert((LD != LD_FOLD && ST != ST_SPLIT) || G % 2 == 0, "paired passes need an even group");
    constexpr size_t bytes = (size_t)kPlanes<A> * RL::N * G * sizeof(double);
    static_assert(bytes <= 160 * 1024, "LDS tile too large");
    static const hipError_t attr = bytes > 64 * 1024
        ? hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fixed<A, RL, G, T, LD, ST, IO>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
        : hipSuccess;
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL((k_fixed<A, RL, G, T, LD, ST, IO>), dim3(a.ngrp * a.ntask), dim3(T), bytes, s, a);
    return hipGetLastError();
}

template <int P> struct PolicyOf;
template <> struct PolicyOf<POL_F64> { using type = F64; };
template <> struct PolicyOf<POL_DDA> { using type = DDA; };
template <> struct PolicyOf<POL_DDS> { using type = DDS; };

#define KB_FIXED(POL, G, T, LD, ST, ...)                                                                            \
    FixedEntry{POL, Radices<__VA_ARGS__>::N, LD, ST, G, T,                                                              \
          &launch_fixed_one<PolicyOf<POL>::type, Radices<__VA_ARGS__>, G, T, LD, ST>}

// ---- shapes chosen at compile time for lengths without a measured table entry: the radix list in the reference's
// order, and the (G, T) with the best product of lane use over the stages and waves per CU, within 80 KB of LDS and
// the register bound on held values (measured tuning may replace the choice at run time).
constexpr int radix_count(int n) {
    int c = 0;
    while (n % 4 == 0) { n /= 4; ++c; }
    while (n % 2 == 0) { n /= 2; ++c; }
    for (int f : {3, 5, 7})
        while (n % f == 0) { n /= f; ++c; }
    return c;
}
constexpr int radix_at(int n, int i) {
    int c = 0;
    while (n % 4 == 0) { n /= 4; if (c++ == i) return 4; }
    while (n % 2 == 0) { n /= 2; if (c++ == i) return 2; }
    for (int f : {3, 5, 7})
        while (n % f == 0) { n /= f; if (c++ == i) return f; }
    return 0;
}
template <int N, class I> struct RadOf;
template <int N, int... I> struct RadOf<N, std::integer_sequence<int, I...>> {
    using type = Radices<radix_at(N, I)...>;
};
template <int N> using RadicesOf = typename RadOf<N, std::make_integer_sequence<int, radix_count(N)>>::type;

struct Shape {
    int G, T;
};
// 256 threads (4 waves) and the G whose stages waste the fewest wave slots (a wave with no butterfly issues
// nothing; a partly filled one costs a full one), within 80 KB of LDS and the held-value bound; ties go to the
// larger G. Measured with the first and last stages on global memory, 4-wave workgroups beat 8-wave ones of the
// same wave efficiency (320 points: 4 x 256 0.374 ms against 8 x 512 0.403; 189: 8 x 256 0.236 against 8 x 512
// 0.252; 245: 5 x 256 0.215 against 9 x 512 0.229); the rule picks 4 x 256 and 5 x 256 for the first and last,
// 7 x 256 for 189.
constexpr Shape auto_shape(int n, bool dd) {
    const int planes = dd ? 4 : 2, hmax = dd ? 8 : 16, nst = radix_count(n);
    constexpr int T = 256;
    Shape best{1, T};
    long long best_num = -1, best_den = 1;
    for (int G = 1; G <= 32; ++G) {
        const long long lds = (long long)planes * 8 * n * G;
        if (lds > 80 * 1024) break;
        int h = 0;
        long long work = 0, slots = 0;
        for (int i = 0; i < nst; ++i) {
            const int r = radix_at(n, i), W = G * (n / r), it = (W + T - 1) / T, rem = W - (it - 1) * T;
            h = it * r > h ? it * r : h;
            work += W;
            slots += (long long)(it - 1) * (T / 64) + (rem + 63) / 64;
        }
        if (h > hmax) continue;
        // efficiency work / (64 slots), compared exactly; ties (and better) take the larger G
        if (best_num < 0 || work * 64 * best_den >= best_num * 64 * slots) {
            best_num = work;
            best_den = slots;
            best = {G, T};
        }
    }
    return best;
}

#define KB_AUTO(POL, N, LD, ST)                                                                                   \
    FixedEntry{POL, N, LD, ST, auto_shape(N, POL != POL_F64).G, auto_shape(N, POL != POL_F64).T,                   \
               &launch_fixed_one<PolicyOf<POL>::type, RadicesOf<N>, auto_shape(N, POL != POL_F64).G,              \
                                 auto_shape(N, POL != POL_F64).T, LD, ST>}

}  // namespace
}  // namespace kbsrc::gpu
