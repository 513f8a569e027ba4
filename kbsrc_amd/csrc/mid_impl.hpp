// The split / kernel-multiply / bin-map / pre-split step for one pair of inverse-input bins (k, hinv - k), on any
// source of the forward complex FFT's bins: the split/fold kernel reads them from global memory, the fused row
// pass from its LDS tile. One body, so both compute the same operations on the same values.
#pragma once
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {
namespace {

template <class A> using MVal = typename A::T;
template <class A> using MCx = cx<typename A::T>;

// complex element i of a global table (twiddles) in the policy's element format, and real element i (H)
template <class A> __device__ __forceinline__ MCx<A> mid_ldc(const double* base, long long i) {
    if constexpr (A::kDD) {
        const double2* p = reinterpret_cast<const double2*>(base) + 2 * i;
        const double2 h = p[0], l = p[1];
        return {dd{h.x, l.x}, dd{h.y, l.y}};
    } else {
        const double2 v = reinterpret_cast<const double2*>(base)[i];
        return {v.x, v.y};
    }
}
template <class A> __device__ __forceinline__ MVal<A> mid_ldr(const double* base, long long i) {
    if constexpr (A::kDD) {
        const double2 v = reinterpret_cast<const double2*>(base)[i];
        return dd{v.x, v.y};
    } else {
        return base[i];
    }
}
// the same with the byte offset formed in 32 bits (below 4 GiB)
template <class A> __device__ __forceinline__ MCx<A> mid_ldc32(const double* base, unsigned i) {
    const char* b = reinterpret_cast<const char*>(base);
    if constexpr (A::kDD) {
        const double2* p = reinterpret_cast<const double2*>(b + (size_t)(i * 32u));
        const double2 h = p[0], l = p[1];
        return {dd{h.x, l.x}, dd{h.y, l.y}};
    } else {
        const double2 v = *reinterpret_cast<const double2*>(b + (size_t)(i * 16u));
        return {v.x, v.y};
    }
}
template <class A> __device__ __forceinline__ MVal<A> mid_ldr32(const double* base, unsigned i) {
    const char* b = reinterpret_cast<const char*>(base);
    if constexpr (A::kDD) {
        const double2 v = *reinterpret_cast<const double2*>(b + (size_t)(i * 16u));
        return dd{v.x, v.y};
    } else {
        return *reinterpret_cast<const double*>(b + (size_t)(i * 8u));
    }
}

// The step's tables: the kernel spectrum H, the forward split twiddles wf and the inverse pre-split twiddles wi,
// entry i at i (natural order), or residue-major for a workgroup that reads only the residues +-r mod T (the fused
// row pass): entry i at (i mod T) L + i / T, so that the lanes of one residue read neighbouring addresses.
struct MidTabs {
    const double* H;
    const double* wf;
    const double* wi;
    template <class A> __device__ __forceinline__ MVal<A> kern(int i) const { return mid_ldr<A>(H, i); }
    template <class A> __device__ __forceinline__ MCx<A> fwd(int i) const { return mid_ldc<A>(wf, i); }
    template <class A> __device__ __forceinline__ MCx<A> inv(int i) const { return mid_ldc<A>(wi, i); }
};
// floor(u / D) for 0 <= u <= kDivSmallMax by one FP32 product and a truncation, no integer multiply: (u + 1/2) (1/D)
// rounded twice (1/D, then the fma) is within (u + 1/2) 2^-23 / D of (u + 1/2) / D, less than the (1/2) / D between
// that value and either integer around it while u + 1/2 < 2^22.
constexpr unsigned kDivSmallMax = (1u << 22) - 2;
template <unsigned D> __device__ __forceinline__ unsigned div_small(unsigned u) {
    constexpr float inv = 1.0f / D;
    return (unsigned)__builtin_fmaf((float)u, inv, 0.5f * inv);
}

// The residue-major form, L entries per residue (a compile-time count): positions and byte offsets in 32 bits, the
// index arithmetic on 24-bit products (indices are non-negative and at most kDivSmallMax, each table below 4 GiB:
// the launcher checks both), so that the table base stays in scalar registers.
template <int T, int L> struct MidTabsRes {
    const double* H;
    const double* wf;
    const double* wi;
    __device__ __forceinline__ unsigned pos(int i) const {
        const unsigned u = (unsigned)i, q = div_small<T>(u);
        return __umul24(u - __umul24(q, T), L) + q;
    }
    template <class A> __device__ __forceinline__ MVal<A> kern(int i) const {
        return mid_ldr32<A>(H, pos(i));
    }
    template <class A> __device__ __forceinline__ MCx<A> fwd(int i) const { return mid_ldc32<A>(wf, pos(i)); }
    template <class A> __device__ __forceinline__ MCx<A> inv(int i) const { return mid_ldc32<A>(wi, pos(i)); }
};

// ---- forward split X[i] from the complex FFT output, z(k) = bin k (src/dsp/fft.hpp:365-407)
template <class A, class Z, class TB>
__device__ __forceinline__ MCx<A> fwd_split_at(Z z, const TB& tb, int h, int i) {
    using Cx = MCx<A>;
    if (i == 0 || i == h) {
        const Cx z0 = z(0);
        return {i == 0 ? A::add(z0.re, z0.im) : A::sub(z0.re, z0.im), A::lift(0.0)};
    }
    const MVal<A> half = A::lift(0.5);
    if constexpr (!A::kDD) {   // FP64: every bin from its own pair (src/dsp/fft.hpp:366-373)
        const Cx Az = z(i), Bz = c_conj<A>(z(h - i));
        const Cx E = c_add<A>(Az, Bz), O = c_mul_mi<A>(c_sub<A>(Az, Bz));
        const Cx v = c_add<A>(E, c_mul<A>(O, tb.template fwd<A>(i)));
        return {A::mul(v.re, half), A::mul(v.im, half)};
    } else {                   // dd: bins above h/2 from their conjugate partner's E and O (src/dsp/fft.hpp:397-407)
        const bool lower = i <= h / 2;
        const int k = lower ? i : h - i, j = h - k;
        const Cx Az = z(k), Bz = c_conj<A>(z(j));
        const Cx E = c_add<A>(Az, Bz), O = c_mul_mi<A>(c_sub<A>(Az, Bz));
        const Cx v = lower ? c_add<A>(E, c_mul<A>(O, tb.template fwd<A>(k)))
                           : c_add<A>(c_conj<A>(E), c_mul<A>(c_conj<A>(O), tb.template fwd<A>(j)));
        return {A::mul(v.re, half), A::mul(v.im, half)};
    }
}

// spectrum bin idx in [0, hinv] after the kernel multiply and the bin map (src/resampler.cpp:64-78)
// general fold (src/resampler.cpp:53-63): Y[idx] sums X~[j mod Nb] H[j] over every j <= Jg with j = +-idx mod Mb, in
// ascending j, so each output bin accumulates in the reference's order without atomics
template <class A, class Z, class TB>
__device__ __forceinline__ MCx<A> gen_term_at(const MidArgs& a, Z z, const TB& tb, int j) {
    const int jj = j % a.Nb;
    const MCx<A> x = jj <= a.Nb / 2 ? fwd_split_at<A>(z, tb, a.h, jj)
                                    : c_conj<A>(fwd_split_at<A>(z, tb, a.h, a.Nb - jj));
    const MVal<A> g = tb.template kern<A>(j);
    return {A::mul(x.re, g), A::mul(x.im, g)};
}
template <class A, class Z, class TB>
__device__ __forceinline__ MCx<A> gen_bin_at(const MidArgs& a, Z z, const TB& tb, int idx) {
    MCx<A> y = {A::lift(0.0), A::lift(0.0)};
    if (idx == 0 || 2 * idx == a.Mb) {   // k == 0 or k == Mb/2: real part doubled (j = 0 itself adds whole)
        for (int j = idx; j <= a.Jg; j += a.Mb) {
            const MCx<A> v = gen_term_at<A>(a, z, tb, j);
            if (j == 0) y = c_add<A>(y, v);
            else y.re = A::add(y.re, A::mul(A::lift(2.0), v.re));
        }
        y.im = A::lift(0.0);
        return y;
    }
    for (int t = 0;; ++t) {   // j = idx + t Mb adds v, j = Mb - idx + t Mb adds conj(v)
        const int j1 = idx + t * a.Mb;
        if (j1 > a.Jg) break;
        y = c_add<A>(y, gen_term_at<A>(a, z, tb, j1));
        const int j2 = a.Mb - idx + t * a.Mb;
        if (j2 > a.Jg) break;
        y = c_add<A>(y, c_conj<A>(gen_term_at<A>(a, z, tb, j2)));
    }
    return y;
}

template <class A, class Z, class TB>
__device__ __forceinline__ MCx<A> mid_bin_at(const MidArgs& a, Z z, const TB& tb, int idx) {
    using Cx = MCx<A>;
    if (a.Jg) return gen_bin_at<A>(a, z, tb, idx);
    if (!a.up) {
        if (idx == a.hinv) {   // output Nyquist bin: twice the real part (src/resampler.cpp:71)
            const Cx x = fwd_split_at<A>(z, tb, a.h, idx);
            return {A::mul(A::mul(A::lift(2.0), x.re), tb.template kern<A>(idx)), A::lift(0.0)};
        }
        const Cx x = fwd_split_at<A>(z, tb, a.h, idx);
        const MVal<A> hk = tb.template kern<A>(idx);
        Cx y = {A::mul(x.re, hk), A::mul(x.im, hk)};
        const int j = a.Mb - idx;   // bins above the output Nyquist fold onto Mb - j
        if (j >= a.K && j <= a.jmax) {
            const Cx xj = fwd_split_at<A>(z, tb, a.h, j);
            const MVal<A> hj = tb.template kern<A>(j);
            const Cx v = {A::mul(xj.re, hj), A::mul(xj.im, hj)};
            y = c_add<A>(y, c_conj<A>(v));
        }
        return y;
    }
    const int kk = idx % a.Nb;   // upsampling: periodic extension of the input spectrum
    const Cx v = kk <= a.Nb / 2 ? fwd_split_at<A>(z, tb, a.h, kk) : c_conj<A>(fwd_split_at<A>(z, tb, a.h, a.Nb - kk));
    const MVal<A> hk = tb.template kern<A>(idx);
    Cx y = {A::mul(v.re, hk), A::mul(v.im, hk)};
    if (idx == a.hinv) y.im = A::lift(0.0);
    return y;
}

// inverse pre-split (src/dsp/fft.hpp:412-458): v-form for bin k with partner j = hinv - k
template <class A> __device__ __forceinline__ MCx<A> pre_v(MCx<A> Y, MCx<A> Yp, MCx<A> w) {
    const MCx<A> B = c_conj<A>(Yp);
    const MCx<A> E = c_add<A>(Y, B), D = c_mul<A>(c_sub<A>(Y, B), c_conj<A>(w));
    return c_conj<A>(MCx<A>{A::sub(E.re, D.im), A::add(E.im, D.re)});
}

// Inverse-input bins k (0 <= k <= hinv / 2) and j = hinv - k: put(i, v) receives bin i's value.
template <class A, class Z, class TB, class Put>
__device__ __forceinline__ void mid_pair(const MidArgs& a, Z z, const TB& tb, int k, Put put) {
    using Cx = MCx<A>;
    const int j = a.hinv - k;
    const Cx Yk = mid_bin_at<A>(a, z, tb, k), Yj = mid_bin_at<A>(a, z, tb, j);
    put(k, pre_v<A>(Yk, Yj, tb.template inv<A>(k)));
    if (k && j != k) {
        if constexpr (!A::kDD) {
            put(j, pre_v<A>(Yj, Yk, tb.template inv<A>(j)));
        } else {   // dd: the partner from the same E and split (src/dsp/fft.hpp:452-456)
            const Cx B = c_conj<A>(Yj);
            const Cx E = c_add<A>(Yk, B), sp = c_sub<A>(Yk, B);
            const Cx Ej = c_conj<A>(E);
            const Cx Dj = c_mul<A>(Cx{A::neg(sp.re), sp.im}, c_conj<A>(tb.template inv<A>(j)));
            put(j, c_conj<A>(Cx{A::sub(Ej.re, Dj.im), A::add(Ej.im, Dj.re)}));
        }
    }
}

}  // namespace
}  // namespace kbsrc::gpu
