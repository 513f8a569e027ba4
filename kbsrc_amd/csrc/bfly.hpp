// Device building blocks shared by the sub-FFT kernels: global element access in the policy's element format and
// the reference's radix-2/3/4/5/7 Stockham butterflies (src/dsp/fft.hpp:62-84,123-145) on register arrays.
#pragma once
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {
namespace {

template <class A> using Val = typename A::T;
template <class A> using Cx = cx<typename A::T>;
template <class A> constexpr int kPlanes = A::kDD ? 4 : 2;
template <class A> constexpr int kElt = A::kDD ? 4 : 2;   // doubles per global complex element

template <class A> __device__ __forceinline__ Cx<A> gld(const double* base, long long i) {
    if constexpr (A::kDD) {
        const double2* p = reinterpret_cast<const double2*>(base) + 2 * i;
        const double2 h = p[0], l = p[1];
        return {dd{h.x, l.x}, dd{h.y, l.y}};
    } else {
        const double2 v = reinterpret_cast<const double2*>(base)[i];
        return {v.x, v.y};
    }
}
template <class A> __device__ __forceinline__ void gst(double* base, long long i, Cx<A> v) {
    if constexpr (A::kDD) {
        double2* p = reinterpret_cast<double2*>(base) + 2 * i;
        p[0] = make_double2(v.re.hi, v.im.hi);
        p[1] = make_double2(v.re.lo, v.im.lo);
    } else {
        reinterpret_cast<double2*>(base)[i] = make_double2(v.re, v.im);
    }
}

// The same with a 32-bit element index: the byte offset i * 32 (dd) or i * 16 (FP64) is formed in 32 bits and
// added to the base, so that a base that is the same for every lane stays in scalar registers and each access is one
// 32-bit vector offset from it. Callers keep the byte offset below 2^32.
template <class A> __device__ __forceinline__ const double2* at32(const double* base, unsigned i) {
    return reinterpret_cast<const double2*>(reinterpret_cast<const char*>(base) +
                                            (size_t)(i * (unsigned)(kElt<A> * sizeof(double))));
}
template <class A> __device__ __forceinline__ Cx<A> gld32(const double* base, unsigned i) {
    const double2* p = at32<A>(base, i);
    if constexpr (A::kDD) {
        const double2 h = p[0], l = p[1];
        return {dd{h.x, l.x}, dd{h.y, l.y}};
    } else {
        const double2 v = p[0];
        return {v.x, v.y};
    }
}
template <class A> __device__ __forceinline__ void gst32(double* base, unsigned i, Cx<A> v) {
    double2* p = const_cast<double2*>(at32<A>(base, i));
    if constexpr (A::kDD) {
        p[0] = make_double2(v.re.hi, v.im.hi);
        p[1] = make_double2(v.re.lo, v.im.lo);
    } else {
        p[0] = make_double2(v.re, v.im);
    }
}

// Radix-R butterfly of the reference's Stockham step on x[0..R): out[u] goes to destination q + u*s. Its twiddle
// u - 1 is element twi + u - 1 of tw.
template <class A, int R>
__device__ __forceinline__ void butterfly(const Cx<A> (&x)[R], Cx<A> (&out)[R], bool twid, const double* tw,
                                          const double* oddc, unsigned twi = 0) {
    if constexpr (R == 4) {
        const Cx<A> t0 = c_add<A>(x[0], x[2]), t1 = c_sub<A>(x[0], x[2]), t2 = c_add<A>(x[1], x[3]);
        const Cx<A> t3 = c_mul_mi<A>(c_sub<A>(x[1], x[3]));
        out[0] = c_add<A>(t0, t2);
        out[1] = c_add<A>(t1, t3);
        out[2] = c_sub<A>(t0, t2);
        out[3] = c_sub<A>(t1, t3);
        if (twid) {
            out[1] = c_mul<A>(out[1], gld32<A>(tw, twi));
            out[2] = c_mul<A>(out[2], gld32<A>(tw, twi + 1));
            out[3] = c_mul<A>(out[3], gld32<A>(tw, twi + 2));
        }
    } else if constexpr (R == 2) {
        out[0] = c_add<A>(x[0], x[1]);
        out[1] = c_sub<A>(x[0], x[1]);
        if (twid) out[1] = c_mul<A>(out[1], gld32<A>(tw, twi));
    } else {   // odd radix by symmetric sums (src/dsp/fft.hpp:62-84)
        constexpr int H = (R - 1) / 2;
        constexpr int ri = R == 3 ? 0 : R == 5 ? 1 : 2;
        Cx<A> S[H], D[H];
#pragma unroll
        for (int t = 1; t <= H; ++t) {
            S[t - 1] = c_add<A>(x[t], x[R - t]);
            D[t - 1] = c_sub<A>(x[t], x[R - t]);
        }
        Cx<A> sum = x[0];
#pragma unroll
        for (int t = 0; t < H; ++t) sum = c_add<A>(sum, S[t]);
        out[0] = sum;
#pragma unroll
        for (int u = 1; u <= H; ++u) {
            Cx<A> ac = x[0], bc;
#pragma unroll
            for (int t = 1; t <= H; ++t) {
                const Cx<A> cs = gld<A>(oddc, ri * 7 + (t * u) % R);
                ac = {A::add(ac.re, A::mul(S[t - 1].re, cs.re)), A::add(ac.im, A::mul(S[t - 1].im, cs.re))};
                if (t == 1)   // the sum starts at zero: zero plus the first product (A::add0)
                    bc = {A::add0(A::mul(D[0].re, cs.im)), A::add0(A::mul(D[0].im, cs.im))};
                else
                    bc = {A::add(bc.re, A::mul(D[t - 1].re, cs.im)), A::add(bc.im, A::mul(D[t - 1].im, cs.im))};
            }
            Cx<A> xu = {A::add(ac.re, bc.im), A::sub(ac.im, bc.re)};
            Cx<A> xr = {A::sub(ac.re, bc.im), A::add(ac.im, bc.re)};
            if (twid) {
                xu = c_mul<A>(xu, gld32<A>(tw, twi + u - 1));
                xr = c_mul<A>(xr, gld32<A>(tw, twi + R - u - 1));
            }
            out[u] = xu;
            out[R - u] = xr;
        }
    }
}

}  // namespace
}  // namespace kbsrc::gpu
