// Double-double (hi + lo FP64 pair) arithmetic shared by the host planner and the gfx950 kernels.
//
// Every function is a fixed sequence of IEEE round-to-nearest FP64 operations, so host and device produce the
// same bits. Contraction into FMA must stay off (the build passes -ffp-contract=off and the pragma below makes
// it local to every file that includes this header); the only fused operation is the explicit product-error
// term fma(a, b, -a*b), which is exact and equals the reference's FMA path (src/dsp/double_double.hpp:20-22).
//
// Two additions exist because the reference uses both and they round differently:
//   add_acc  two TwoSums (hi and lo), the scalar dd operator+      (src/dsp/double_double.hpp:67-70)
//   add_slp  one TwoSum on the high parts, the 4-lane DD4 operator+ (src/dsp/simd.hpp:45-50)
// The reference runs its four-step sub-FFTs with add_slp and every other dd step with add_acc.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#define KB_HD __host__ __device__ __forceinline__

namespace kbsrc::gpu {

struct dd {
    double hi, lo;
};
KB_HD dd mkdd(double h) { return dd{h, 0.0}; }

// QuickTwoSum: exact when |a| >= |b| (src/dsp/double_double.hpp:65)
KB_HD dd quick2(double a, double b) {
    double s = a + b;
    return dd{s, b - (s - a)};
}
// Knuth TwoSum: exact for any a, b (src/dsp/double_double.hpp:66)
KB_HD dd knuth2(double a, double b) {
    double s = a + b, v = s - a;
    return dd{s, (a - (s - v)) + (b - v)};
}
KB_HD dd add_acc(dd a, dd b) {
    dd s = knuth2(a.hi, b.hi), t = knuth2(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = quick2(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return quick2(s.hi, s.lo);
}
KB_HD dd add_slp(dd a, dd b) {
    double s = a.hi + b.hi, v = s - a.hi;
    double e = (a.hi - (s - v)) + (b.hi - v);
    e = e + (a.lo + b.lo);
    return quick2(s, e);
}
KB_HD dd ddneg(dd a) { return dd{-a.hi, -a.lo}; }
// product, error term by FMA (src/dsp/double_double.hpp:73-76)
KB_HD dd ddmul(dd a, dd b) {
    double p = a.hi * b.hi, e = __builtin_fma(a.hi, b.hi, -p);
    e = e + (a.hi * b.lo + a.lo * b.hi);
    return quick2(p, e);
}
KB_HD dd ddmul1(dd a, double b) {   // dd x double (src/dsp/double_double.hpp:77-79)
    double p = a.hi * b, e = __builtin_fma(a.hi, b, -p);
    e = e + a.lo * b;
    return quick2(p, e);
}
KB_HD dd sub_acc(dd a, dd b) { return add_acc(a, ddneg(b)); }
// add_slp({0, 0}, b) and add_acc({0, 0}, b) for b a ddmul result, in two operations: ddmul's result is normalised
// (fl(b.hi + b.lo) == b.hi), so both adds return b, except that a negative zero component comes back positive, which
// adding 0.0 to each component reproduces (checked over subnormal, zero and cancelling operands by
// tests/dd_zero_add.cpp).
KB_HD dd zero_add_product(dd b) { return dd{b.hi + 0.0, b.lo + 0.0}; }
KB_HD dd dddiv1(dd a, double b) {   // dd / double, three quotient digits (src/dsp/double_double.hpp:80-84)
    double q1 = a.hi / b;
    dd r = sub_acc(a, ddmul1(mkdd(b), q1));
    double q2 = r.hi / b;
    r = sub_acc(r, ddmul1(mkdd(b), q2));
    double q3 = r.hi / b;
    return add_acc(quick2(q1, q2), mkdd(q3));
}
KB_HD dd dddiv(dd a, dd b) {        // dd / dd long division (src/dsp/double_double.hpp:88-92)
    double q1 = a.hi / b.hi;
    dd r = sub_acc(a, ddmul1(b, q1));
    double q2 = r.hi / b.hi;
    r = sub_acc(r, ddmul1(b, q2));
    double q3 = r.hi / b.hi;
    return add_acc(quick2(q1, q2), mkdd(q3));
}

template <class T> struct cx {
    T re, im;
};

// Arithmetic policies. T is the value type; add/sub/mul are the reference's operations for that type.
struct F64 {   // Fast mode: plain FP64
    using T = double;
    static constexpr bool kDD = false;
    KB_HD static T add(T a, T b) { return a + b; }
    KB_HD static T sub(T a, T b) { return a - b; }
    KB_HD static T mul(T a, T b) { return a * b; }
    KB_HD static T neg(T a) { return -a; }
    KB_HD static T lift(double v) { return v; }
    KB_HD static T from(dd v) { return v.hi; }   // from_dd<double> (src/dsp/fft.hpp:16)
    KB_HD static T add0(T b) { return 0.0 + b; }   // add(lift(0), b) for b a mul result
};
struct DDA {   // dd with the accurate add
    using T = dd;
    static constexpr bool kDD = true;
    KB_HD static T add(T a, T b) { return add_acc(a, b); }
    KB_HD static T sub(T a, T b) { return add_acc(a, ddneg(b)); }
    KB_HD static T mul(T a, T b) { return ddmul(a, b); }
    KB_HD static T neg(T a) { return ddneg(a); }
    KB_HD static T lift(double v) { return mkdd(v); }
    KB_HD static T from(dd v) { return v; }
    KB_HD static T add0(T b) { return zero_add_product(b); }   // add(lift(0), b) for b a mul result
};
struct DDS {   // dd with the one-TwoSum add of the reference's 4-lane sub-FFTs
    using T = dd;
    static constexpr bool kDD = true;
    KB_HD static T add(T a, T b) { return add_slp(a, b); }
    KB_HD static T sub(T a, T b) { return add_slp(a, ddneg(b)); }
    KB_HD static T mul(T a, T b) { return ddmul(a, b); }
    KB_HD static T neg(T a) { return ddneg(a); }
    KB_HD static T lift(double v) { return mkdd(v); }
    KB_HD static T from(dd v) { return v; }
    KB_HD static T add0(T b) { return zero_add_product(b); }   // add(lift(0), b) for b a mul result
};

template <class A> KB_HD cx<typename A::T> c_add(cx<typename A::T> a, cx<typename A::T> b) {
    return {A::add(a.re, b.re), A::add(a.im, b.im)};
}
template <class A> KB_HD cx<typename A::T> c_sub(cx<typename A::T> a, cx<typename A::T> b) {
    return {A::sub(a.re, b.re), A::sub(a.im, b.im)};
}
// a * w = {re*wr - im*wi, re*wi + im*wr} (src/dsp/fft.hpp:46-51)
template <class A> KB_HD cx<typename A::T> c_mul(cx<typename A::T> a, cx<typename A::T> w) {
    return {A::sub(A::mul(a.re, w.re), A::mul(a.im, w.im)), A::add(A::mul(a.re, w.im), A::mul(a.im, w.re))};
}
template <class A> KB_HD cx<typename A::T> c_mul_mi(cx<typename A::T> a) { return {a.im, A::neg(a.re)}; }  // a*(-i)
template <class A> KB_HD cx<typename A::T> c_conj(cx<typename A::T> a) { return {a.re, A::neg(a.im)}; }

}  // namespace kbsrc::gpu
