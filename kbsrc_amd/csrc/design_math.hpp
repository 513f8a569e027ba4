// Scalar double-double functions of the kernel design, compiled for the host (twiddle tables) and the GPU (taps):
// one source, so both give the same bits. FP64 only: IEEE division and square root, explicit FMA, no contraction.
#pragma once
#include "dd.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {

// pi as a dd (src/dsp/double_double.hpp:110-111)
#define KB_PI_DD (dd{3.141592653589793116e+00, 1.224646799147353207e-16})

// cos and sin of x in [0, pi/4] by their Taylor series in dd (src/dsp/double_double.hpp:113-121)
KB_HD void sincos_taylor(dd x, dd& c, dd& s) {
    const dd x2 = ddmul(x, x);
    dd ts = x, sum_s = x, tc = mkdd(1.0), sum_c = mkdd(1.0);
    for (int k = 1; k < 40; ++k) {
        tc = dddiv1(ddneg(ddmul(tc, x2)), double((2 * k - 1) * (2 * k)));
        sum_c = add_acc(sum_c, tc);
        ts = dddiv1(ddneg(ddmul(ts, x2)), double((2 * k) * (2 * k + 1)));
        sum_s = add_acc(sum_s, ts);
        if (__builtin_fabs(ts.hi) < 1e-40 && __builtin_fabs(tc.hi) < 1e-40) break;
    }
    c = sum_c;
    s = sum_s;
}

// e^{+i 2 pi k / N}: exact octant reduction of 8k / N, then the series on the remainder
// (src/dsp/double_double.hpp:122-135)
KB_HD void cis(long long k, long long N, dd& c, dd& s) {
    k %= N;
    if (k < 0) k += N;
    const long long eight = 8 * k, oct = eight / N, rem = eight - oct * N;
    const bool mirrored = oct & 1;
    long long quad = oct / 2;
    const dd pi = KB_PI_DD;
    const dd quarter_pi = {pi.hi / 4, pi.lo / 4};
    const dd phi = dddiv1(ddmul1(quarter_pi, double(mirrored ? (N - rem) : rem)), double(N));
    dd cv, sv;
    sincos_taylor(phi, cv, sv);
    if (mirrored) {
        sv = ddneg(sv);
        quad += 1;
    }
    switch (quad & 3) {
        case 0: c = cv; s = sv; break;
        case 1: c = ddneg(sv); s = cv; break;
        case 2: c = ddneg(cv); s = ddneg(sv); break;
        default: c = sv; s = ddneg(cv); break;
    }
}

// cis(2 pi (jB + l) / N) from the table values cis(2 pi jB / N) = (ac, as) and cis(2 pi l / N) = (bc, bs): one dd
// complex product (src/dsp/fft.hpp:36-42)
KB_HD void cis_compose(dd ac, dd as, dd bc, dd bs, dd& c, dd& s) {
    c = sub_acc(ddmul(ac, bc), ddmul(as, bs));
    s = add_acc(ddmul(ac, bs), ddmul(as, bc));
}

// sqrt of a dd: one Newton step from the FP64 root (src/dsp/double_double.hpp:94-98)
KB_HD dd dd_sqrt(dd a) {
    if (a.hi <= 0) return mkdd(0.0);
    const double y = __builtin_sqrt(a.hi);
    const dd r = sub_acc(a, ddmul1(mkdd(y), y));
    return quick2(y, r.hi / (2 * y));
}

// I0(x) = sum (x^2/4)^k / (k!)^2 (src/dsp/double_double.hpp:100-107)
KB_HD dd bessel_i0(dd x) {
    const dd q = dddiv1(ddmul(x, x), 4.0);
    dd term = mkdd(1.0), sum = mkdd(1.0);
    for (int k = 1; k < 2000; ++k) {
        term = dddiv1(ddmul(term, q), double(k) * double(k));
        sum = add_acc(sum, term);
        if (term.hi < sum.hi * 1e-34) break;
    }
    return sum;
}

// Tap n of the one-sided Kaiser-windowed sinc with relative cutoff a/b over L taps (src/design.cpp:17-37):
// ratio * sinc(pi n a / b) * I0(beta sqrt(1 - (n/L)^2)) / I0(beta). The sine's angle is reduced exactly as
// (n a mod 2b) / 2b turns; ratio = a/b and i0_beta = I0(beta) are the caller's (computed once).
KB_HD dd kaiser_tap(long long n, long long L, long long a, long long b, double beta, dd ratio, dd i0_beta) {
    dd sinc = mkdd(1.0);
    if (n) {
        const long long m = (long long)(((__int128)n * a) % (2 * (__int128)b));
        dd c, s;
        cis(m, 2 * b, c, s);
        const dd angle = dddiv1(ddmul(KB_PI_DD, ddmul1(mkdd((double)n), (double)a)), (double)b);
        sinc = dddiv(s, angle);
    }
    const dd u = dddiv1(mkdd((double)n), (double)L);
    const dd win = dddiv(bessel_i0(ddmul(mkdd(beta), dd_sqrt(sub_acc(mkdd(1.0), ddmul(u, u))))), i0_beta);
    return ddmul(ddmul(ratio, sinc), win);
}

}  // namespace kbsrc::gpu
