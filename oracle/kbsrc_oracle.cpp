// Oracle: a plain, scalar CPU restatement of KBSRC 1.3.0's resampling path, used ONLY as a checker by tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py. Nothing in the product links or calls it.
//
// It restates, operation for operation, what the reference computes, so that its output is bit-identical to
// the reference's (pinned against the compiled reference under oracle/_ref and the reference's SHA-256 pins):
//   double-double arithmetic          src/dsp/double_double.hpp:20-135, src/dsp/simd.hpp:41-60,88-128
//   twiddles (dd_cis, CisTable)       src/dsp/double_double.hpp:113-135, src/dsp/fft.hpp:29-43
//   Stockham sub-FFT                  src/dsp/fft.hpp:62-151
//   four-step / Bluestein CFFT        src/dsp/fft.hpp:215-314
//   real FFT splits                   src/dsp/fft.hpp:316-462
//   kernel taps and spectrum          src/design.cpp:17-49
//   Burg edge extension               src/lpc.cpp:13-41
//   planning, staging, engine         src/resampler.cpp:34-228
// The code is written from the algorithm (scalar, one value at a time), not from the reference's SIMD code:
// a reference DD4 lane operation is applied here to one scalar dd with the same IEEE operations in the same
// order. Build: oracle/Makefile (g++ -ffp-contract=off, FMA required for the product-error term).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------------------------------------------------
// double-double primitives
struct DD { double hi, lo; };
static inline DD mk(double h) { return {h, 0.0}; }
// QuickTwoSum (double_double.hpp:65)
static inline DD quick2(double a, double b) { double s = a + b; return {s, b - (s - a)}; }
// Knuth TwoSum (double_double.hpp:66)
static inline DD knuth2(double a, double b) {
    double s = a + b, v = s - a;
    return {s, (a - (s - v)) + (b - v)};
}
// accurate ("two-two-sum") add, scalar dd operator+ (double_double.hpp:67-70)
static inline DD add_a(DD a, DD b) {
    DD s = knuth2(a.hi, b.hi), t = knuth2(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = quick2(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return quick2(s.hi, s.lo);
}
// "sloppy" add of the 4-lane DD4 type (simd.hpp:45-50 / 88-94): one TwoSum on the high parts
static inline DD add_s(DD a, DD b) {
    double s = a.hi + b.hi, v = s - a.hi;
    double e = (a.hi - (s - v)) + (b.hi - v);
    e = e + (a.lo + b.lo);
    return quick2(s, e);
}
static inline DD neg(DD a) { return {-a.hi, -a.lo}; }
// product with the FMA error term (double_double.hpp:73-79; simd.hpp:56-60)
static inline DD mul(DD a, DD b) {
    double p = a.hi * b.hi, e = __builtin_fma(a.hi, b.hi, -p);
    e = e + (a.hi * b.lo + a.lo * b.hi);
    return quick2(p, e);
}
static inline DD mul1(DD a, double b) {   // dd * double (double_double.hpp:77-79)
    double p = a.hi * b, e = __builtin_fma(a.hi, b, -p);
    e = e + a.lo * b;
    return quick2(p, e);
}
static inline DD sub_a(DD a, DD b) { return add_a(a, neg(b)); }
static inline DD div1(DD a, double b) {   // dd / double, three correction steps (double_double.hpp:80-84)
    double q1 = a.hi / b;
    DD r = sub_a(a, mul1(mk(b), q1));
    double q2 = r.hi / b;
    r = sub_a(r, mul1(mk(b), q2));
    double q3 = r.hi / b;
    return add_a(quick2(q1, q2), mk(q3));
}
static inline DD div(DD a, DD b) {        // dd / dd long division (double_double.hpp:88-92)
    double q1 = a.hi / b.hi;
    DD r = sub_a(a, mul1(b, q1));
    double q2 = r.hi / b.hi;
    r = sub_a(r, mul1(b, q2));
    double q3 = r.hi / b.hi;
    return add_a(quick2(q1, q2), mk(q3));
}
static inline DD sqrt_dd(DD a) {          // one Newton step (double_double.hpp:94-98)
    if (a.hi <= 0) return mk(0.0);
    double y = std::sqrt(a.hi);
    DD r = sub_a(a, mul1(mk(y), y));
    return quick2(y, r.hi / (2 * y));
}
static DD bessel_i0(DD x) {               // power series (double_double.hpp:100-107)
    DD q = div1(mul(x, x), 4.0), t = mk(1.0), s = mk(1.0);
    for (int k = 1; k < 2000; k++) {
        t = div1(mul(t, q), double(k) * double(k));
        s = add_a(s, t);
        if (t.hi < s.hi * 1e-34) break;
    }
    return s;
}
static const DD PI_DD = {3.141592653589793116e+00, 1.224646799147353207e-16};

// cos/sin on [0, pi/4] by Taylor series in dd (double_double.hpp:113-121)
static void sincos_small(DD x, DD& c, DD& s) {
    DD x2 = mul(x, x), t = x, ss = x, cc = mk(1.0), tc = mk(1.0);
    for (int k = 1; k < 40; k++) {
        tc = div1(neg(mul(tc, x2)), double((2 * k - 1) * (2 * k)));
        cc = add_a(cc, tc);
        t = div1(neg(mul(t, x2)), double((2 * k) * (2 * k + 1)));
        ss = add_a(ss, t);
        if (std::fabs(t.hi) < 1e-40 && std::fabs(tc.hi) < 1e-40) break;
    }
    c = cc; s = ss;
}
// e^{+i 2 pi k / N} with exact octant reduction (double_double.hpp:122-135)
static void cis(long long k, long long N, DD& c, DD& s) {
    k %= N; if (k < 0) k += N;
    long long m = 8 * k, oct = m / N, rem = m - oct * N;
    bool odd = oct & 1; long long q = oct / 2;
    DD phi = div1(mul1(DD{PI_DD.hi / 4, PI_DD.lo / 4}, double(odd ? (N - rem) : rem)), double(N));
    DD cc, sn; sincos_small(phi, cc, sn);
    if (odd) { sn = neg(sn); q += 1; }
    switch (q & 3) {
        case 0: c = cc; s = sn; break;
        case 1: c = neg(sn); s = cc; break;
        case 2: c = neg(cc); s = neg(sn); break;
        default: c = sn; s = neg(cc); break;
    }
}
// table of cis(2 pi k / N) built from ~2 sqrt(N) series values and one dd complex product (fft.hpp:29-43)
struct CisTab {
    long long N = 0, B = 1; std::vector<DD> ac, as, bc, bs;
    explicit CisTab(long long n) : N(n) {
        B = (long long)std::ceil(std::sqrt((double)n));
        long long J = n / B + 1;
        ac.resize(J); as.resize(J); bc.resize(B); bs.resize(B);
        for (long long j = 0; j < J; j++) cis(j * B, n, ac[j], as[j]);
        for (long long l = 0; l < B; l++) cis(l, n, bc[l], bs[l]);
    }
    void get(long long k, DD& c, DD& s) const {
        k %= N; if (k < 0) k += N;
        long long j = k / B, l = k % B;
        if (l == 0) { c = ac[j]; s = as[j]; return; }
        if (j == 0) { c = bc[l]; s = bs[l]; return; }
        c = sub_a(mul(ac[j], bc[l]), mul(as[j], bs[l]));
        s = add_a(mul(ac[j], bs[l]), mul(as[j], bc[l]));
    }
};

// ---------------------------------------------------------------------------------------------------------
// Arithmetic policies: one scalar value type plus its add. F64 = fast mode; DDS = dd with the DD4 (sloppy)
// add, used inside the four-step sub-FFTs of Exact; DDA = dd with the scalar (accurate) add, used everywhere
// else in dd. Multiplication is the same for both dd flavours.
struct F64 {
    using T = double;
    static T add(T a, T b) { return a + b; }
    static T sub(T a, T b) { return a - b; }
    static T mul(T a, T b) { return a * b; }
    static T ng(T a) { return -a; }
    static T from(DD v) { return v.hi; }
    static T from_d(double v) { return v; }
};
struct DDA {
    using T = DD;
    static T add(T a, T b) { return add_a(a, b); }
    static T sub(T a, T b) { return add_a(a, neg(b)); }
    static T mul(T a, T b) { return oracle::mul(a, b); }
    static T ng(T a) { return neg(a); }
    static T from(DD v) { return v; }
    static T from_d(double v) { return mk(v); }
};
struct DDS {
    using T = DD;
    static T add(T a, T b) { return add_s(a, b); }
    static T sub(T a, T b) { return add_s(a, neg(b)); }
    static T mul(T a, T b) { return oracle::mul(a, b); }
    static T ng(T a) { return neg(a); }
    static T from(DD v) { return v; }
    static T from_d(double v) { return mk(v); }
};

template <class T> struct Cx { T re, im; };
template <class A> static inline Cx<typename A::T> cadd(Cx<typename A::T> a, Cx<typename A::T> b) {
    return {A::add(a.re, b.re), A::add(a.im, b.im)};
}
template <class A> static inline Cx<typename A::T> csub(Cx<typename A::T> a, Cx<typename A::T> b) {
    return {A::sub(a.re, b.re), A::sub(a.im, b.im)};
}
// complex product a * w (fft.hpp:46-51): {re*wr - im*wi, re*wi + im*wr}
template <class A> static inline Cx<typename A::T> cmul(Cx<typename A::T> a, Cx<typename A::T> w) {
    return {A::sub(A::mul(a.re, w.re), A::mul(a.im, w.im)), A::add(A::mul(a.re, w.im), A::mul(a.im, w.re))};
}
template <class T> static inline Cx<T> mul_mi(Cx<T> a, T (*ng)(T)) { return {a.im, ng(a.re)}; }
template <class A> static inline Cx<typename A::T> conj(Cx<typename A::T> a) { return {a.re, A::ng(a.im)}; }

static bool smooth7(long long n) {
    if (n < 1) return false;
    for (int p : {2, 3, 5, 7}) while (n % p == 0) n /= p;
    return n == 1;
}
static long long next_smooth7(long long n) { while (!smooth7(n)) n++; return n; }

// ---------------------------------------------------------------------------------------------------------
// Stockham sub-FFT, exponent sign -1 (fft.hpp:86-151). Radices: all 4s, then one 2, then 3s, 5s, 7s.
template <class A> struct Stockham {
    using T = typename A::T; using C = Cx<T>;
    int n = 0;
    struct Stage { int r, m, s; std::vector<C> tw; };
    std::vector<Stage> st;
    T cr[3][7], sr[3][7];
    void init(int n_) {
        n = n_; st.clear();
        std::vector<int> rs; int t = n;
        while (t % 4 == 0) { rs.push_back(4); t /= 4; }
        while (t % 2 == 0) { rs.push_back(2); t /= 2; }
        for (int p : {3, 5, 7}) while (t % p == 0) { rs.push_back(p); t /= p; }
        if (t != 1) throw std::runtime_error("FFT length not 7-smooth");
        CisTab ct(n);
        int rem = n, s = 1;
        for (int r : rs) {
            Stage g; g.r = r; g.m = rem / r; g.s = s; g.tw.resize((size_t)g.m * (r - 1));
            for (int p = 0; p < g.m; p++)
                for (int u = 1; u < r; u++) {
                    DD c, sn; ct.get(-(long long)s * p * u, c, sn);
                    g.tw[(size_t)p * (r - 1) + u - 1] = {A::from(c), A::from(sn)};
                }
            st.push_back(g); rem /= r; s *= r;
        }
        const int rr[3] = {3, 5, 7};
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < rr[i]; k++) { DD c, sn; cis(k, rr[i], c, sn); cr[i][k] = A::from(c); sr[i][k] = A::from(sn); }
    }
    template <int R>
    void odd(const C* src, size_t sm, C* dst, int s, int q, const T* c_, const T* s_, const C* tw, bool p0) const {
        constexpr int H = (R - 1) / 2;
        C a0 = src[q], S[H], D[H];
        for (int t = 1; t <= H; t++) {
            C x = src[q + t * sm], y = src[q + (R - t) * sm];
            S[t - 1] = cadd<A>(x, y); D[t - 1] = csub<A>(x, y);
        }
        C sum = a0;
        for (int t = 0; t < H; t++) sum = cadd<A>(sum, S[t]);
        dst[q] = sum;
        for (int u = 1; u <= H; u++) {
            C a = a0, b = {A::from_d(0.0), A::from_d(0.0)};
            for (int t = 1; t <= H; t++) {
                int mi = (t * u) % R;
                a = {A::add(a.re, A::mul(S[t - 1].re, c_[mi])), A::add(a.im, A::mul(S[t - 1].im, c_[mi]))};
                b = {A::add(b.re, A::mul(D[t - 1].re, s_[mi])), A::add(b.im, A::mul(D[t - 1].im, s_[mi]))};
            }
            C xu = {A::add(a.re, b.im), A::sub(a.im, b.re)}, xr = {A::sub(a.re, b.im), A::add(a.im, b.re)};
            if (p0) { dst[q + u * s] = xu; dst[q + (R - u) * s] = xr; }
            else { dst[q + u * s] = cmul<A>(xu, tw[u - 1]); dst[q + (R - u) * s] = cmul<A>(xr, tw[R - u - 1]); }
        }
    }
    C* run(C* x, C* y) const {
        C* a = x; C* b = y;
        for (const Stage& g : st) {
            const int r = g.r, m = g.m, s = g.s;
            const size_t sm = (size_t)s * m;
            for (int p = 0; p < m; p++) {
                const C* tw = &g.tw[(size_t)p * (r - 1)];
                const C* src = a + (size_t)s * p; C* dst = b + (size_t)s * r * p;
                for (int q = 0; q < s; q++) {
                    if (r == 4) {
                        C a0 = src[q], a1 = src[q + sm], a2 = src[q + 2 * sm], a3 = src[q + 3 * sm];
                        C t0 = cadd<A>(a0, a2), t1 = csub<A>(a0, a2), t2 = cadd<A>(a1, a3);
                        C t3 = mul_mi<T>(csub<A>(a1, a3), A::ng);
                        dst[q] = cadd<A>(t0, t2);
                        C o1 = cadd<A>(t1, t3), o2 = csub<A>(t0, t2), o3 = csub<A>(t1, t3);
                        if (p == 0) { dst[q + s] = o1; dst[q + 2 * s] = o2; dst[q + 3 * s] = o3; }
                        else {
                            dst[q + s] = cmul<A>(o1, tw[0]); dst[q + 2 * s] = cmul<A>(o2, tw[1]);
                            dst[q + 3 * s] = cmul<A>(o3, tw[2]);
                        }
                    } else if (r == 2) {
                        C a0 = src[q], a1 = src[q + sm];
                        dst[q] = cadd<A>(a0, a1);
                        C d = csub<A>(a0, a1);
                        dst[q + s] = p ? cmul<A>(d, tw[0]) : d;
                    } else if (r == 3) odd<3>(src, sm, dst, s, q, cr[0], sr[0], tw, p == 0);
                    else if (r == 5) odd<5>(src, sm, dst, s, q, cr[1], sr[1], tw, p == 0);
                    else odd<7>(src, sm, dst, s, q, cr[2], sr[2], tw, p == 0);
                }
            }
            std::swap(a, b);
        }
        return a;
    }
};

// ---------------------------------------------------------------------------------------------------------
// Complex FFT (fft.hpp:215-314). S: storage/outer arithmetic (DDA or F64). Sub-FFT arithmetic Sub and the
// "lane" flag reproduce the reference's choice of compute type:
//   lane = true  (C = DD4 or V4): sub-FFTs in Sub (DDS for dd, F64 for double), the four-step twiddle is
//                 applied to every k1 != 0 including column 0, and dd results are renormalised by QuickTwoSum
//                 when stored (Lane<dd, DD4>::unpack, fft.hpp:200-204);
//   lane = false (C = dd, kernel design): only instantiated with Sub == S (DDA); the twiddle skips column 0
//                 and values are stored as computed.
// Lengths up to 2048, or with no divisor N1 >= 8 below sqrt(n), run one scalar Stockham in S.
template <class S, class Sub> struct CFFT {
    using T = typename S::T; using Cs = Cx<T>;
    int n = 0, N1 = 0, N2 = 0; bool four = false, lane = true;
    Stockham<Sub> s1, s2; Stockham<S> sd; std::vector<Cs> tw;
    int m = 0; std::vector<CFFT> inner; std::vector<Cs> bc, Vf;
    void init(int n_, bool lane_) {
        n = n_; lane = lane_; m = 0; inner.clear();
        if (!smooth7(n)) {   // Bluestein (fft.hpp:228-240)
            m = (int)next_smooth7(2LL * n - 1);
            inner.resize(1); inner[0].init(m, lane);
            bc.resize(n);
            std::vector<Cx<DD>> v(m, Cx<DD>{mk(0.0), mk(0.0)}), w(m);
            for (long long t = 0; t < n; t++) {
                long long e = (long long)((__int128)t * t % (2LL * n));
                DD c, sn; cis(e, 2LL * n, c, sn);
                bc[t] = {S::from(c), S::from(sn)};
                v[t] = {c, sn}; if (t) v[m - t] = {c, sn};
            }
            CFFT<DDA, DDA> fd; fd.init(m, false);
            Cx<DD>* V = fd.run(v.data(), w.data());
            Vf.resize(m);
            for (int k = 0; k < m; k++) Vf[k] = {S::from(div1(V[k].re, (double)m)), S::from(div1(V[k].im, (double)m))};
            return;
        }
        four = n > 2048;
        if (!four) { sd.init(n); return; }
        int best = 1;
        for (int d = 1; (long long)d * d <= n; d++) if (n % d == 0) best = d;
        N1 = best; N2 = n / best;
        if (N1 < 8) { four = false; sd.init(n); return; }
        s1.init(N1); s2.init(N2); tw.resize((size_t)n);
        CisTab ct(n);
        for (int k1 = 0; k1 < N1; k1++)
            for (int n2 = 0; n2 < N2; n2++) {
                DD c, sn; ct.get(-(long long)k1 * n2, c, sn);
                tw[(size_t)k1 * N2 + n2] = {S::from(c), S::from(sn)};
            }
    }
    static Cs store(Cs v, bool renorm) {
        if constexpr (std::is_same_v<T, DD>) {
            if (renorm) return {quick2(v.re.hi, v.re.lo), quick2(v.im.hi, v.im.lo)};
        }
        return v;
    }
    // x: input (n, destroyed), y: n. Returns the result pointer.
    Cs* run(Cs* x, Cs* y) const {
        if (m) {   // X_k = conj(b_k) sum_j x_j conj(b_j) b_{k-j} (fft.hpp:268-278)
            std::vector<Cs> u(m), u2(m);
            for (int j = 0; j < n; j++) u[j] = cmul<S>(x[j], conj<S>(bc[j]));
            for (int j = n; j < m; j++) u[j] = {S::from_d(0.0), S::from_d(0.0)};
            Cs* U = inner[0].run(u.data(), u2.data());
            for (int k = 0; k < m; k++) U[k] = conj<S>(cmul<S>(U[k], Vf[k]));
            Cs* other = U == u.data() ? u2.data() : u.data();
            Cs* Cc = inner[0].run(U, other);
            for (int k = 0; k < n; k++) y[k] = cmul<S>(conj<S>(Cc[k]), conj<S>(bc[k]));
            return y;
        }
        if (!four) return sd.run(x, y);
        std::vector<Cs> col(std::max(N1, N2)), wk(std::max(N1, N2));
        for (int c = 0; c < N2; c++) {             // column pass: N1-point FFTs down the columns
            for (int r = 0; r < N1; r++) col[r] = x[(size_t)r * N2 + c];
            Cs* res = s1.run(col.data(), wk.data());
            for (int k1 = 0; k1 < N1; k1++) {
                Cs v = res[k1];
                if (k1 && (lane || c)) v = cmul<Sub>(v, tw[(size_t)k1 * N2 + c]);
                x[(size_t)k1 * N2 + c] = store(v, lane);
            }
        }
        for (int r = 0; r < N1; r++) {             // row pass, stored transposed into natural order
            for (int i = 0; i < N2; i++) col[i] = x[(size_t)r * N2 + i];
            Cs* res = s2.run(col.data(), wk.data());
            for (int k2 = 0; k2 < N2; k2++) y[(size_t)N1 * k2 + r] = store(res[k2], lane);
        }
        return y;
    }
};

// real FFT of even length N via a complex FFT of N/2 (fft.hpp:352-462)
template <class S, class Sub> struct RFFT {
    using T = typename S::T; using Cs = Cx<T>;
    int N = 0, h = 0; CFFT<S, Sub> c; std::vector<Cs> w;   // w[k] = e^{-2 pi i k / N}
    void init(int N_, bool lane) {
        N = N_; h = N / 2;
        if (N % 2) throw std::runtime_error("RFFT length must be even");
        c.init(h, lane); w.resize(h + 1);
        CisTab ct(N);
        for (int k = 0; k <= h; k++) { DD cc, s; ct.get(-k, cc, s); w[k] = {S::from(cc), S::from(s)}; }
    }
    // X[k] = (E + O w_k)/2, E = z_k + conj(z_{h-k}), O = -i (z_k - conj(z_{h-k}))
    void forward_split(const Cs* z, Cs* X) const {
        X[0] = {S::add(z[0].re, z[0].im), S::from_d(0.0)};
        X[h] = {S::sub(z[0].re, z[0].im), S::from_d(0.0)};
        const T half = S::from_d(0.5);
        if constexpr (std::is_same_v<T, double>) {
            for (int k = 1; k < h; k++) {
                Cs A = z[k], B = conj<S>(z[h - k]);
                Cs E = cadd<S>(A, B), O = mul_mi<T>(csub<S>(A, B), S::ng);
                Cs v = cadd<S>(E, cmul<S>(O, w[k]));
                X[k] = {S::mul(v.re, half), S::mul(v.im, half)};
            }
        } else {
            for (int k = 1; k <= h / 2; k++) {
                const int j = h - k;
                Cs A = z[k], B = conj<S>(z[j]);
                Cs E = cadd<S>(A, B), O = mul_mi<T>(csub<S>(A, B), S::ng);
                Cs v = cadd<S>(E, cmul<S>(O, w[k]));
                X[k] = {S::mul(v.re, half), S::mul(v.im, half)};
                if (j != k) {
                    Cs u = cadd<S>(conj<S>(E), cmul<S>(conj<S>(O), w[j]));
                    X[j] = {S::mul(u.re, half), S::mul(u.im, half)};
                }
            }
        }
    }
    template <class U> void forward(const U* r, Cs* X) const {
        std::vector<Cs> z(h), y(h);
        for (int i = 0; i < h; i++) {
            if constexpr (std::is_same_v<U, DD>) z[i] = {r[2 * i], r[2 * i + 1]};
            else z[i] = {S::from_d(r[2 * i]), S::from_d(r[2 * i + 1])};
        }
        Cs* zz = c.run(z.data(), y.data());
        forward_split(zz, X);
    }
    // Hermitian bins X[0..h] -> z (h complex) for the inverse (conjugated-input trick, fft.hpp:412-458)
    void inverse_split(const Cs* X, Cs* z) const {
        if constexpr (std::is_same_v<T, double>) {
            for (int k = 0; k < h; k++) {
                Cs A = X[k], B = conj<S>(X[h - k]);
                Cs E = cadd<S>(A, B), D = cmul<S>(csub<S>(A, B), conj<S>(w[k]));
                Cs v = {S::sub(E.re, D.im), S::add(E.im, D.re)};
                z[k] = conj<S>(v);
            }
        } else {
            for (int k = 0; k <= h / 2; k++) {
                const int j = h - k;
                Cs A = X[k], B = conj<S>(X[j]);
                Cs E = cadd<S>(A, B), sp = csub<S>(A, B);
                Cs D = cmul<S>(sp, conj<S>(w[k]));
                Cs v = {S::sub(E.re, D.im), S::add(E.im, D.re)};
                z[k] = conj<S>(v);
                if (k && j != k) {
                    Cs Ej = conj<S>(E), Dj = cmul<S>(Cs{S::ng(sp.re), sp.im}, conj<S>(w[j]));
                    Cs u = {S::sub(Ej.re, Dj.im), S::add(Ej.im, Dj.re)};
                    z[j] = conj<S>(u);
                }
            }
        }
    }
    void inverse(const Cs* X, T* r) const {   // unnormalised: N times the true inverse
        std::vector<Cs> z(h), y(h);
        inverse_split(X, z.data());
        Cs* zz = c.run(z.data(), y.data());
        for (int i = 0; i < h; i++) { r[2 * i] = zz[i].re; r[2 * i + 1] = S::ng(zz[i].im); }
    }
};

// ---------------------------------------------------------------------------------------------------------
// kernel design (design.cpp:17-49)
// one tap; ib = I0(beta), ab = a/b
DD kernel_tap(long long n, int L, long long a, long long b, double beta, DD ib, DD ab) {
    DD s = mk(1.0);
    if (n) {   // sin(pi n a/b) / (pi n a/b) with n a reduced mod 2b exactly
        long long m = (long long)(((__int128)n * a) % (2 * (__int128)b));
        DD c, sn; cis(m, 2 * b, c, sn);
        s = div(sn, div1(mul(PI_DD, mul1(mk((double)n), (double)a)), (double)b));
    }
    DD u = div1(mk((double)n), (double)L);
    DD w = div(bessel_i0(mul(mk(beta), sqrt_dd(sub_a(mk(1.0), mul(u, u))))), ib);
    return mul(mul(ab, s), w);
}
std::vector<DD> kernel_taps(int L, long long a, long long b, double beta) {
    std::vector<DD> h(L + 1);
    const DD ib = bessel_i0(mk(beta)), ab = div1(mk((double)a), (double)b);
    for (int n = 0; n <= L; n++) h[n] = kernel_tap(n, L, a, b, beta, ib, ab);
    return h;
}
std::vector<DD> kernel_spectrum(const std::vector<DD>& h, int Nlen, double scale) {
    RFFT<DDA, DDA> R; R.init(Nlen, false);
    const int Lk = (int)h.size() - 1;
    std::vector<DD> r(Nlen, mk(0.0));
    r[0] = h[0];
    for (int n = 1; n <= Lk; n++) { r[n] = h[n]; r[Nlen - n] = h[n]; }
    std::vector<Cx<DD>> X(Nlen / 2 + 1);
    R.forward(r.data(), X.data());
    std::vector<DD> H(Nlen / 2 + 1);
    for (int k = 0; k <= Nlen / 2; k++) H[k] = div1(X[k].re, scale);
    return H;
}

// Burg LPC extrapolation (lpc.cpp:13-41): dd recursion, float64 prediction, stability guard
std::vector<double> burg(const double* x, long long len, long long n, int order = 32, int train = 4096) {
    std::vector<double> out(n, 0.0);
    long long m = std::min<long long>(len, train);
    const double* seg = x + len - m;
    if (m < 2 * order + 8) return out;
    bool nz = false; double mx = 0;
    for (long long i = 0; i < m; i++) { if (seg[i] != 0) nz = true; mx = std::max(mx, std::fabs(seg[i])); }
    if (!nz) return out;
    std::vector<DD> f, bk, a{mk(1.0)};
    for (long long i = 1; i < m; i++) f.push_back(mk(seg[i]));
    for (long long i = 0; i + 1 < m; i++) bk.push_back(mk(seg[i]));
    for (int it = 0; it < order; it++) {
        DD num = mk(0.0), den = mk(0.0);
        for (size_t i = 0; i < f.size(); i++) {
            num = add_a(num, mul(f[i], bk[i]));
            den = add_a(den, add_a(mul(f[i], f[i]), mul(bk[i], bk[i])));
        }
        if (den.hi <= 0) break;
        DD k = div(neg(mul1(num, 2.0)), den);
        a.push_back(mk(0.0));
        std::vector<DD> ar(a.rbegin(), a.rend());
        for (size_t i = 0; i < a.size(); i++) a[i] = add_a(a[i], mul(k, ar[i]));
        std::vector<DD> f2(f.size() - 1), b2(f.size() - 1);
        for (size_t i = 0; i + 1 < f.size(); i++) { f2[i] = add_a(f[i + 1], mul(k, bk[i + 1])); b2[i] = add_a(bk[i], mul(k, f[i])); }
        f.swap(f2); bk.swap(b2);
    }
    const int p = (int)a.size() - 1;
    std::vector<double> c(p);
    for (int i = 0; i < p; i++) c[i] = -(a[p - i].hi + a[p - i].lo);
    std::vector<double> buf(p + n);
    for (int i = 0; i < p; i++) buf[i] = seg[m - p + i];
    for (long long i = 0; i < n; i++) {
        double s = 0;
        for (int j = 0; j < p; j++) s += c[j] * buf[i + j];
        buf[p + i] = s;
    }
    const double lim = 4 * mx;
    for (long long i = 0; i < n; i++) if (!std::isfinite(buf[p + i]) || std::fabs(buf[p + i]) > lim) return out;
    for (long long i = 0; i < n; i++) out[i] = buf[p + i];
    return out;
}

// ---------------------------------------------------------------------------------------------------------
// planning (resampler.cpp:94-167, 187-209)
struct Plan {
    long long P = 0, Q = 0, nout = 0, nin = 0, Lk = 0, c2 = 0, gc = 0, fk = 0, mn = 0, L = 0, fi = 0, fo = 0;
    long long E = 0, pre = 0, span = 0, nblk = 0, base = 0, tot = 0, o = 0;
    int ch = 0, nb = 0, marg = 0, Nb = 0, Mb = 0, hop = 0;
    bool up = false, equal = false, gen = false;
    double fml = 0;
};
Plan make_plan(long long frames, int channels, int rate_in, int rate_out, int precision, double length, double beta,
               int block) {
    Plan p;
    if (channels < 1 || frames < 0) throw std::invalid_argument("bad input buffer");
    if (rate_in < 1 || rate_out < 1) throw std::invalid_argument("sample rates must be positive");
    if (!(length > 0 && length <= 10) || !(beta > 0 && beta <= 200) || block < 0)
        throw std::invalid_argument("option out of range");
    long long fi = rate_in, fo = rate_out, g = std::gcd(fi, fo);
    p.P = fo / g; p.Q = fi / g; p.up = fo > fi; p.nin = frames; p.ch = channels; p.fi = fi; p.fo = fo;
    const long long cap = LLONG_MAX / 16;
    const __int128 n128 = ((__int128)frames * p.P + p.Q - 1) / p.Q;
    if (frames > cap || n128 > cap) throw std::invalid_argument("input too long");
    p.nout = (long long)n128;
    if (fo == fi) { p.equal = true; return p; }
    p.fk = p.up ? fo : fi; p.mn = std::min(fi, fo);
    const double lk = std::floor(length * (double)p.fk);
    if (!(lk >= 1 && lk <= INT_MAX)) throw std::invalid_argument("option out of range");
    p.Lk = (long long)lk;
    p.fml = std::sqrt(beta * beta + 9.8696044010893586) / (2 * 3.14159265358979323846 * length);
    const double drop = std::ceil(2 * p.fml);
    if (!((double)p.mn - drop > 2 * p.fml)) throw std::invalid_argument("length too short for these sample rates");
    p.c2 = p.mn - (long long)drop; p.gc = std::gcd(p.c2, p.fk);
    const long long P = p.P, Q = p.Q;
    p.L = (p.Lk * fi + p.fk - 1) / p.fk + 1;
    long long margl = std::max((p.L + Q + Q - 1) / Q, p.up ? (p.Lk + P + P - 1) / P : 0LL);
    if ((block ? block : 10 * margl) > INT_MAX / 2 / std::max(P, Q))
        throw std::invalid_argument("block, rates or length too large");
    p.E = p.L + Q; p.pre = (p.E + 64 + Q - 1) / Q * Q; p.span = p.pre + frames + p.E + 64;
    auto fix = [&](int v) { while (v % 2 || !smooth7(v / 2)) v++; return v; };
    const bool fast_auto = !block && precision == 0;
    auto round_block = [&](int v) { while (v % 2 || !smooth7(v / 2) || (fast_auto && (v / 2) % 7 == 0)) v++; return v; };
    int nb = block ? fix(block)
                   : round_block((int)(fast_auto ? std::clamp(51200 / std::max(P, Q), 4 * margl, 10 * margl) : 10 * margl));
    if (nb <= 2 * margl) throw std::invalid_argument("option out of range");
    const int marg = (int)margl;
    while (!block && nb > 4 * marg) {   // short inputs: smaller blocks (part of the output definition)
        long long nbk = (p.span + (long long)(nb - 2 * marg) * Q - 1) / ((long long)(nb - 2 * marg) * Q);
        if ((long long)channels * ((nbk + 3) / 4) >= 16) break;
        int nn = round_block(std::max(4 * marg, nb * 3 / 4));
        if (nn >= nb) break;
        nb = nn;
    }
    p.nb = nb; p.marg = marg; p.Nb = nb * (int)Q; p.Mb = nb * (int)P; p.hop = nb - 2 * marg;
    p.nblk = (p.span + (long long)p.hop * Q - 1) / ((long long)p.hop * Q);
    p.base = (long long)marg * Q; p.tot = p.base + p.nblk * p.hop * Q + p.Nb;
    p.o = p.pre / Q * P;
    p.gen = (double)(std::max(fi, fo) - p.mn) / 2 < 3 * p.fml;
    return p;
}

// staging and edge extension (resampler.cpp:168-186)
std::vector<std::vector<double>> stage(const Plan& p, const double* in, bool lpc) {
    std::vector<std::vector<double>> xx(p.ch);
    double fd[32];
    for (int i = 0; i < 32; i++) { DD co, sn; cis(i, 64, co, sn); fd[i] = mul1(add_a(mk(1.0), co), 0.5).hi; }
    for (int c = 0; c < p.ch; c++) {
        xx[c].assign(p.tot, 0.0);
        double* X = xx[c].data() + p.base + p.pre;
        for (long long i = 0; i < p.nin; i++) X[i] = in[i * p.ch + c];
        if (lpc) {
            const long long ext = p.E + 32;
            std::vector<double> tl = burg(X, p.nin, ext);
            for (int i = 0; i < 32; i++) tl[ext - 32 + i] *= fd[i];
            const long long m = std::min<long long>(p.nin, 4096);
            std::vector<double> xr(m);
            for (long long i = 0; i < m; i++) xr[i] = X[m - 1 - i];
            std::vector<double> hd = burg(xr.data(), m, ext);
            for (int i = 0; i < 32; i++) hd[ext - 32 + i] *= fd[i];
            for (long long i = 0; i < ext; i++) { X[p.nin + i] = tl[i]; X[-1 - i] = hd[i]; }
        }
    }
    return xx;
}

// kernel spectrum for the plan; Jg > 0 selects the general fold
std::vector<DD> design(const Plan& p, double length, double beta, int& Jg) {
    Jg = 0;
    const long long Nb = p.Nb, Mb = p.Mb, rate_in = p.fi;
    if (p.gen) {
        long long Nk = std::max(Nb, Mb) * 5LL / 4; Nk += Nk & 1;
        while (!smooth7(Nk / 2)) Nk += 2;
        __int128 A = (__int128)p.c2 * Nb, B = (__int128)Nk * rate_in;
        long long g2 = (long long)std::gcd((unsigned long long)p.c2 * (unsigned long long)Nb,
                                           (unsigned long long)Nk * (unsigned long long)rate_in);
        double fk2 = (double)Nk * rate_in / Nb;
        std::vector<DD> h2 = kernel_taps((int)std::floor(length * fk2), (long long)(A / g2), (long long)(B / g2), beta);
        Jg = (int)std::min<long long>(Nk / 2, (long long)std::ceil((p.mn / 2.0 + 4 * p.fml) * Nb / rate_in));
        return kernel_spectrum(h2, (int)Nk, (double)Nb);
    }
    std::vector<DD> h = kernel_taps((int)p.Lk, p.c2 / p.gc, p.fk / p.gc, beta);
    return kernel_spectrum(h, p.up ? p.Mb : p.Nb, (double)p.Nb);
}

// overlap-save engine (resampler.cpp:34-88), one (channel, block) task after another
template <class S, class Sub>
void engine(const Plan& pl, const std::vector<std::vector<double>>& xx, double* out, const std::vector<DD>& Hdd, int Jg) {
    using T = typename S::T; using Cs = Cx<T>;
    const long long P = pl.P, Q = pl.Q;
    const int Nb = pl.Nb, Mb = pl.Mb, K = Mb / 2 + 1;
    RFFT<S, Sub> Fw, Iv; Fw.init(Nb, true); Iv.init(Mb, true);
    std::vector<T> H(Hdd.size());
    for (size_t k = 0; k < Hdd.size(); k++) H[k] = S::from(Hdd[k]);
    std::vector<T> r(Mb); std::vector<Cs> X(Nb / 2 + 1), Y(K);
    const T zero = S::from_d(0.0);
    for (int c = 0; c < pl.ch; c++)
        for (long long b = 0; b < pl.nblk; b++) {
            const double* s = xx[c].data() + b * pl.hop * Q;
            Fw.forward(s, X.data());
            if (Jg) {   // general fold: every bin of the finer-grid response folded mod Mb, ascending j
                for (int k = 0; k < K; k++) Y[k] = {zero, zero};
                for (int j = 0; j <= Jg; j++) {
                    int jj = j % Nb;
                    Cs x = jj <= Nb / 2 ? X[jj] : conj<S>(X[Nb - jj]);
                    T g = H[j];
                    Cs v = {S::mul(x.re, g), S::mul(x.im, g)};
                    int k = j % Mb;
                    if (j == 0) Y[0] = cadd<S>(Y[0], v);
                    else if (k == 0 || 2 * k == Mb) Y[k].re = S::add(Y[k].re, S::mul(S::from_d(2.0), v.re));
                    else if (2 * k < Mb) Y[k] = cadd<S>(Y[k], v);
                    else Y[Mb - k] = cadd<S>(Y[Mb - k], conj<S>(v));
                }
                Y[0].im = zero; Y[Mb / 2].im = zero;
            } else if (!pl.up) {   // downsampling: bins above the output Nyquist fold onto Mb - j
                for (int k = 0; k < K; k++) Y[k] = {S::mul(X[k].re, H[k]), S::mul(X[k].im, H[k])};
                const int jmax = std::min(Nb / 2, Mb - 1);
                for (int j = K; j <= jmax; j++) {
                    Cs v = {S::mul(X[j].re, H[j]), S::mul(X[j].im, H[j])};
                    Y[Mb - j] = cadd<S>(Y[Mb - j], conj<S>(v));
                }
                Y[Mb / 2] = {S::mul(S::mul(S::from_d(2.0), X[Mb / 2].re), H[Mb / 2]), zero};
            } else {               // upsampling: the input spectrum extended periodically
                for (int k = 0; k < K; k++) {
                    int kk = k % Nb;
                    Cs v = kk <= Nb / 2 ? X[kk] : conj<S>(X[Nb - kk]);
                    Y[k] = {S::mul(v.re, H[k]), S::mul(v.im, H[k])};
                }
                Y[Mb / 2].im = zero;
            }
            Iv.inverse(Y.data(), r.data());
            double* dst = out + (size_t)c * pl.nout;
            for (long long i = 0; i < (long long)pl.hop * P; i++) {
                long long d = b * pl.hop * P + i - pl.o;
                if (d >= 0 && d < pl.nout) {
                    if constexpr (std::is_same_v<T, DD>) dst[d] = r[pl.marg * P + i].hi + r[pl.marg * P + i].lo;
                    else dst[d] = r[pl.marg * P + i];
                }
            }
        }
}

}  // namespace oracle

// ---------------------------------------------------------------------------------------------------------
// C entry points for tests (ctypes). Status 0 = ok, 1 = invalid argument, 2 = other error.
static thread_local std::string g_err;
static int fail_with(const std::exception& e, int code) { g_err = e.what(); return code; }

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// Plan integers: [P, Q, nout, nb, marg, Nb, Mb, hop, nblk, o, tot, base, pre, E, L, Lk, c2, gen, up, equal]
int oracle_plan(long long frames, int ch, int rin, int rout, int precision, double length, double beta, int block,
                long long* v) {
    try {
        oracle::Plan p = oracle::make_plan(frames, ch, rin, rout, precision, length, beta, block);
        long long a[20] = {p.P, p.Q, p.nout, p.nb, p.marg, p.Nb, p.Mb, p.hop, p.nblk, p.o, p.tot, p.base, p.pre,
                           p.E, p.L, p.Lk, p.c2, p.gen, p.up, p.equal};
        std::memcpy(v, a, sizeof a);
        return 0;
    } catch (const std::invalid_argument& e) { return fail_with(e, 1); }
    catch (const std::exception& e) { return fail_with(e, 2); }
}

// precision: 0 = fast (float64), 1 = exact (double-double). out: ch x nout planar.
int oracle_resample(const double* in, long long frames, int ch, int rin, int rout, int precision, double length,
                    double beta, int block, int lpc, double* out) {
    try {
        if (frames > 0 && !in) throw std::invalid_argument("bad input buffer");
        oracle::Plan p = oracle::make_plan(frames, ch, rin, rout, precision, length, beta, block);
        if (p.equal) {
            for (int c = 0; c < ch; c++) for (long long i = 0; i < frames; i++) out[(size_t)c * p.nout + i] = in[i * ch + c];
            return 0;
        }
        auto xx = oracle::stage(p, in, lpc != 0);
        int Jg = 0;
        std::vector<oracle::DD> H = oracle::design(p, length, beta, Jg);
        if (precision == 0) oracle::engine<oracle::F64, oracle::F64>(p, xx, out, H, Jg);
        else oracle::engine<oracle::DDA, oracle::DDS>(p, xx, out, H, Jg);
        return 0;
    } catch (const std::invalid_argument& e) { return fail_with(e, 1); }
    catch (const std::exception& e) { return fail_with(e, 2); }
}

// cis(2 pi k / N) as [c.hi, c.lo, s.hi, s.lo]
void oracle_cis(long long k, long long N, double* v) {
    oracle::DD c, s; oracle::cis(k, N, c, s);
    v[0] = c.hi; v[1] = c.lo; v[2] = s.hi; v[3] = s.lo;
}

// CisTable(N).get(k) as [c.hi, c.lo, s.hi, s.lo] for k = k0 .. k0+count-1
void oracle_cis_table(long long N, long long k0, long long count, double* v) {
    oracle::CisTab t(N);
    for (long long i = 0; i < count; i++) {
        oracle::DD c, s; t.get(k0 + i, c, s);
        v[4 * i] = c.hi; v[4 * i + 1] = c.lo; v[4 * i + 2] = s.hi; v[4 * i + 3] = s.lo;
    }
}

// one-sided taps h[0..L] as (hi, lo) pairs
int oracle_kernel_taps(int L, long long a, long long b, double beta, double* v) {
    try {
        auto h = oracle::kernel_taps(L, a, b, beta);
        for (int i = 0; i <= L; i++) { v[2 * i] = h[i].hi; v[2 * i + 1] = h[i].lo; }
        return 0;
    } catch (const std::exception& e) { return fail_with(e, 2); }
}

// kernel spectrum of taps (hi, lo pairs, L+1 of them) in length Nlen, divided by scale: Nlen/2+1 (hi, lo) pairs
// taps h[idx[i]] only, for spot checks of very long kernels
int oracle_kernel_taps_at(int L, long long a, long long b, double beta, const long long* idx, int count, double* v) {
    try {
        const oracle::DD ib = oracle::bessel_i0(oracle::mk(beta)), ab = oracle::div1(oracle::mk((double)a), (double)b);
        for (int i = 0; i < count; i++) {
            if (idx[i] < 0 || idx[i] > L) throw std::invalid_argument("tap index out of range");
            const oracle::DD t = oracle::kernel_tap(idx[i], L, a, b, beta, ib, ab);
            v[2 * i] = t.hi; v[2 * i + 1] = t.lo;
        }
        return 0;
    } catch (const std::exception& e) { return fail_with(e, 2); }
}

int oracle_kernel_spectrum(const double* taps, int L, int Nlen, double scale, double* v) {
    try {
        std::vector<oracle::DD> h(L + 1);
        for (int i = 0; i <= L; i++) h[i] = {taps[2 * i], taps[2 * i + 1]};
        auto H = oracle::kernel_spectrum(h, Nlen, scale);
        for (int k = 0; k <= Nlen / 2; k++) { v[2 * k] = H[k].hi; v[2 * k + 1] = H[k].lo; }
        return 0;
    } catch (const std::exception& e) { return fail_with(e, 2); }
}

// Burg extension of n samples after x[0..len)
void oracle_burg(const double* x, long long len, long long n, double* v) {
    auto o = oracle::burg(x, len, n);
    std::memcpy(v, o.data(), sizeof(double) * n);
}

// Complex FFT of length n as the engine runs it. kind 0: Exact (dd, lane four-step), 1: Fast (f64), 2: dd as in
// the kernel design (no lanes). in/out: n elements of [re.hi, re.lo, im.hi, im.lo] (lo ignored for Fast).
int oracle_cfft(int n, int kind, const double* in, double* out) {
    try {
        using namespace oracle;
        if (kind == 1) {
            CFFT<F64, F64> f; f.init(n, true);
            std::vector<Cx<double>> x(n), y(n);
            for (int i = 0; i < n; i++) x[i] = {in[4 * i], in[4 * i + 2]};
            Cx<double>* r = f.run(x.data(), y.data());
            for (int i = 0; i < n; i++) { out[4 * i] = r[i].re; out[4 * i + 1] = 0; out[4 * i + 2] = r[i].im; out[4 * i + 3] = 0; }
        } else {
            CFFT<DDA, DDS> fe; CFFT<DDA, DDA> fd;
            std::vector<Cx<DD>> x(n), y(n);
            for (int i = 0; i < n; i++) x[i] = {{in[4 * i], in[4 * i + 1]}, {in[4 * i + 2], in[4 * i + 3]}};
            Cx<DD>* r;
            if (kind == 0) { fe.init(n, true); r = fe.run(x.data(), y.data()); }
            else { fd.init(n, false); r = fd.run(x.data(), y.data()); }
            for (int i = 0; i < n; i++) { out[4 * i] = r[i].re.hi; out[4 * i + 1] = r[i].re.lo; out[4 * i + 2] = r[i].im.hi; out[4 * i + 3] = r[i].im.lo; }
        }
        return 0;
    } catch (const std::exception& e) { return fail_with(e, 2); }
}

}  // extern "C"
