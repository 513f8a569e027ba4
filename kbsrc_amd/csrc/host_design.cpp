// Host planning and design. See host_design.hpp for what each function follows in the reference.
#include "host_design.hpp"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <thread>

#pragma clang fp contract(off)

namespace kbsrc::gpu {

const dd kPi = {3.141592653589793116e+00, 1.224646799147353207e-16};

bool smooth7(long long n) {
    if (n < 1) return false;
    for (long long f : {2LL, 3LL, 5LL, 7LL})
        while (n % f == 0) n /= f;
    return n == 1;
}
long long next_smooth7(long long n) {
    while (!smooth7(n)) ++n;
    return n;
}

Plan make_plan(long long frames, int channels, int rate_in, int rate_out, bool fast, double length, double beta,
               int block) {
    if (channels < 1 || frames < 0) throw std::invalid_argument("bad input buffer");
    if (rate_in < 1 || rate_out < 1) throw std::invalid_argument("sample rates must be positive");
    if (!(length > 0 && length <= 10) || !(beta > 0 && beta <= 200) || block < 0)
        throw std::invalid_argument("option out of range");
    Plan p;
    p.fast = fast;
    p.fi = rate_in;
    p.fo = rate_out;
    const long long g = std::gcd(p.fi, p.fo);
    p.P = p.fo / g;
    p.Q = p.fi / g;
    p.up = p.fo > p.fi;
    p.nin = frames;
    p.ch = channels;
    // output length ceil(frames * P / Q) in 128-bit, refused far below the long long limit
    const long long cap = LLONG_MAX / 16;
    const __int128 n128 = ((__int128)frames * p.P + p.Q - 1) / p.Q;
    if (frames > cap || n128 > cap) throw std::invalid_argument("input too long");
    p.nout = (long long)n128;
    if (p.fo == p.fi) {
        p.equal = true;
        return p;
    }
    const long long P = p.P, Q = p.Q;
    p.fk = p.up ? p.fo : p.fi;   // kernel sampling rate: the higher one
    p.mn = std::min(p.fi, p.fo);
    const double lk = std::floor(length * (double)p.fk);
    if (!(lk >= 1 && lk <= INT_MAX)) throw std::invalid_argument("option out of range");
    p.Lk = (long long)lk;
    // Kaiser main-lobe half-width; the cutoff sits ceil(2 fml)/2 Hz below the lower Nyquist frequency
    p.fml = std::sqrt(beta * beta + 9.8696044010893586) / (2 * 3.14159265358979323846 * length);
    const double drop = std::ceil(2 * p.fml);
    if (!((double)p.mn - drop > 2 * p.fml)) throw std::invalid_argument("length too short for these sample rates");
    p.c2 = p.mn - (long long)drop;
    p.gc = std::gcd(p.c2, p.fk);
    p.L = (p.Lk * p.fi + p.fk - 1) / p.fk + 1;   // kernel reach in input samples
    const long long margl = std::max((p.L + Q + Q - 1) / Q, p.up ? (p.Lk + P + P - 1) / P : 0LL);
    if ((block ? block : 10 * margl) > INT_MAX / 2 / std::max(P, Q))
        throw std::invalid_argument("block, rates or length too large");
    p.E = p.L + Q;
    p.pre = (p.E + 64 + Q - 1) / Q * Q;
    p.span = p.pre + frames + p.E + 64;
    // nb is even with a 7-smooth half; automatic Fast blocks aim at ~51200-point transforms and avoid radix 7
    const bool fast_auto = !block && fast;
    auto fix = [](int v) {
        while (v % 2 || !smooth7(v / 2)) ++v;
        return v;
    };
    auto round_block = [&](int v) {
        while (v % 2 || !smooth7(v / 2) || (fast_auto && (v / 2) % 7 == 0)) ++v;
        return v;
    };
    int nb = block ? fix(block)
                   : round_block((int)(fast_auto ? std::clamp(51200 / std::max(P, Q), 4 * margl, 10 * margl)
                                                 : 10 * margl));
    if (nb <= 2 * margl) throw std::invalid_argument("option out of range");
    const int marg = (int)margl;
    // Short inputs: shrink automatic blocks until 8 workers would find work. Part of the output definition.
    while (!block && nb > 4 * marg) {
        const long long step = (long long)(nb - 2 * marg) * Q;
        const long long nbk = (p.span + step - 1) / step;
        if ((long long)channels * ((nbk + 3) / 4) >= 16) break;
        const int nn = round_block(std::max(4 * marg, nb * 3 / 4));
        if (nn >= nb) break;
        nb = nn;
    }
    p.nb = nb;
    p.marg = marg;
    p.Nb = nb * (int)Q;
    p.Mb = nb * (int)P;
    p.hop = nb - 2 * marg;
    p.nblk = (p.span + (long long)p.hop * Q - 1) / ((long long)p.hop * Q);
    p.base = (long long)marg * Q;
    p.tot = p.base + p.nblk * p.hop * Q + p.Nb;
    p.xstride = (p.tot + 31) / 32 * 32;
    p.o = p.pre / Q * P;
    p.gen = (double)(std::max(p.fi, p.fo) - p.mn) / 2 < 3 * p.fml;
    return p;
}

// --- twiddles -----------------------------------------------------------------------------------------------
static void sincos_taylor(dd x, dd& c, dd& s) {   // x in [0, pi/4]
    const dd x2 = ddmul(x, x);
    dd ts = x, sum_s = x, tc = mkdd(1.0), sum_c = mkdd(1.0);
    for (int k = 1; k < 40; ++k) {
        tc = dddiv1(ddneg(ddmul(tc, x2)), double((2 * k - 1) * (2 * k)));
        sum_c = add_acc(sum_c, tc);
        ts = dddiv1(ddneg(ddmul(ts, x2)), double((2 * k) * (2 * k + 1)));
        sum_s = add_acc(sum_s, ts);
        if (std::fabs(ts.hi) < 1e-40 && std::fabs(tc.hi) < 1e-40) break;
    }
    c = sum_c;
    s = sum_s;
}

void cis(long long k, long long N, dd& c, dd& s) {
    k %= N;
    if (k < 0) k += N;
    // 2 pi k / N = (pi/4) * (8k / N): octant index and exact remainder
    const long long eight = 8 * k, oct = eight / N, rem = eight - oct * N;
    const bool mirrored = oct & 1;
    long long quad = oct / 2;
    const dd quarter_pi = {kPi.hi / 4, kPi.lo / 4};
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

CisTable::CisTable(long long n) : N_(n) {
    B_ = (long long)std::ceil(std::sqrt((double)n));
    const long long J = n / B_ + 1;
    ac_.resize(J); as_.resize(J); bc_.resize(B_); bs_.resize(B_);
    for (long long j = 0; j < J; ++j) cis(j * B_, n, ac_[j], as_[j]);
    for (long long l = 0; l < B_; ++l) cis(l, n, bc_[l], bs_[l]);
}

void CisTable::get(long long k, dd& c, dd& s) const {
    k %= N_;
    if (k < 0) k += N_;
    const long long j = k / B_, l = k % B_;
    if (l == 0) { c = ac_[j]; s = as_[j]; return; }
    if (j == 0) { c = bc_[l]; s = bs_[l]; return; }
    c = sub_acc(ddmul(ac_[j], bc_[l]), ddmul(as_[j], bs_[l]));
    s = add_acc(ddmul(ac_[j], bs_[l]), ddmul(as_[j], bc_[l]));
}

// --- kernel design ------------------------------------------------------------------------------------------
dd dd_sqrt(dd a) {   // Newton step from the FP64 root (src/dsp/double_double.hpp:94-98)
    if (a.hi <= 0) return mkdd(0.0);
    const double y = std::sqrt(a.hi);
    const dd r = sub_acc(a, ddmul1(mkdd(y), y));
    return quick2(y, r.hi / (2 * y));
}

dd bessel_i0(dd x) {   // sum (x^2/4)^k / (k!)^2 (src/dsp/double_double.hpp:100-107)
    const dd q = dddiv1(ddmul(x, x), 4.0);
    dd term = mkdd(1.0), sum = mkdd(1.0);
    for (int k = 1; k < 2000; ++k) {
        term = dddiv1(ddmul(term, q), double(k) * double(k));
        sum = add_acc(sum, term);
        if (term.hi < sum.hi * 1e-34) break;
    }
    return sum;
}

std::vector<dd> kernel_taps(int L, long long a, long long b, double beta) {
    std::vector<dd> h(L + 1);
    const dd i0_beta = bessel_i0(mkdd(beta));
    const dd ratio = dddiv1(mkdd((double)a), (double)b);
    auto tap = [&](int n) {
        dd sinc = mkdd(1.0);
        if (n) {   // sin(pi n a / b) / (pi n a / b), the angle reduced exactly as (n a mod 2b) / 2b turns
            const long long m = (long long)(((__int128)n * a) % (2 * (__int128)b));
            dd c, s;
            cis(m, 2 * b, c, s);
            const dd angle = dddiv1(ddmul(kPi, ddmul1(mkdd((double)n), (double)a)), (double)b);
            sinc = dddiv(s, angle);
        }
        const dd u = dddiv1(mkdd((double)n), (double)L);
        const dd win = dddiv(bessel_i0(ddmul(mkdd(beta), dd_sqrt(sub_acc(mkdd(1.0), ddmul(u, u))))), i0_beta);
        h[n] = ddmul(ddmul(ratio, sinc), win);
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<long long>(hw, std::max(1, (L + 1) / 512));
    std::atomic<int> next(0);
    auto work = [&]() {
        for (int c; (c = next.fetch_add(256)) <= L;)
            for (int n = c; n <= std::min(L, c + 255); ++n) tap(n);
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    return h;
}

std::vector<dd> circular_kernel(const std::vector<dd>& h, int n) {
    std::vector<dd> r(n, mkdd(0.0));
    const int Lk = (int)h.size() - 1;
    r[0] = h[0];
    for (int i = 1; i <= Lk; ++i) {
        r[i] = h[i];
        r[n - i] = h[i];
    }
    return r;
}

// --- Burg edge extension ------------------------------------------------------------------------------------
std::vector<double> burg_extrap(const double* x, long long len, long long n, int order, int train) {
    std::vector<double> out(n, 0.0);
    const long long m = std::min<long long>(len, train);
    const double* seg = x + len - m;
    if (m < 2 * order + 8) return out;
    bool any = false;
    double peak = 0;
    for (long long i = 0; i < m; ++i) {
        any |= seg[i] != 0;
        peak = std::max(peak, std::fabs(seg[i]));
    }
    if (!any) return out;
    // forward and backward prediction errors, then the reflection recursion, all in dd
    std::vector<dd> fe(m - 1), be(m - 1), a{mkdd(1.0)};
    for (long long i = 0; i + 1 < m; ++i) {
        fe[i] = mkdd(seg[i + 1]);
        be[i] = mkdd(seg[i]);
    }
    for (int it = 0; it < order; ++it) {
        dd num = mkdd(0.0), den = mkdd(0.0);
        for (size_t i = 0; i < fe.size(); ++i) {
            num = add_acc(num, ddmul(fe[i], be[i]));
            den = add_acc(den, add_acc(ddmul(fe[i], fe[i]), ddmul(be[i], be[i])));
        }
        if (den.hi <= 0) break;
        const dd kref = dddiv(ddneg(ddmul1(num, 2.0)), den);
        a.push_back(mkdd(0.0));
        const std::vector<dd> rev(a.rbegin(), a.rend());
        for (size_t i = 0; i < a.size(); ++i) a[i] = add_acc(a[i], ddmul(kref, rev[i]));
        std::vector<dd> f2(fe.size() - 1), b2(fe.size() - 1);
        for (size_t i = 0; i + 1 < fe.size(); ++i) {
            f2[i] = add_acc(fe[i + 1], ddmul(kref, be[i + 1]));
            b2[i] = add_acc(be[i], ddmul(kref, fe[i]));
        }
        fe.swap(f2);
        be.swap(b2);
    }
    const int ord = (int)a.size() - 1;
    std::vector<double> coef(ord);
    for (int i = 0; i < ord; ++i) coef[i] = -(a[ord - i].hi + a[ord - i].lo);
    std::vector<double> buf(ord + n);
    for (int i = 0; i < ord; ++i) buf[i] = seg[m - ord + i];
    for (long long i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < ord; ++j) s += coef[j] * buf[i + j];
        buf[ord + i] = s;
    }
    const double lim = 4 * peak;
    for (long long i = 0; i < n; ++i)
        if (!std::isfinite(buf[ord + i]) || std::fabs(buf[ord + i]) > lim) return out;
    for (long long i = 0; i < n; ++i) out[i] = buf[ord + i];
    return out;
}

void edge_fade(double fd[32]) {
    for (int i = 0; i < 32; ++i) {
        dd c, s;
        cis(i, 64, c, s);
        fd[i] = ddmul1(add_acc(mkdd(1.0), c), 0.5).hi;
    }
}

void edge_extensions(const Plan& p, const double* in, int channel, std::vector<double>& head,
                     std::vector<double>& tail) {
    const long long ext = p.E + 32, nin = p.nin, ch = p.ch;
    const long long m = std::min<long long>(nin, 4096);
    double fd[32];
    edge_fade(fd);
    // tail: the last m samples predict forward; head: the first m samples reversed predict backward
    std::vector<double> last(m), first_rev(m);
    for (long long i = 0; i < m; ++i) {
        last[i] = in[(nin - m + i) * ch + channel];
        first_rev[i] = in[(m - 1 - i) * ch + channel];
    }
    tail = burg_extrap(last.data(), m, ext);
    head = burg_extrap(first_rev.data(), m, ext);
    for (int i = 0; i < 32; ++i) {
        tail[ext - 32 + i] *= fd[i];
        head[ext - 32 + i] *= fd[i];
    }
}

}  // namespace kbsrc::gpu
