// Host planning and design. See host_design.hpp for what each function follows in the reference.
#include "host_design.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <stdexcept>

#pragma clang fp contract(off)

namespace kbsrc::gpu {

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

// --- twiddles (cis itself is in design_math.hpp, shared with the GPU) ----------------------------------------
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
    cis_compose(ac_[j], as_[j], bc_[l], bs_[l], c, s);
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
    const long long m = edge_frames(p);
    edge_extensions(p, in, in + (p.nin - m) * p.ch, channel, head, tail);
}

std::vector<double> edge_extension(const Plan& p, const double* first, const double* last_frames, int channel,
                                   bool at_tail) {
    const long long ext = p.E + 32, ch = p.ch;
    const long long m = edge_frames(p);
    double fd[32];
    edge_fade(fd);
    // tail: the last m samples predict forward; head: the first m samples reversed predict backward
    std::vector<double> train(m);
    for (long long i = 0; i < m; ++i)
        train[i] = at_tail ? last_frames[i * ch + channel] : first[(m - 1 - i) * ch + channel];
    std::vector<double> e = burg_extrap(train.data(), m, ext);
    for (int i = 0; i < 32; ++i) e[ext - 32 + i] *= fd[i];
    return e;
}
void edge_extensions(const Plan& p, const double* first, const double* last_frames, int channel,
                     std::vector<double>& head, std::vector<double>& tail) {
    tail = edge_extension(p, first, last_frames, channel, true);
    head = edge_extension(p, first, last_frames, channel, false);
}

}  // namespace kbsrc::gpu
