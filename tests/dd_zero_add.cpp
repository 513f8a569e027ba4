// Host check of the identity the odd-radix butterflies use to start their sine sums at the first product instead of
// adding it to zero: for b = ddmul(D, c), where D is a dd difference (the butterfly's D = x - y, one-TwoSum or
// accurate add) and c a dd constant, add_slp({0, 0}, b) and add_acc({0, 0}, b) both equal {b.hi + 0.0, b.lo + 0.0}
// bit for bit: ddmul's result is normalised (fl(hi + lo) == hi), so either add returns it with a negative zero
// component turned positive, which is what adding 0.0 to each component does. Random operands over the whole
// exponent range, with exact zeros of both signs, subnormals, equal operands (D = 0) and near-cancellations. Prints
// the case count; exit 1 and the first failing case on a mismatch.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "dd.hpp"

using kbsrc::gpu::dd;
namespace K = kbsrc::gpu;

static uint64_t bits(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    return u;
}
static bool same(dd a, dd b) { return bits(a.hi) == bits(b.hi) && bits(a.lo) == bits(b.lo); }

int main(int argc, char** argv) {
    const long n = argc > 1 ? std::atol(argv[1]) : 2000000;
    std::mt19937_64 rng(12345);
    std::uniform_int_distribution<int> kind(0, 9), ex(-1080, 1020);
    std::uniform_real_distribution<double> mant(0.5, 1.0);
    auto num = [&]() -> double {
        switch (kind(rng)) {
            case 0: return 0.0;
            case 1: return -0.0;
            case 2: return std::ldexp(mant(rng), -1074 + (int)(rng() % 60)) * ((rng() & 1) ? -1 : 1);   // subnormal
            default: return std::ldexp(mant(rng), ex(rng)) * ((rng() & 1) ? -1 : 1);
        }
    };
    auto ddnum = [&]() -> dd {   // a normalised dd (quick2 of a value and a smaller one), or a plain double
        const double h = num();
        if (rng() % 4 == 0) return dd{h, 0.0};
        const double l = std::ldexp(mant(rng), std::ilogb(h == 0 ? 1.0 : h) - 54 - (int)(rng() % 8)) * ((rng() & 1) ? -1 : 1);
        return K::quick2(h, h == 0 ? 0.0 : l);
    };
    const dd zero{0.0, 0.0};
    long checked = 0;
    for (long i = 0; i < n; ++i) {
        const dd x = ddnum(), c = ddnum();
        dd y = ddnum();
        switch (rng() % 4) {   // D = x - y: equal operands, near-cancellation, or unrelated
            case 0: y = x; break;
            case 1: y = dd{x.hi, -x.lo}; break;
            case 2: y = dd{std::nextafter(x.hi, 0.0), x.lo}; break;
            default: break;
        }
        for (int add = 0; add < 2; ++add) {
            const dd D = add ? K::add_acc(x, K::ddneg(y)) : K::add_slp(x, K::ddneg(y));
            const dd b = K::ddmul(D, c);
            if (!std::isfinite(b.hi) || !std::isfinite(b.lo)) continue;
            const dd s = K::add_slp(zero, b), t = K::add_acc(zero, b), z{b.hi + 0.0, b.lo + 0.0};
            if (!same(s, z) || !same(t, z)) {
                std::printf("mismatch: D=(%a,%a) c=(%a,%a) b=(%a,%a) slp=(%a,%a) acc=(%a,%a)\n", D.hi, D.lo, c.hi,
                            c.lo, b.hi, b.lo, s.hi, s.lo, t.hi, t.lo);
                return 1;
            }
            ++checked;
        }
    }
    std::printf("checked %ld\n", checked);
    return 0;
}
