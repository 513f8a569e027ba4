// The int8 matrix-core core of the split-precision (Ozaki) radix-16 DFT stage, shared by the standalone stage
// (split_dft.hip) and the LDS sub-FFT pass of the split form (kernels.hip).
//
// y = F x for vectors x of 16 complex dd values, F[k][n] = e^{-2 pi i k n / 16}, as one real 32 x 32 product
// [Re y; Im y] = M x' with x' = (Re x[0], Im x[0], Re x[1], ...). Each vector is scaled by a power of two to
// |z| < 1/2 and cut into S base-128 digits (each chunk of up to 4 a signed leading digit in [-64, 64] and unsigned
// ones in [0, 127]), z = sum_j b_j 2^{-7 (j + 1)}; M / 2 into balanced digits a_i in [-65, 65] (on the host). The
// products a_i b_j with i + j <= S + 1 run on v_mfma_i32_16x16x32_i8 (exact int32 sums of 32 terms), summed per
// weight 2^{-7 (i + j)} in int32 (exact: at most S products of |a| <= 65, |b| <= 127 over 32 terms), and the
// weighted sums are recombined in FP64 (exact groups, one TwoSum chain) into dd. Error: the digits kept (about
// 2^{-7 S} of the vector's largest value) and the products dropped.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dd.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {
namespace split {

typedef int v4i __attribute__((ext_vector_type(4)));

// Wave layout of v_mfma_i32_16x16x32_i8 (checked on the device against an integer product): lane l holds A row
// l % 16, k = 8 (l / 16) .. + 7; B column l % 16, same k; D column l % 16, rows 4 (l / 16) .. + 3.
//
// One wave, 16 vectors (column col = lane % 16 of the product). On entry lane l holds the 8 reals k = 8 q .. 8 q + 7
// (q = l / 16) of its column's vector, real k = Re x[k / 2] for even k, Im x[k / 2] for odd k, as (hi, lo). On return it holds
// y[4 q + r] = (re[r], im[r]) for r = 0..3, as normalised dd. Every lane of the wave must call it.
template <int S>
__device__ __forceinline__ void split_core(const double (&hi)[8], const double (&lo)[8], const long* __restrict__ adig,
                                           int lane, dd (&re)[4], dd (&im)[4]) {
    // the vector's scale: largest |hi| over its 32 reals (4 lanes share a column)
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) m = __builtin_fmax(m, __builtin_fabs(hi[i]));
    m = __builtin_fmax(m, __shfl_xor(m, 16));
    m = __builtin_fmax(m, __shfl_xor(m, 32));
    int ex = 0;
    if (m > 0) (void)__builtin_frexp(m, &ex);
    const int e = m > 0 ? ex + 1 : 0;   // |x| 2^{-e} < 1/2
    // S base-128 digits of each real v = (hi + lo) 2^{-e}, most significant first, in chunks of up to 4 digits:
    // t = 2^{7 nd} h, c = rint(t), h = t - c, all exact, |h| <= 1/2 and |c| <= 2^{7 nd - 1}; the chunk's digits
    // are c's 7-bit fields, the leading one signed (c >> 7 (nd - 1), in [-64, 64]) and the others unsigned in
    // [0, 127] (two's complement), so v = sum_j d_j 128^{-(j+1)} + remainder. Above 6 digits the low part joins
    // the remainder by an exact TwoSum after each chunk (its scaled size stays below 2^-26); the last chunk is
    // rounded, so the remainder dropped is at most (1/2 + 2^-26) 128^{-S}. Byte i of b[j] is digit j of real
    // 8 q + i: each chunk's digit bytes are transposed from the eight reals' chunk words by byte permutes.
    constexpr int NC = (S + 3) / 4;
    long b[S];
    double h[8], l[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        h[i] = __builtin_ldexp(hi[i], -e);
        l[i] = __builtin_ldexp(lo[i], -e);
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const int nd = k == NC - 1 ? S - 4 * k : 4;
        unsigned x[8];   // byte m: digit m of the chunk (m = 0 most significant)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double t = __builtin_ldexp(h[i], 7 * nd);
            const double c = __builtin_rint(t);
            h[i] = t - c;
            const unsigned ci = (unsigned)(int)c;
            unsigned w = (ci >> (7 * (nd - 1))) & 0xffu;
#pragma unroll
            for (int m = 1; m < nd; ++m) w |= ((ci >> (7 * (nd - 1 - m))) & 0x7fu) << (8 * m);
            x[i] = w;
            if (S > 6 && k < NC - 1) {
                const dd z = knuth2(h[i], __builtin_ldexp(l[i], 7 * nd));
                h[i] = z.hi;
                l[i] = z.lo;
            }
        }
        // 4 x 4 byte transposes (reals 0-3 into the low dwords, 4-7 into the high ones): y[m] byte i = x[i] byte m
        unsigned y[2][4];
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            const unsigned* X = x + 4 * hf;
            const unsigned P = __builtin_amdgcn_perm(X[1], X[0], 0x05010400u), Q = __builtin_amdgcn_perm(X[1], X[0], 0x07030602u);
            const unsigned R = __builtin_amdgcn_perm(X[3], X[2], 0x05010400u), U = __builtin_amdgcn_perm(X[3], X[2], 0x07030602u);
            y[hf][0] = __builtin_amdgcn_perm(R, P, 0x05040100u);
            y[hf][1] = __builtin_amdgcn_perm(R, P, 0x07060302u);
            y[hf][2] = __builtin_amdgcn_perm(U, Q, 0x05040100u);
            y[hf][3] = __builtin_amdgcn_perm(U, Q, 0x07060302u);
        }
#pragma unroll
        for (int m = 0; m < nd; ++m)
            b[4 * k + m] = (long)(((unsigned long)y[1][m] << 32) | (unsigned long)y[0][m]);
    }
    // products a_i b_j, i + j <= S + 1 (1-based), summed per weight t = i + j (0-based: a_i b_{t - i}, i <= t),
    // one M-tile (real, then imaginary rows) at a time; the sum of weight t weighs 2^{1 + e - 7 (t + 2)} (M / 2 and
    // v carry 2^{-1} and 2^{-e}). Weights run in groups of 4, each group's four sums folded into the FP64 result
    // before the next group starts, so only four accumulators are live.
#pragma nounroll
    for (int mt = 0; mt < 2; ++mt) {
        long a[S];
#pragma unroll
        for (int i = 0; i < S; ++i) a[i] = adig[(i * 2 + mt) * 64 + lane];
        double s[4], err[4];
#pragma unroll
        for (int t0 = 0; t0 < S; t0 += 4) {
            const int tn = t0 + 4 < S ? t0 + 4 : S;
            v4i acc[4];
#pragma unroll
            for (int t = t0; t < tn; ++t) {
                acc[t - t0] = v4i{0, 0, 0, 0};
#pragma unroll
                for (int i = 0; i <= t; ++i)
                    acc[t - t0] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a[i], b[t - i], acc[t - t0], 0, 0, 0);
            }
            // the group exactly (|G| < 2^43, every Horner step exact in FP64), then one step of an exact TwoSum
            // chain with the error terms summed: the dd result is within ~2^-105 of the exact sum
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double G = (double)acc[0][r];
#pragma unroll
                for (int t = t0 + 1; t < tn; ++t) G = __builtin_fma(G, 128.0, (double)acc[t - t0][r]);
                const double g = __builtin_ldexp(G, 1 + e - 7 * (tn + 1));
                if (t0 == 0) {
                    s[r] = g;
                    err[r] = 0.0;
                } else {
                    const dd z = knuth2(s[r], g);
                    s[r] = z.hi;
                    err[r] += z.lo;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const dd v = quick2(s[r], err[r]);
            if (mt == 0) re[r] = v;
            else im[r] = v;
        }
    }
}

}  // namespace split
}  // namespace kbsrc::gpu
