// Split-precision (Ozaki) form of a radix-16 DFT stage on the int8 matrix cores. Not on the bit-exact path: the
// plain engine computes every butterfly in dd, bit for bit as the reference does; the split form of a conversion
// runs the first two radix-4 stages of its sub-FFTs (those whose length 16 divides) as this stage instead, and a
// standalone radix-16 DFT measures the stage against the plain dd one.
//
// y = F x for vectors x of 16 complex dd values, F[k][n] = e^{-2 pi i k n / 16}, as one real 32 x 32 product
// [Re y; Im y] = M x' with x' = (Re x[0], Im x[0], Re x[1], ...). Each vector is scaled by a power of two to |z| < 1/2 and cut into S base-128
// digits in [-64, 64], z = sum_j b_j 2^{-7 (j + 1)}; M / 2 likewise into a_i (on the host). The products a_i b_j
// with i + j <= S + 1 run on v_mfma_i32_16x16x32_i8 (exact int32 sums of 32 terms), summed per weight
// 2^{-7 (i + j)} in int32 (exact: at most S products of |a|, |b| <= 65 over 32 terms), and the weighted sums are
// recombined in FP64 (exact groups, one TwoSum chain) into dd. Error: the digits kept (about 2^{-7 S} of the
// vector's largest value) and the products dropped. FP64 dd keeps the stage's twiddle multiply.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bfly.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {

namespace {

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
    // S base-128 digits of each real v = (hi + lo) 2^{-e}, most significant first, by nearest-integer peeling:
    // t = 128 h, d = rint(t), h = t - d, all exact, so v = sum_j d_j 128^{-(j+1)} + remainder and every |d_j| <= 64.
    // The low part joins the remainder by an exact TwoSum every 6 digits (before its scaled size reaches 2^-12);
    // the last digit is rounded, so the remainder dropped is at most (1/2 + 2^-19) 128^{-S}. Byte i of b[j] is
    // digit j of real 8q + i.
    long b[S];
#pragma unroll
    for (int j = 0; j < S; ++j) b[j] = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double h = __builtin_ldexp(hi[i], -e), l = __builtin_ldexp(lo[i], -e);
#pragma unroll
        for (int j = 0; j < S; ++j) {
            if (j > 0 && j % 6 == 0) {
                const dd z = knuth2(h, __builtin_ldexp(l, 42));
                h = z.hi;
                l = z.lo;
            }
            const double t = h * 128.0;
            const double d = __builtin_rint(t);
            h = t - d;
            b[j] |= (long)(uint8_t)(int8_t)(int)d << (8 * i);
        }
    }
    // products a_i b_j, i + j <= S + 1 (1-based), summed per weight t = i + j, one M-tile (real, then imaginary
    // rows) at a time; acc[t] weighs 2^{1 + e - 7 (t + 2)} (M / 2 and v carry 2^{-1} and 2^{-e})
#pragma nounroll
    for (int mt = 0; mt < 2; ++mt) {
        v4i acc[S];
#pragma unroll
        for (int t = 0; t < S; ++t) acc[t] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const long a = adig[(i * 2 + mt) * 64 + lane];
#pragma unroll
            for (int j = 0; i + j + 2 <= S + 1; ++j)
                acc[i + j] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a, b[j], acc[i + j], 0, 0, 0);
        }
        // exact groups of 4 weights (|G| < 2^43, every Horner step exact in FP64), then an exact TwoSum chain with
        // the error terms summed: the dd result is within ~2^-105 of the exact sum
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double s = 0.0, err = 0.0;
#pragma unroll
            for (int t0 = 0; t0 < S; t0 += 4) {
                const int tn = t0 + 4 < S ? t0 + 4 : S;
                double G = (double)acc[t0][r];
#pragma unroll
                for (int t = t0 + 1; t < tn; ++t) G = __builtin_fma(G, 128.0, (double)acc[t][r]);
                const double g = __builtin_ldexp(G, 1 + e - 7 * (tn + 1));
                if (t0 == 0) {
                    s = g;
                } else {
                    const dd z = knuth2(s, g);
                    s = z.hi;
                    err += z.lo;
                }
            }
            const dd v = quick2(s, err);
            if (mt == 0) re[r] = v;
            else im[r] = v;
        }
    }
}

// the 8 reals of lane l's share of vector x (16 dd elements {re.hi, im.hi, re.lo, im.lo}, element n at x + n * step
// elements): elements 4 q .. 4 q + 3, re and im, k = 2 (n - 4 q) + part, two 16-byte loads each; zeros when the
// lane's column has no vector
__device__ __forceinline__ void load_reals(const double* x, long long step, bool live, int q, double (&hi)[8],
                                           double (&lo)[8]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (live) {
            const double2* e = reinterpret_cast<const double2*>(x + (4 * q + i) * step * 4);
            const double2 h = e[0], l = e[1];
            hi[2 * i] = h.x;
            hi[2 * i + 1] = h.y;
            lo[2 * i] = l.x;
            lo[2 * i + 1] = l.y;
        } else {
            hi[2 * i] = hi[2 * i + 1] = lo[2 * i] = lo[2 * i + 1] = 0.0;
        }
    }
}

template <int S>
__global__ __launch_bounds__(256) void k_split_dft16(const double* __restrict__ x, long long nvec,
                                                     const long* __restrict__ adig, double* __restrict__ y) {
    const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15;
    const long long v = ((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16 + col;
    const bool live = v < nvec;
    double hi[8], lo[8];
    load_reals(x + (live ? v * 64 : 0), 1, live, q, hi, lo);
    dd re[4], im[4];
    split_core<S>(hi, lo, adig, lane, re, im);
    if (!live) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) gst<DDA>(y, v * 16 + 4 * q + r, cx<dd>{re[r], im[r]});
}

// The first two radix-4 Stockham stages of a sub-FFT of length n (16 | n) as one radix-16 stage: for each sub-FFT
// t (elements t n .. t n + n - 1) and column p < n / 16, y[t n + 16 p + k] = w_n^{p k} DFT16(x[t n + p + (n / 16) m])[k],
// the twiddle w_n^{p k} = tw16[p * 15 + k - 1] applied in dd (accurate add) for p k != 0.
template <int S>
__global__ __launch_bounds__(256) void k_split16_stage(const double* __restrict__ x, double* __restrict__ y, int n,
                                                       long long nvec, const double* __restrict__ tw16,
                                                       const long* __restrict__ adig) {
    const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15;
    const long long v = ((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16 + col;
    const bool live = v < nvec;
    const int nv = n >> 4;
    const long long t = live ? v / nv : 0;
    const int p = live ? (int)(v - t * nv) : 0;
    double hi[8], lo[8];
    load_reals(x + (t * n + p) * 4, nv, live, q, hi, lo);
    dd re[4], im[4];
    split_core<S>(hi, lo, adig, lane, re, im);
    if (!live) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = 4 * q + r;
        cx<dd> z{re[r], im[r]};
        if (p != 0 && k != 0) z = c_mul<DDA>(z, gld<DDA>(tw16, (long long)p * 15 + k - 1));
        gst<DDA>(y, t * n + 16 * p + k, z);
    }
}

template <int S> hipError_t launch_s(const double* x, long long nvec, const long* adig, double* y, hipStream_t s) {
    const long long waves = (nvec + 15) / 16, blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_split_dft16<S>), dim3((unsigned)blocks), dim3(256), 0, s, x, nvec, adig, y);
    return hipGetLastError();
}

template <int S>
hipError_t launch_st(const double* x, double* y, int n, long long nsub, const double* tw16, const long* adig,
                     hipStream_t s) {
    const long long nvec = nsub * (n / 16), waves = (nvec + 15) / 16, blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_split16_stage<S>), dim3((unsigned)blocks), dim3(256), 0, s, x, y, n, nvec, tw16, adig);
    return hipGetLastError();
}

}  // namespace

bool split_slices_ok(int slices) {
    return (slices >= 4 && slices <= 8) || slices == 10 || slices == 12 || slices == 14 || slices == 15 ||
           slices == 16;
}

hipError_t launch_split_dft16(const double* x, long long nvec, int slices, const long* adig, double* y,
                              hipStream_t s) {
    if (nvec < 1) return hipSuccess;
    switch (slices) {
        case 4: return launch_s<4>(x, nvec, adig, y, s);
        case 5: return launch_s<5>(x, nvec, adig, y, s);
        case 6: return launch_s<6>(x, nvec, adig, y, s);
        case 7: return launch_s<7>(x, nvec, adig, y, s);
        case 8: return launch_s<8>(x, nvec, adig, y, s);
        case 10: return launch_s<10>(x, nvec, adig, y, s);
        case 12: return launch_s<12>(x, nvec, adig, y, s);
        case 14: return launch_s<14>(x, nvec, adig, y, s);
        case 15: return launch_s<15>(x, nvec, adig, y, s);
        case 16: return launch_s<16>(x, nvec, adig, y, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_split16_stage(const double* x, double* y, int n, long long nsub, int slices, const double* tw16,
                                const long* adig, hipStream_t s) {
    if (n < 16 || n % 16 || nsub < 1) return hipErrorInvalidValue;
    switch (slices) {
        case 4: return launch_st<4>(x, y, n, nsub, tw16, adig, s);
        case 5: return launch_st<5>(x, y, n, nsub, tw16, adig, s);
        case 6: return launch_st<6>(x, y, n, nsub, tw16, adig, s);
        case 7: return launch_st<7>(x, y, n, nsub, tw16, adig, s);
        case 8: return launch_st<8>(x, y, n, nsub, tw16, adig, s);
        case 10: return launch_st<10>(x, y, n, nsub, tw16, adig, s);
        case 12: return launch_st<12>(x, y, n, nsub, tw16, adig, s);
        case 14: return launch_st<14>(x, y, n, nsub, tw16, adig, s);
        case 15: return launch_st<15>(x, y, n, nsub, tw16, adig, s);
        case 16: return launch_st<16>(x, y, n, nsub, tw16, adig, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace kbsrc::gpu
