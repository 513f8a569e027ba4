// Split-precision (Ozaki) form of a radix-16 DFT stage on the int8 matrix cores (the core is split_core.hpp): the
// standalone stage over vectors of 16 complex dd values, measured against the plain dd stage, and the same stage
// as the first two radix-4 stages of sub-FFTs stored one after another (the global-memory route of the split form
// of a conversion; sub-FFTs that fit LDS take it inside their pass, kernels.hip). Not on the bit-exact path: the
// plain engine computes every butterfly in dd, bit for bit as the reference does. FP64 dd keeps the stage's
// twiddle multiply.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bfly.hpp"
#include "split_core.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {

namespace {

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
    split::split_core<S>(hi, lo, adig, lane, re, im);
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
    split::split_core<S>(hi, lo, adig, lane, re, im);
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
