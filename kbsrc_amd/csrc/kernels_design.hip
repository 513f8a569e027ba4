// Kernel design on the GPU: the Kaiser-sinc taps, one thread per tap, and their circular placement as the real
// input of the spectrum FFT. Same scalar dd code as the host (design_math.hpp).
#include "kernels.hpp"
#include "design_math.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {

namespace {

__global__ __launch_bounds__(256) void k_taps(TapArgs a) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n > a.L) return;
    const dd v = kaiser_tap(n, a.L, a.a, a.b, a.beta, dd{a.ratio_hi, a.ratio_lo}, dd{a.i0_hi, a.i0_lo});
    reinterpret_cast<double2*>(a.h)[n] = double2{v.hi, v.lo};
}

// r[0] = h[0]; for i = 1..Lk: r[i] = h[i], then r[n - i] = h[i] (src/design.cpp:42-43). Position p keeps the
// value of the last write in that order: h[max(p, n - p)] where both apply.
__global__ __launch_bounds__(256) void k_circ(const double* h, long long Lk, long long n, double* r) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    long long src = -1;
    if (p == 0) {
        src = 0;
    } else {
        if (p <= Lk) src = p;
        if (n - p <= Lk && n - p > src) src = n - p;
    }
    const double2 v = src >= 0 ? reinterpret_cast<const double2*>(h)[src] : double2{0.0, 0.0};
    reinterpret_cast<double2*>(r)[p] = v;
}

// H in the engine's element format: dd pairs as they are, or FP64 = the pair's hi (src/dsp/fft.hpp:16)
__global__ __launch_bounds__(256) void k_hi(const double* pairs, long long n, double* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pairs[2 * i];
}

dim3 grid_for(long long items) { return dim3((unsigned)((items + 255) / 256)); }

}  // namespace

hipError_t launch_taps(const TapArgs& a, hipStream_t s) {
    if (a.L < 0 || (a.L + 256) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_taps, grid_for(a.L + 1), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_circ(const double* h, long long Lk, long long n, double* r, hipStream_t s) {
    if (n < 1 || Lk < 0 || Lk >= n || (n + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_circ, grid_for(n), dim3(256), 0, s, h, Lk, n, r);
    return hipGetLastError();
}

hipError_t launch_hi(const double* pairs, long long n, double* out, hipStream_t s) {
    if (n < 1) return hipSuccess;
    hipLaunchKernelGGL(k_hi, grid_for(n), dim3(256), 0, s, pairs, n, out);
    return hipGetLastError();
}

}  // namespace kbsrc::gpu
