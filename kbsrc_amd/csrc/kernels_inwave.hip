// The plain dd radix-16 stage (two radix-4 Stockham stages of 16-point sub-FFTs, src/dsp/fft.hpp:86-151) with the
// exchange between its stages done inside the wave by v_permlane32_swap instead of through LDS: measured beside the
// LDS pass (kbsrc_plain_dft16), same operations, same bits.
#include "fixed_impl.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {
namespace {

// lanes 32..63 of a trade places with lanes 0..31 of b (v_permlane32_swap_b32 on each 32-bit half)
__device__ __forceinline__ void swap_halves(double& a, double& b) {
    const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
    const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ void swap_halves(Cx<DDS>& a, Cx<DDS>& b) {
    swap_halves(a.re.hi, b.re.hi);
    swap_halves(a.re.lo, b.re.lo);
    swap_halves(a.im.hi, b.im.hi);
    swap_halves(a.im.lo, b.im.lo);
}

// Lanes l and l + 32 of a wave share vector l. The lane of role L = l / 32 runs the stage-0 butterflies b = L + 2j
// (j = 0, 1: inputs b + 4t from global memory, outputs Y[j][u] for elements 4b + u), then the stage-1 butterflies
// b' = L + 2m, whose inputs t = 0..3 are output u = b' of stage-0 butterfly t. Swapping register halves between
// Y[j][2m] and Y[j][2m + 1] leaves, in every lane, Y[j][2m + a] = output u = L + 2m of butterfly t = 2j + a: the
// stage-1 inputs at register indices that do not depend on the lane.
__global__ __launch_bounds__(256) void k_dft16_inwave(const double* __restrict__ x, long long nvec,
                                                      double* __restrict__ y, const double* __restrict__ tw,
                                                      const double* __restrict__ oddc) {
    using A = DDS;
    const int lane = (int)(threadIdx.x & 63), L = lane >> 5;
    const long long v = ((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 32 + (lane & 31);
    const long long vl = v < nvec ? v : nvec - 1;   // every lane takes part in the swaps
    const double* xv = x + vl * 64;
    Cx<A> Y[2][4];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int b = L + 2 * j;
        Cx<A> in[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) in[t] = gld<A>(xv, b + 4 * t);
        butterfly<A, 4>(in, Y[j], b != 0, tw + (long long)b * 3 * kElt<A>, oddc);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        swap_halves(Y[j][0], Y[j][1]);
        swap_halves(Y[j][2], Y[j][3]);
    }
    if (v >= nvec) return;
    double* yv = y + v * 64;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const int bp = L + 2 * m;
        Cx<A> in[4], out[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) in[t] = Y[t >> 1][2 * m + (t & 1)];
        butterfly<A, 4>(in, out, false, tw, oddc);
#pragma unroll
        for (int u = 0; u < 4; ++u) gst<A>(yv, bp + 4 * u, renorm_dd<A>(out[u]));
    }
}

}  // namespace

hipError_t launch_dft16_inwave(const double* x, long long nvec, double* y, const double* tw, const double* oddc,
                               hipStream_t s) {
    if (nvec < 1) return hipSuccess;
    const long long blocks = (nvec + 127) / 128;   // 4 waves of 32 vectors
    if (blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_dft16_inwave, dim3((unsigned)blocks), dim3(256), 0, s, x, nvec, y, tw, oddc);
    return hipGetLastError();
}

}  // namespace kbsrc::gpu
