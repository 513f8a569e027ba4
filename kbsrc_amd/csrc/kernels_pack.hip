// Output packing on the GPU: planar float64 samples -> interleaved little-endian WAV sample bytes, without
// dither (src/io/wav.cpp:52-86,235-250). Integer formats scale by 2^(8B-1), round to nearest even (the
// reference's nearbyint under the default rounding mode), clip to [-2^(8B-1), 2^(8B-1) - 1] counting the clipped
// samples, and write NaN as zero.
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace kbsrc::gpu {

namespace {

template <int B> __device__ __forceinline__ void put_int(unsigned char* p, double y, unsigned long long* clipped) {
    constexpr double sc = B == 2 ? 32768.0 : B == 3 ? 8388608.0 : 2147483648.0;
    double v = __builtin_rint(y * sc);
    bool clip = false;
    if (v > sc - 1) {
        v = sc - 1;
        clip = true;
    } else if (v < -sc) {
        v = -sc;
        clip = true;
    }
    if (clip) atomicAdd(clipped, 1ULL);
    const long long iv = v != v ? 0 : (long long)v;   // NaN is written as zero
#pragma unroll
    for (int t = 0; t < B; ++t) p[t] = (unsigned char)((iv >> (8 * t)) & 255);
}

__global__ __launch_bounds__(256) void k_pack(PackArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;   // sample of the chunk, interleaved
    if (idx >= a.frames * a.ch) return;
    const long long i = a.frame0 + idx / a.ch;
    const int c = (int)(idx % a.ch);
    const double y = a.y[c * a.nout + i];
    unsigned char* p = a.out + idx * a.bps;
    switch (a.format) {
        case PACK_F64: {
            *reinterpret_cast<double*>(p) = y;   // 8-byte aligned: idx * 8
            break;
        }
        case PACK_F32: *reinterpret_cast<float*>(p) = (float)y; break;
        case PACK_S32: put_int<4>(p, y, a.clipped); break;
        case PACK_S24: put_int<3>(p, y, a.clipped); break;
        default: put_int<2>(p, y, a.clipped); break;
    }
}

}  // namespace

int pack_bytes(int format) {
    switch (format) {
        case PACK_F64: return 8;
        case PACK_F32: case PACK_S32: return 4;
        case PACK_S24: return 3;
        case PACK_S16: return 2;
        default: return 0;
    }
}

hipError_t launch_pack(PackArgs a, hipStream_t s) {
    a.bps = pack_bytes(a.format);
    const long long n = a.frames * a.ch;
    if (!a.bps || n < 0 || (n + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace kbsrc::gpu
