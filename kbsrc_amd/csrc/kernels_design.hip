// Kernel design and plan tables on the GPU: the Kaiser-sinc taps, one thread per tap, their circular placement as
// the real input of the spectrum FFT, and the long twiddle and chirp tables of the FFT plans. Same scalar dd code
// as the host (design_math.hpp).
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

__device__ __forceinline__ void put_cx(double* out, long long i, dd c, dd s, int as_dd) {
    if (as_dd) {
        reinterpret_cast<double4*>(out)[i] = double4{c.hi, s.hi, c.lo, s.lo};
    } else {
        reinterpret_cast<double2*>(out)[i] = double2{c.hi, s.hi};
    }
}

__device__ __forceinline__ void get_pair(const double* t, long long j, dd& c, dd& s) {
    const double4 v = reinterpret_cast<const double4*>(t)[j];
    c = dd{v.x, v.y};
    s = dd{v.z, v.w};
}

// CisTable's two short tables (src/dsp/fft.hpp:29-35) for length N and block B: a[j] = cis(j B, N) for j < J,
// b[l] = cis(l, N) for l < B, each {c.hi, c.lo, s.hi, s.lo}; the host's dd code (design_math.hpp), the same bits
__global__ __launch_bounds__(256) void k_cis_short(long long N, long long B, long long J, double* a, double* b) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= J + B) return;
    dd c, s;
    double* o;
    if (t < J) {
        cis(t * B, N, c, s);
        o = a + 4 * t;
    } else {
        cis(t - J, N, c, s);
        o = b + 4 * (t - J);
    }
    reinterpret_cast<double4*>(o)[0] = double4{c.hi, c.lo, s.hi, s.lo};
}

// CisTable::get (src/dsp/fft.hpp:36-42) for every element of a table
__global__ __launch_bounds__(256) void k_cis_table(TableArgs a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    long long k = a.mode == TAB_FOUR ? -((i / a.N2) * (i % a.N2)) : -i;
    k %= a.N;
    if (k < 0) k += a.N;
    const long long j = k / a.B, l = k - j * a.B;
    dd c, s;
    if (l == 0) {
        get_pair(a.a, j, c, s);
    } else if (j == 0) {
        get_pair(a.b, l, c, s);
    } else {
        dd ac, as, bc, bs;
        get_pair(a.a, j, ac, as);
        get_pair(a.b, l, bc, bs);
        cis_compose(ac, as, bc, bs, c, s);
    }
    put_cx(a.out, i, c, s, a.as_dd);
}

__global__ __launch_bounds__(256) void k_chirp(long long n, long long m, int as_dd, double* bc, double* v) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long e = (long long)((__int128)t * t % (2 * (__int128)n));
    dd c, s;
    cis(e, 2 * n, c, s);
    put_cx(bc, t, c, s, as_dd);
    put_cx(v, t, c, s, 1);
    if (t) put_cx(v, m - t, c, s, 1);
}

__global__ __launch_bounds__(256) void k_div_m(const double* V, long long m, int as_dd, double* out) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const double4 v = reinterpret_cast<const double4*>(V)[q];
    put_cx(out, q, dddiv1(dd{v.x, v.z}, (double)m), dddiv1(dd{v.y, v.w}, (double)m), as_dd);
}

dim3 grid_for(long long items) { return dim3((unsigned)((items + 255) / 256)); }

}  // namespace

hipError_t launch_cis_short(long long N, long long B, long long J, double* a, double* b, hipStream_t s) {
    if (N < 1 || B < 1 || J < 1 || (J + B + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cis_short, grid_for(J + B), dim3(256), 0, s, N, B, J, a, b);
    return hipGetLastError();
}

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

hipError_t launch_cis_table(const TableArgs& a, hipStream_t s) {
    if (a.count < 0 || a.N < 1 || a.B < 1 || a.N2 < 1 || (a.count + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    if (a.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cis_table, grid_for(a.count), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_chirp(long long n, long long m, int as_dd, double* bc, double* v, hipStream_t s) {
    if (n < 1 || m < 2 * n - 1 || (n + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_chirp, grid_for(n), dim3(256), 0, s, n, m, as_dd, bc, v);
    return hipGetLastError();
}

hipError_t launch_div_m(const double* V, long long m, int as_dd, double* out, hipStream_t s) {
    if (m < 1 || (m + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_div_m, grid_for(m), dim3(256), 0, s, V, m, as_dd, out);
    return hipGetLastError();
}

hipError_t launch_hi(const double* pairs, long long n, double* out, hipStream_t s) {
    if (n < 1) return hipSuccess;
    hipLaunchKernelGGL(k_hi, grid_for(n), dim3(256), 0, s, pairs, n, out);
    return hipGetLastError();
}

}  // namespace kbsrc::gpu
