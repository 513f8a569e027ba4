// Specialised sub-FFT passes of the split-precision form: the compile-time pass of kernels_fixed.hip with its first
// two radix-4 stages replaced by the int8-MFMA radix-16 stage (split_core.hpp), for the sub-FFT lengths of the
// benchmark conversions (96 kHz -> 44.1 kHz forward rows, 320 points; 44.1 kHz -> 48 kHz inverse columns, 224).
// The load and store are the specialised pass's; the matrix stage reads one LDS tile and writes a second, the
// remaining Stockham stages run in place there. The same operations as the generic split pass and the
// global-memory route, so the same bits.
#include "fixed_impl.hpp"
#include "split_core.hpp"

namespace kbsrc::gpu {

namespace {

template <class RL, int G, int T, int... I>
__device__ __forceinline__ void tail_stages(const Lds<DDS, RL::N * G>& L, int tid, const double* tw,
                                            const double* oddc, const Prune& pr, std::integer_sequence<int, I...>) {
    (stage<DDS, RL, G, T, I + 2>(L, tid, tw, oddc, pr), ...);
}

// two waves per SIMD: the matrix stage's slicing needs up to ~200 VGPRs at 15 slices
template <class RL, int G, int T, int LD, int ST, int SPL>
__global__ __launch_bounds__(T, 2) void k_fixed_split(PassArgs a) {
    using A = DDS;
    constexpr int N = RL::N, CAP = N * G, NV = G * (N / 16);
    extern __shared__ double lds[];   // two tiles of kPlanes * CAP doubles
    const Lds<A, CAP> L{lds}, L2{lds + kPlanes<A> * CAP};
    const int tid = threadIdx.x;
    const int wg = xcd_wg();
    const int task = wg / a.ngrp, grp = wg - task * a.ngrp;
    const int j0 = grp * G;
    const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
    auto jof = [&](int g) -> int { return j0 + g < a.nsub ? j0 + g : -1; };

    // load, lanes sub-FFT-fastest (as the specialised pass)
    for (int w = tid; w < G * N; w += T) {
        const int g = w % G, e = w / G;
        const int j = jof(g);
        if (j < 0) continue;
        const long long i = e * a.lde + (long long)j * a.ldg;
        Cx<A> v;
        if constexpr (LD == LD_REAL_F64) {
            const double2 pr = reinterpret_cast<const double2*>(a.src + io.in_off)[i];
            v = {A::lift(pr.x), A::lift(pr.y)};
        } else {
            v = gld<A>(a.src + task * a.src_tstride, i);
        }
        L.put(e * G + g, v);
    }
    __syncthreads();

    // the radix-16 stage: vector (g, p) = elements p + (N / 16) m of local sub-FFT g; y[16 p + k] = w_N^{p k} DFT16
    {
        constexpr int nv = N / 16;
        const int lane = tid & 63, q = lane >> 4, col = lane & 15;
        for (int base = (tid >> 6) * 16; base < NV; base += (T >> 6) * 16) {   // uniform over the wave
            const int v = base + col;
            const int g = v < NV ? v % G : 0, p = v < NV ? v / G : 0;
            const bool live = v < NV && jof(g) >= 0;
            double hi[8], lo[8];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (live) {
                    const Cx<A> c = L.get((p + nv * (4 * q + i)) * G + g);
                    hi[2 * i] = c.re.hi;
                    hi[2 * i + 1] = c.im.hi;
                    lo[2 * i] = c.re.lo;
                    lo[2 * i + 1] = c.im.lo;
                } else {
                    hi[2 * i] = hi[2 * i + 1] = lo[2 * i] = lo[2 * i + 1] = 0.0;
                }
            }
            dd re[4], im[4];
            split::split_core<SPL>(hi, lo, a.adig, lane, re, im);
            if (live) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int k = 4 * q + r;
                    cx<dd> z{re[r], im[r]};
                    if (p != 0 && k != 0) z = c_mul<DDA>(z, gld<DDA>(a.tw16, (long long)p * 15 + k - 1));
                    L2.put((16 * p + k) * G + g, z);
                }
            }
        }
    }
    __syncthreads();
    Prune pr;
    if constexpr (ST == ST_CPLX) {
        if (a.zkeep >= 0) pr = Prune{a.zkeep, a.zh, a.ste, a.stg, j0};
    }
    tail_stages<RL, G, T>(L2, tid, a.tw, a.oddc, pr, std::make_integer_sequence<int, RL::nst - 2>{});

    // store (as the specialised pass): four-step twiddle, renormalisation, complex buffer or output samples
    for (int w = tid; w < G * N; w += T) {
        const int g = w % G, e = w / G;
        const int j = jof(g);
        if (j < 0) continue;
        const long long jg = j;
        const long long i = e * a.ste + jg * a.stg;
        if (ST == ST_CPLX && z_unused(a, i)) continue;
        Cx<A> v = L2.get(e * G + g);
        if (a.twmode && e != 0 && (a.twmode == 1 || jg != 0)) v = c_mul<A>(v, gld<A>(a.tw4, i));
        if (a.renorm) v = renorm_dd<A>(v);
        if constexpr (ST == ST_CPLX) {
            gst<A>(a.dst + task * a.dst_tstride, i, v);
        } else {
            double* o = a.out + io.out_off;
            const long long x0 = 2 * i - a.margP;
            if (x0 >= 0 && x0 < a.hopP) {
                const long long d = io.blk * a.hopP + x0 - a.o;
                if (d >= 0 && d < io.nout) o[d] = v.re.hi + v.re.lo;
            }
            const long long x1 = x0 + 1;
            if (x1 >= 0 && x1 < a.hopP) {
                const long long d = io.blk * a.hopP + x1 - a.o;
                if (d >= 0 && d < io.nout) {
                    const dd ni = A::neg(v.im);
                    o[d] = ni.hi + ni.lo;
                }
            }
        }
    }
}

template <class RL, int G, int T, int LD, int ST, int SPL> hipError_t launch_split_one(const PassArgs& a, hipStream_t s) {
    static_assert(RL::nst >= 3 && RL::r[0] == 4 && RL::r[1] == 4, "the split stage replaces two radix-4 stages");
    static_assert(RL::hold(G, T) <= 8, "too many values held per thread");
    static_assert(LD != LD_FOLD && ST != ST_SPLIT, "no fused split passes");
    constexpr size_t bytes = (size_t)2 * kPlanes<DDS> * RL::N * G * sizeof(double);
    static_assert(bytes <= 160 * 1024, "LDS tiles too large");
    static const hipError_t attr = bytes > 64 * 1024
        ? hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fixed_split<RL, G, T, LD, ST, SPL>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
        : hipSuccess;
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL((k_fixed_split<RL, G, T, LD, ST, SPL>), dim3(a.ngrp * a.ntask), dim3(T), bytes, s, a);
    return hipGetLastError();
}

#define KB_SPLIT(G, T, LD, ST, SPL, ...) \
    SplitFixedEntry{Radices<__VA_ARGS__>::N, LD, ST, SPL, G, T, &launch_split_one<Radices<__VA_ARGS__>, G, T, LD, ST, SPL>}

// first matching entry wins; the others stay for KBSRC_SPLIT_VARIANT sweeps. Measured on the 180 s 96 kHz ->
// 44.1 kHz conversion (ms per conversion, plain 1.90-2.00): G = 2, T = 192: 2.01-2.02 (6 slices), 2.14-2.15 (15);
// G = 4, T = 256: 2.01, 2.21-2.22; G = 4, T = 320: 1.99-2.01, 2.50-2.51 (one workgroup per CU at 15 slices).
const SplitFixedEntry kSplit[] = {
    KB_SPLIT(2, 192, LD_CPLX, ST_CPLX, 6, 4, 4, 4, 5),
    KB_SPLIT(2, 192, LD_CPLX, ST_CPLX, 7, 4, 4, 4, 5),
    KB_SPLIT(2, 192, LD_CPLX, ST_CPLX, 8, 4, 4, 4, 5),
    KB_SPLIT(2, 192, LD_CPLX, ST_CPLX, 15, 4, 4, 4, 5),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 6, 4, 4, 4, 5),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 15, 4, 4, 4, 5),
    KB_SPLIT(4, 320, LD_CPLX, ST_CPLX, 6, 4, 4, 4, 5),
    KB_SPLIT(4, 320, LD_CPLX, ST_CPLX, 15, 4, 4, 4, 5),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 6, 4, 4, 2, 7),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 7, 4, 4, 2, 7),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 8, 4, 4, 2, 7),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 15, 4, 4, 2, 7),
};

}  // namespace

const SplitFixedEntry* find_split_fixed(int n, int ld, int st, int slices) {
    const char* v = std::getenv("KBSRC_SPLIT_VARIANT");   // sweep: take the k-th matching entry
    int skip = v ? std::atoi(v) : 0;
    for (const SplitFixedEntry& e : kSplit)
        if (e.n == n && e.ld == ld && e.st == st && e.slices == slices && skip-- <= 0) return &e;
    return nullptr;
}

}  // namespace kbsrc::gpu
