// Specialised sub-FFT passes of the split-precision form (included by the split kernel tables): the compile-time
// pass of fixed_impl.hpp with its first two radix-4 stages replaced by the int8-MFMA radix-16 stage
// (split_core.hpp). The matrix stage reads its inputs from the pass's source and writes one LDS tile, the
// remaining Stockham stages run in place there and the last one writes the pass's destination. The same
// operations as the generic split pass and the global-memory route, so the same bits.
#pragma once
#include "fixed_impl.hpp"
#include "split_core.hpp"

namespace kbsrc::gpu {

namespace {

template <class RL, int G, int T, int... I>
__device__ __forceinline__ void mid_stages(const Lds<DDS, RL::N * G>& L, int tid, const OwnTw& tw,
                                           const double* oddc, const Prune& pr, std::integer_sequence<int, I...>) {
    (stage<DDS, RL, G, T, I + 2>(L, tid, tw, oddc, pr), ...);
}

// two waves per SIMD: the matrix stage's slicing needs up to ~150 VGPRs at 15 slices
template <class RL, int G, int T, int LD, int ST, int SPL>
__global__ __launch_bounds__(T, 2) void k_fixed_split(PassArgs a) {
    using A = DDS;
    constexpr int N = RL::N, CAP = N * G, NV = G * (N / 16);
    extern __shared__ double lds[];   // kPlanes * CAP doubles
    const Lds<A, CAP> L{lds};
    const int tid = threadIdx.x;
    const int wg = xcd_wg();
    const int task = wg / a.ngrp, grp = wg - task * a.ngrp;
    const int j0 = grp * G;
    const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
    auto jof = [&](int g) -> int { return j0 + g < a.nsub ? j0 + g : -1; };
    const PassView view = pass_view<LD>(a, io, task);

    // the radix-16 stage, its inputs straight from the pass's source: vector (g, p) = elements p + (N / 16) m of
    // local sub-FFT g; y[16 p + k] = w_N^{p k} DFT16, written to the LDS tile
    {
        constexpr int nv = N / 16;
        const int lane = tid & 63, q = lane >> 4, col = lane & 15;
        for (int base = (tid >> 6) * 16; base < NV; base += (T >> 6) * 16) {   // uniform over the wave
            const int v = base + col;
            const int g = v < NV ? v % G : 0, p = v < NV ? v / G : 0;
            const int j = v < NV ? jof(g) : -1;
            const bool live = j >= 0;
            double hi[8], lo[8];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (live) {
                    const Cx<A> c = pass_load<A, LD>(view, p + nv * (4 * q + i), j);
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
                    if (p != 0 && k != 0) z = c_mul<DDA>(z, gld32<DDA>(a.tw16, (unsigned)(p * 15 + k - 1)));
                    L.put((16 * p + k) * G + g, z);
                }
            }
        }
    }
    __syncthreads();
    Prune pr;
    if constexpr (ST == ST_CPLX) pr = Prune{a.wstart, a.wlen};
    const OwnTw tw{a.tw};
    mid_stages<RL, G, T>(L, tid, tw, a.oddc, pr, std::make_integer_sequence<int, RL::nst - 3>{});
    // the last stage, its outputs through the specialised pass's store (four-step twiddle, renormalisation, complex
    // buffer or output samples)
    last_stage<A, RL, G, T>(L, tid, tw, a.oddc, pr, jof,
                            [&](int e, int jj, Cx<A> v) { pass_store<A, ST>(a, view, e, jj, v); });
}

template <class RL, int G, int T, int LD, int ST, int SPL> hipError_t launch_split_one(const PassArgs& a, hipStream_t s) {
    static_assert(RL::nst >= 3 && RL::r[0] == 4 && RL::r[1] == 4, "the split stage replaces two radix-4 stages");
    static_assert(RL::hold(G, T) <= 8, "too many values held per thread");
    static_assert(LD != LD_FOLD && ST != ST_SPLIT, "no fused split passes");
    constexpr size_t bytes = (size_t)kPlanes<DDS> * RL::N * G * sizeof(double);
    static_assert(bytes <= 160 * 1024, "LDS tile too large");
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

}  // namespace
}  // namespace kbsrc::gpu
