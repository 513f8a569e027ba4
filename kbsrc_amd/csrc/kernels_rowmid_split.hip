// The split-precision form of the fused forward row pass + split/multiply/fold kernel (kernels_rowmid.hip): the
// workgroup's row sub-FFTs run as the split form's specialised row pass does (split_fixed_impl.hpp: the first two
// radix-4 stages as one int8-MFMA radix-16 stage reading the column pass's output, the other stages in place in
// LDS), the last stage too in LDS, then the same split/fold step over the rows (rowmid_impl.hpp). Same operations
// as the split row pass followed by the split/fold kernel, so the same bits.
#include <cstdlib>
#include <numeric>

#include "rowmid_impl.hpp"
#include "split_fixed_impl.hpp"

namespace kbsrc::gpu {

namespace {

template <class RL, int G, int T, int... I>
__device__ __forceinline__ void later_stages(const Lds<DDS, RL::N * G>& L, int tid, const OwnTw& tw,
                                             const double* oddc, const Prune& pr, std::integer_sequence<int, I...>) {
    (stage<DDS, RL, G, T, I + 2>(L, tid, tw, oddc, pr), ...);
}

// two waves per SIMD, as the split row pass: the matrix stage's slicing needs up to ~150 VGPRs at 15 slices
template <class RL, int N1, int TT, int U, int T, int SPL, bool RES>
__global__ __launch_bounds__(T, 2) void k_rowmid_split(PassArgs a, MidArgs m, int ngrp, int tgroup) {
    using A = DDS;
    using RU = RowUnits<N1, TT, U>;
    constexpr int N = RL::N, G = RU::G, CAP = N * G, NV = G * (N / 16);
    extern __shared__ double lds[];
    const Lds<A, CAP> L{lds};
    const int tid = threadIdx.x;
    int task, grp;
    rowmid_order(xcd_wg(), a.ntask, ngrp, tgroup, task, grp);
    const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
    const PassView view = pass_view<LD_CPLX>(a, io, task);
    auto jof = [&](int g) -> int { return RU::row(g, grp); };

    // the radix-16 stage, its inputs straight from the column pass's output: vector (g, p) = elements p + (N / 16) m
    // of local row g; y[16 p + k] = w_N^{p k} DFT16, written to the LDS tile (k_fixed_split)
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
                    const Cx<A> c = pass_load<A, LD_CPLX>(view, p + nv * (4 * q + i), j);
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
    const Prune pr{a.wstart, a.wlen};
    const OwnTw tw{a.tw};
    later_stages<RL, G, T>(L, tid, tw, a.oddc, pr, std::make_integer_sequence<int, RL::nst - 2>{});
    rowmid_tail<A, RL, N1, TT, U, T, false, RES>(L, a, m, task, grp, tid);
}

template <class K> hipError_t lds_attr(K kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

template <class RL, int N1, int TT, int U, int T, int SPL, int TG, bool RES>
hipError_t launch_rowmid_split_one(const PassArgs& a, const MidArgs& m, hipStream_t s) {
    using RU = RowUnits<N1, TT, U>;
    static_assert(N1 % TT == 0, "a unit is whole residue classes of the rows");
    static_assert(RL::nst >= 3 && RL::r[0] == 4 && RL::r[1] == 4, "the split stage replaces two radix-4 stages");
    static_assert(RL::hold(RU::G, T) <= 8, "too many values held per thread");
    constexpr size_t bytes = (size_t)kPlanes<DDS> * RL::N * RU::G * sizeof(double);
    static_assert(bytes <= 160 * 1024, "LDS tile too large");
    if (a.sd.n != RL::N || a.sd.nst != RL::nst || a.nsub != N1 || a.twmode != 0 || a.split != SPL || !a.adig ||
        !a.tw16 || m.Jg || m.Mb != 2 * m.hinv || m.h != N1 * RL::N || std::gcd(N1, m.hinv) != TT || a.ntask < 1 ||
        m.ntask != a.ntask || (RES && (!m.Hr || !m.wfr || !m.wir || m.Lr != RU::NQ * RL::N + 1 ||
                (long long)m.Lr * TT * 32 > 0xFFFFFFFFLL)) ||
        m.Mb > (int)kDivSmallMax || m.h > (int)kDivSmallMax)   // the bin indices of the tail (div_small)
        return hipErrorInvalidValue;
    for (int i = 0; i < RL::nst; ++i)
        if (a.sd.st[i].r != RL::r[i] || a.sd.st[i].s != RL::stride(i)) return hipErrorInvalidValue;
    static const hipError_t attr = lds_attr(&k_rowmid_split<RL, N1, TT, U, T, SPL, RES>, bytes);
    if (attr != hipSuccess) return attr;
    const int ngrp = (RU::nunit + U - 1) / U;
    const long long nwg = (long long)a.ntask * ngrp;
    if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
    int tgroup = TG;
    if (const char* v = std::getenv("KBSRC_ROWMID_TGROUP")) tgroup = std::max(1, std::atoi(v));   // tuning knob
    hipLaunchKernelGGL((k_rowmid_split<RL, N1, TT, U, T, SPL, RES>), dim3((unsigned)nwg), dim3(T), bytes, s, a, m, ngrp,
                       std::min(tgroup, a.ntask));
    return hipGetLastError();
}

struct RowMidSplitEntry {
    int n2, n1, hinv, slices;
    hipError_t (*launch)(const PassArgs&, const MidArgs&, hipStream_t);
};

#define KB_ROWMID_SPLIT(N1, HINV, U, T, SPL, TG, RES, ...)                                                         \
    RowMidSplitEntry{Radices<__VA_ARGS__>::N, N1, HINV, SPL,                                                       \
                     &launch_rowmid_split_one<Radices<__VA_ARGS__>, N1, std::gcd(N1, HINV), U, T, SPL, TG, RES>}

// The plain form's benchmark shape (stereo 96 kHz -> 44.1 kHz exact: 320-point rows, row pairs), at the slice counts
// of the split form's sweep; the workgroup shape of the split row pass (4 rows on 256 threads).
const RowMidSplitEntry kRowMidSplit[] = {
    KB_ROWMID_SPLIT(315, 46305, 2, 256, 6, 8, true, 4, 4, 4, 5),
    KB_ROWMID_SPLIT(315, 46305, 2, 256, 7, 8, true, 4, 4, 4, 5),
    KB_ROWMID_SPLIT(315, 46305, 2, 256, 8, 8, true, 4, 4, 4, 5),
    KB_ROWMID_SPLIT(315, 46305, 2, 256, 15, 8, true, 4, 4, 4, 5),
};

}  // namespace

bool rowmid_split_available(int n2, int n1, int hinv, int slices) {
    for (const RowMidSplitEntry& e : kRowMidSplit)
        if (e.n2 == n2 && e.n1 == n1 && e.hinv == hinv && e.slices == slices) return true;
    return false;
}

hipError_t launch_rowmid_split(const PassArgs& a, const MidArgs& m, hipStream_t s) {
    for (const RowMidSplitEntry& e : kRowMidSplit)
        if (e.n2 == a.sd.n && e.n1 == a.nsub && e.hinv == m.hinv && e.slices == a.split) return e.launch(a, m, s);
    return hipErrorInvalidValue;
}

}  // namespace kbsrc::gpu
