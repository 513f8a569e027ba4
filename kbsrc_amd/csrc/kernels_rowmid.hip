// The forward row pass and the split/multiply/fold step as one kernel, for plans whose four-step layout lets a
// workgroup own every forward bin that its inverse-input bins need.
//
// The forward complex FFT of h = N1 N2 points leaves bin i = k1 + N1 k2 in row k1 (the row pass's sub-FFT k1,
// element k2). The split/fold step of the inverse-input pair (k, hinv - k) reads forward bins i and h - i (the real
// split), Mb - i (the fold when downsampling; upsampling reads bin i mod Nb, or Nb minus it) and hinv - i (the
// pre-split's partner): with t = gcd(N1, hinv), and h, Nb and Mb = 2 hinv multiples of t, every one of them lies in a
// row congruent to +-k mod t. The rows r + t m and (t - r) + t m
// (m < N1 / t) form a unit. A workgroup runs the row sub-FFTs of U units in LDS (the row pass's operations: first
// stage from the column pass's output, the rest in place, renormalised as the row pass stores them), then, reading
// those bins from LDS, the split/fold step of every pair (k, hinv - k) with k = +-r mod t, k <= hinv / 2, writing
// the inverse FFT's input. Same operations on the same values as the row pass followed by the split/fold kernel
// (mid_impl.hpp), so the same bits; the forward spectrum never goes through HBM.
//
// Stereo 96 kHz -> 44.1 kHz (N1 = 315, N2 = 320, hinv = 46305): t = 315, each unit is the row pair (r, 315 - r).
#include <cstdlib>
#include <numeric>

#include "rowmid_impl.hpp"

namespace kbsrc::gpu {

namespace {

template <class A, class RL, int N1, int TT, int U, int T, bool NTS, bool RES>
__global__ __launch_bounds__(T, 4) void k_rowmid(PassArgs a, MidArgs m, int ngrp, int tgroup) {
    using RU = RowUnits<N1, TT, U>;
    constexpr int N = RL::N, G = RU::G, CAP = N * G;
    extern __shared__ double lds[];
    const Lds<A, CAP> L{lds};
    const int tid = threadIdx.x;
    int task, grp;
    rowmid_order(xcd_wg(), a.ntask, ngrp, tgroup, task, grp);
    const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
    const PassView view = pass_view<LD_CPLX>(a, io, task);
    auto jof = [&](int g) -> int { return RU::row(g, grp); };
    auto load_at = [&](int e, int j) -> Cx<A> { return pass_load<A, LD_CPLX>(view, e, j); };
    const OwnTw tw{a.tw};
    const Prune pr{a.wstart, a.wlen};
    first_stage<A, RL, G, T>(L, tid, tw, a.oddc, jof, load_at);
    inner_stages<A, RL, G, T>(L, tid, tw, a.oddc, pr, std::make_integer_sequence<int, RL::nst - 1>{});
    rowmid_tail<A, RL, N1, TT, U, T, NTS, RES>(L, a, m, task, grp, tid);
}

// residue-major copy of a table (MidArgs::Hr, wfr, wir)
__global__ __launch_bounds__(256) void k_residue_major(const double* src, long long n, int e, int t, int L,
                                                       double* dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long q = i / t, pos = (i - q * t) * L + q;
    for (int c = 0; c < e; ++c) dst[pos * e + c] = src[i * e + c];
}

template <class K> hipError_t lds_attr(K kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

int gcd_int(int a, int b) {
    while (b) {
        const int t = a % b;
        a = b;
        b = t;
    }
    return a;
}

template <class A, class RL, int N1, int TT, int U, int T, bool NTS, int TG, bool RES>
hipError_t launch_rowmid_one(const PassArgs& a, const MidArgs& m, hipStream_t s) {
    using RU = RowUnits<N1, TT, U>;
    static_assert(N1 % TT == 0, "a unit is whole residue classes of the rows");
    static_assert(RL::hold(RU::G, T) <= (A::kDD ? 8 : 16), "too many values held per thread");
    constexpr size_t bytes = (size_t)kPlanes<A> * RL::N * RU::G * sizeof(double);
    static_assert(bytes <= 160 * 1024, "LDS tile too large");
    // the pass and the split/fold step must be the ones this kernel was built for
    if (a.sd.n != RL::N || a.sd.nst != RL::nst || a.nsub != N1 || a.twmode != 0 || m.Jg || m.Mb != 2 * m.hinv ||
        m.h != N1 * RL::N || gcd_int(N1, m.hinv) != TT || a.ntask < 1 || m.ntask != a.ntask ||
        (RES && (!m.Hr || !m.wfr || !m.wir || m.Lr != RU::NQ * RL::N + 1 ||
                (long long)m.Lr * TT * 32 > 0xFFFFFFFFLL)) ||
        m.Mb > (int)kDivSmallMax || m.h > (int)kDivSmallMax)   // the bin indices of the tail (div_small)
        return hipErrorInvalidValue;
    for (int i = 0; i < RL::nst; ++i)
        if (a.sd.st[i].r != RL::r[i] || a.sd.st[i].s != RL::stride(i)) return hipErrorInvalidValue;
    static const hipError_t attr = lds_attr(&k_rowmid<A, RL, N1, TT, U, T, NTS, RES>, bytes);
    if (attr != hipSuccess) return attr;
    const int ngrp = (RU::nunit + U - 1) / U;
    const long long nwg = (long long)a.ntask * ngrp;
    if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
    int tgroup = TG;
    if (const char* v = std::getenv("KBSRC_ROWMID_TGROUP")) tgroup = std::max(1, std::atoi(v));   // tuning knob
    hipLaunchKernelGGL((k_rowmid<A, RL, N1, TT, U, T, NTS, RES>), dim3((unsigned)nwg), dim3(T), bytes, s, a, m, ngrp,
                       std::min(tgroup, a.ntask));
    return hipGetLastError();
}

struct RowMidEntry {
    int policy, n2, n1, hinv;
    bool res;
    hipError_t (*launch)(const PassArgs&, const MidArgs&, hipStream_t);
};

// An entry serves one plan shape: the row pass (policy, N2-point rows, N1 of them) and the inverse input's hinv, for
// which it was measured faster than the two kernels (the split/fold step's share of the work grows with hinv).
#define KB_ROWMID(POL, N1, HINV, U, T, NTS, TG, RES, ...)                                                          \
    RowMidEntry{POL, Radices<__VA_ARGS__>::N, N1, HINV, RES,                                                       \
                &launch_rowmid_one<PolicyOf<POL>::type, Radices<__VA_ARGS__>, N1, std::gcd(N1, HINV), U, T, NTS, TG,  \
                                   RES>}

// Several entries for one shape: the first is the default, KBSRC_ROWMID_VARIANT=k takes the k-th. TG: tasks per
// group of the workgroup order (rowmid_order), measured: exact 1.617, 1.601, 1.595 and 1.589 ms per conversion at 1,
// 2, 4 and 8; fast 0.423, 0.432, 0.449 and 0.486 ms per launch. RES: the split/fold step reads the residue-major
// tables (MidArgs::Hr, wfr, wir), whose entries for one residue are adjacent: the fast plan's fused launch 0.425 ->
// 0.324 ms (its table loads were one cache line per lane), exact 0.625 -> 0.61-0.62 ms, the other shapes 0-2%. With
// the residue-major tables, measured against the defaults below: fast 0.373 ms on 256 threads, 0.326 on 768, 0.350
// in task groups of 4, 0.445 for two units on 1024 threads; exact 0.62-0.63 ms for 3 pairs on 512 threads, 0.612 in
// task groups of 4, against 0.60-0.61; two pairs on 320 threads at 5 waves per SIMD (84 VGPRs, no spills) 0.70 ms,
// at 4 waves 0.76, two pairs on 256 threads at 5 waves (a 52-byte spill) 0.65.
const RowMidEntry kRowMid[] = {
    // stereo 96 kHz -> 44.1 kHz exact: 320-point rows, row pairs (r, 315 - r); per launch (214 tasks) 0.62-0.635 ms
    // for 2 pairs on 256 threads (0.654-0.670 before the in-place renormalisation and the task groups; then 3 pairs
    // on 512 threads the same, 0.67 with non-temporal stores, 0.688 for 4 on 512, 0.742 for 1 on 128, 0.80-0.81 for 2
    // on 320, three workgroups per CU, and 3 on 384); the two kernels it replaces: 0.665-0.675
    KB_ROWMID(POL_DDS, 315, 46305, 2, 256, false, 8, true, 4, 4, 4, 5),
    KB_ROWMID(POL_DDS, 315, 46305, 2, 256, false, 8, false, 4, 4, 4, 5),
    KB_ROWMID(POL_DDS, 315, 46305, 3, 512, false, 8, false, 4, 4, 4, 5),
    KB_ROWMID(POL_DDS, 315, 46305, 3, 384, false, 8, false, 4, 4, 4, 5),
    KB_ROWMID(POL_DDS, 315, 46305, 2, 256, true, 8, false, 4, 4, 4, 5),
    KB_ROWMID(POL_DDS, 315, 46305, 3, 512, true, 8, false, 4, 4, 4, 5),
    KB_ROWMID(POL_DDS, 315, 46305, 4, 512, false, 8, false, 4, 4, 4, 5),
    // stereo 96 kHz -> 44.1 kHz fast (FP64): 200-point rows, units of 16 rows (r + 25 m, 25 - r + 25 m); per launch
    // (860 tasks) 0.423 ms at 512 threads, 0.430 at 768, 0.451 at 256, 0.459 at 640, 0.687 with non-temporal stores
    KB_ROWMID(POL_F64, 200, 18375, 1, 512, false, 1, true, 4, 2, 5, 5),
    KB_ROWMID(POL_F64, 200, 18375, 1, 512, false, 1, false, 4, 2, 5, 5),
    KB_ROWMID(POL_F64, 200, 18375, 1, 256, false, 1, false, 4, 2, 5, 5),
    KB_ROWMID(POL_F64, 200, 18375, 1, 640, false, 1, false, 4, 2, 5, 5),
    KB_ROWMID(POL_F64, 200, 18375, 1, 768, false, 1, false, 4, 2, 5, 5),
    KB_ROWMID(POL_F64, 200, 18375, 1, 512, true, 1, false, 4, 2, 5, 5),
    // other downsampling plans of long inputs between 44.1 and 192 kHz, exact (60 s stereo conversions, fused
    // against two kernels): 88.2 kHz -> 44.1 kHz 0.566 against 0.588 ms, 176.4 -> 44.1 0.866 / 0.911, 192 -> 96
    // 1.309 / 1.337, 192 -> 48 1.075 / 1.101, 176.4 -> 48 0.975 / 0.995
    KB_ROWMID(POL_DDS, 280, 44800, 2, 256, false, 8, true, 4, 4, 4, 5),  // 88.2 kHz -> 44.1 kHz: row pairs
    KB_ROWMID(POL_DDS, 400, 44800, 2, 256, false, 8, true, 4, 4, 4, 7),  // 176.4 kHz -> 44.1 kHz: row pairs
    KB_ROWMID(POL_DDS, 392, 48020, 1, 512, false, 8, true, 2, 5, 7, 7),  // 192 kHz -> 48 kHz: units of 4 rows
    KB_ROWMID(POL_DDS, 392, 96040, 2, 512, false, 8, true, 2, 5, 7, 7),  // 192 kHz -> 96 kHz: row pairs
    KB_ROWMID(POL_DDS, 405, 48600, 2, 384, false, 8, true, 3, 3, 7, 7),  // 176.4 kHz -> 48 kHz: row pairs
    // not registered (same bits, slower than the two kernels, where the split/fold step's items take two or three
    // rounds of the workgroup): 176.4 kHz -> 96 kHz (405 x 441, hinv 97200) 1.26 against 1.228 ms, 96 -> 48 kHz (280
    // x 343, units of 4 rows) 0.72 / 0.685, 96 -> 88.2 kHz (250 x 400, units of 4 rows) 0.90 / 0.827; stereo 44.1 kHz
    // -> 48 kHz exact (upsampling; 245-point rows, units of 6 rows r + 63 m and 63 - r + 63 m): 0.469-0.565 ms per
    // launch on 320-512 threads against 0.434 for the two kernels, its radix-5 and -7 stages filling 41-57% of the
    // lanes of six rows (with the residue-major tables 0.452-0.521 ms against 0.427; conversion 1.117-1.124 against
    // 1.110 ms)
};

}  // namespace

bool rowmid_residue_tables(int policy, int n2, int n1, int hinv) {
    for (const RowMidEntry& e : kRowMid)
        if (e.policy == policy && e.n2 == n2 && e.n1 == n1 && e.hinv == hinv && e.res) return true;
    return false;
}

hipError_t launch_residue_major(const double* src, long long n, int e, int t, int L, double* dst, hipStream_t s) {
    if (n < 1 || e < 1 || t < 1 || (long long)L * t < n) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_residue_major, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, n, e, t, L, dst);
    return hipGetLastError();
}

bool rowmid_available(int policy, int n2, int n1, int hinv) {
    for (const RowMidEntry& e : kRowMid)
        if (e.policy == policy && e.n2 == n2 && e.n1 == n1 && e.hinv == hinv) return true;
    return false;
}

hipError_t launch_rowmid(int policy, const PassArgs& a, const MidArgs& m, hipStream_t s) {
    int skip = 0;   // KBSRC_ROWMID_VARIANT=k: the k-th entry of the shape, the first where it has fewer
    if (const char* v = std::getenv("KBSRC_ROWMID_VARIANT")) skip = std::max(0, std::atoi(v));
    const RowMidEntry* first = nullptr;
    for (const RowMidEntry& e : kRowMid)
        if (e.policy == policy && e.n2 == a.sd.n && e.n1 == a.nsub && e.hinv == m.hinv) {
            if (!first) first = &e;
            if (skip-- <= 0) return e.launch(a, m, s);
        }
    return first ? first->launch(a, m, s) : hipErrorInvalidValue;
}

}  // namespace kbsrc::gpu
