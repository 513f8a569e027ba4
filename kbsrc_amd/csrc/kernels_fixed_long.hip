// Long sub-FFTs as two specialised passes (the 3150-point forward columns of the 10 MHz -> 48 kHz plan, ...).
//
// A sub-FFT of n = R M points whose Stockham stages split after stage k (R = the product of the first k radices):
// the first k stages never mix the M columns c + M l (l < R), the rest never mix the R classes q + R j (j < M)
// (src/dsp/fft.hpp:115-150 stage indexing). So the transform runs as
//   head  every column c: an R-point sub-FFT of stages 0 .. k-1, its butterfly p being the long transform's
//         butterfly c + M p (twiddle from the long transform's stage table);
//   tail  every class q: an M-point sub-FFT of stages k .. with strides s / R, the same butterfly indices p and
//         the long transform's tables;
// the same operations on the same values as one pass over the whole sub-FFT, so the same bits. Against a 100 KB
// tile per workgroup (one workgroup per CU, stages whose butterfly counts leave up to half the lanes idle), the two
// parts have small tiles, many sub-FFTs per workgroup and several workgroups per CU, for one extra round trip
// through HBM.
//
// Both parts take their G sub-FFTs of a workgroup from G neighbouring long sub-FFTs j (the pass's sub-FFT index)
// with the same column c (head) or class q (tail), so lanes reach neighbouring j: the layout the one-pass kernel
// reads and writes. The head reads the pass's source and writes the scratch buffer, element x = c R + e of long
// sub-FFT j at x nsub + j (per task); the tail reads it and writes through the pass's store (four-step twiddle,
// renormalisation, unread bins).
#include "fixed_impl.hpp"

namespace kbsrc::gpu {

namespace {

template <class A, class RH, int G, int T, int LD>
__global__ __launch_bounds__(T, 4) void k_long_head(PassArgs a, double* scr, long long scr_ts, int ngrp) {
    constexpr int R = RH::N;
    extern __shared__ double lds[];
    const Lds<A, R * G> L{lds};
    const int tid = threadIdx.x, wg = xcd_wg();
    const int M = a.sd.n / R;
    const int tc = wg / ngrp, grp = wg - tc * ngrp;   // (task, column c) and the group of long sub-FFTs
    const int task = tc / M, c = tc - task * M, j0 = grp * G;
    const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
    const PassView view = pass_view<LD>(a, io, task);
    double* out = scr + task * scr_ts;
    const int nsub = a.nsub;
    auto jof = [&](int g) -> int { return j0 + g < nsub ? j0 + g : -1; };
    auto load_at = [&](int l, int j) -> Cx<A> { return pass_load<A, LD>(view, c + M * l, j); };
    auto store_at = [&](int e, int j, Cx<A> v) { gst32<A>(out, (unsigned)((c * R + e) * nsub + j), v); };
    const LongTw tw{a.tw, a.sd.st, c, 0, M};
    const Prune none;
    first_stage<A, RH, G, T>(L, tid, tw, a.oddc, jof, load_at);
    inner_stages<A, RH, G, T>(L, tid, tw, a.oddc, none, std::make_integer_sequence<int, RH::nst - 2>{});
    last_stage<A, RH, G, T>(L, tid, tw, a.oddc, none, jof, store_at);
}

// Tail workgroups run tasks fastest (tfast): the tasks of one (class, group) read the same four-step twiddles
// (indexed like the store, the same table for every task) one after another, from L2.
template <class A, class RT, int G, int T, int ST>
__global__ __launch_bounds__(T, 4) void k_long_tail(PassArgs a, const double* scr, long long scr_ts, int k, int ngrp,
                                                   int tfast) {
    constexpr int M = RT::N;
    extern __shared__ double lds[];
    const Lds<A, M * G> L{lds};
    const int tid = threadIdx.x, wg = xcd_wg();
    const int R = a.sd.n / M;
    int task, q, grp;
    if (tfast) {
        task = wg % a.ntask;
        const int qg = wg / a.ntask;
        q = qg / ngrp;
        grp = qg - q * ngrp;
    } else {
        const int tq = wg / ngrp;   // (task, class q) and the group of long sub-FFTs
        grp = wg - tq * ngrp;
        task = tq / R;
        q = tq - task * R;
    }
    const int j0 = grp * G;
    const TaskIO io = task_io(a.tdesc, a.task0 + task, a.nblk, a.in_cstride, a.in_bstride, a.nout);
    const PassView view = pass_view<LD_CPLX>(a, io, task);
    const double* in = scr + task * scr_ts;
    const int nsub = a.nsub;
    auto jof = [&](int g) -> int { return j0 + g < nsub ? j0 + g : -1; };
    auto load_at = [&](int jj, int j) -> Cx<A> { return gld32<A>(in, (unsigned)((jj * R + q) * nsub + j)); };
    auto store_at = [&](int e, int j, Cx<A> v) { pass_store<A, ST>(a, view, e * R + q, j, v); };
    const LongTw tw{a.tw, a.sd.st + k, 0, 0, 1};
    const Prune none;
    first_stage<A, RT, G, T>(L, tid, tw, a.oddc, jof, load_at);
    inner_stages<A, RT, G, T>(L, tid, tw, a.oddc, none, std::make_integer_sequence<int, RT::nst - 2>{});
    last_stage<A, RT, G, T>(L, tid, tw, a.oddc, none, jof, store_at);
}

template <class K> hipError_t lds_attr(K kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

// The head's radix list followed by the tail's must be the pass's stage list, strides included.
template <class RH, class RT> bool stages_match(const SubDesc& sd) {
    constexpr int k = RH::nst, R = RH::N;
    if (sd.nst != RH::nst + RT::nst || sd.n != RH::N * RT::N) return false;
    for (int i = 0; i < k; ++i)
        if (sd.st[i].r != RH::r[i] || sd.st[i].s != RH::stride(i)) return false;
    for (int i = 0; i < RT::nst; ++i)
        if (sd.st[k + i].r != RT::r[i] || sd.st[k + i].s != R * RT::stride(i)) return false;
    return true;
}

template <class A, class RH, int GH, int TH, class RT, int GT, int TT, int LD, int ST>
hipError_t launch_long_two(const PassArgs& a, double* scr, long long scr_ts, hipStream_t s) {
    static_assert(RH::nst >= 2 && RT::nst >= 2, "each part runs its first and last stages on global memory");
    static_assert(RH::hold(GH, TH) <= (A::kDD ? 8 : 16) && RT::hold(GT, TT) <= (A::kDD ? 8 : 16),
                  "too many values held per thread");
    if (!stages_match<RH, RT>(a.sd) || a.nsub < 1 || a.ntask < 1) return hipErrorInvalidValue;
    constexpr int R = RH::N, M = RT::N;
    constexpr size_t hb = (size_t)kPlanes<A> * R * GH * sizeof(double), tb = (size_t)kPlanes<A> * M * GT * sizeof(double);
    static_assert(hb <= 160 * 1024 && tb <= 160 * 1024, "LDS tile too large");
    static const hipError_t ah = lds_attr(&k_long_head<A, RH, GH, TH, LD>, hb);
    static const hipError_t at = lds_attr(&k_long_tail<A, RT, GT, TT, ST>, tb);
    if (ah != hipSuccess) return ah;
    if (at != hipSuccess) return at;
    const int gh = (a.nsub + GH - 1) / GH, gt = (a.nsub + GT - 1) / GT;
    const long long wh = (long long)a.ntask * M * gh, wt = (long long)a.ntask * R * gt;
    if (wh > 0x7fffffffLL || wt > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_long_head<A, RH, GH, TH, LD>), dim3((unsigned)wh), dim3(TH), hb, s, a, scr, scr_ts, gh);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    const char* order = std::getenv("KBSRC_LONG_ORDER");   // tuning knob: 0 = tasks slowest in the tail
    const int tfast = order && order[0] == '0' ? 0 : 1;
    hipLaunchKernelGGL((k_long_tail<A, RT, GT, TT, ST>), dim3((unsigned)wt), dim3(TT), tb, s, a, (const double*)scr,
                       scr_ts, RH::nst, gt, tfast);
    return hipGetLastError();
}

#define KB_LONG(POL, LD, ST, GH, TH, GT, TT, HEAD, TAIL)                                                             \
    LongEntry{POL, HEAD::N * TAIL::N, LD, ST, HEAD::N,                                                            \
              &launch_long_two<PolicyOf<POL>::type, HEAD, GH, TH, TAIL, GT, TT, LD, ST>}

struct LongEntry {
    int policy, n, ld, st, R;
    hipError_t (*launch)(const PassArgs&, double*, long long, hipStream_t);
};

// the parts' radix lists
using R233 = Radices<2, 3, 3>;
using R557 = Radices<5, 5, 7>;
using R55 = Radices<5, 5>;
using R555 = Radices<5, 5, 5>;

// Several entries for one (policy, n, ld, st): the first is the default, KBSRC_LONG_VARIANT=k takes the k-th.
const LongEntry kLong[] = {
    // 3150-point forward columns of the 10 MHz -> 48 kHz exact plan: 18 x 175 (90 x 35 spills in its tail). (G, T)
    // per part so that the stages' butterfly counts fill their rounds: 18 points x 85 on 256 threads (765 and 510
    // butterflies, 3 and 2 full rounds), 175 x 7 on 256 (245 and 175): 11.9 ms per 31-task pass (one pass: 16.2 ms);
    // 18 x 128 on 576, 175 x 14 on 512: 12.9 ms
    KB_LONG(POL_DDS, LD_REAL_F64, ST_CPLX, 85, 256, 7, 256, R233, R557),
    KB_LONG(POL_DDS, LD_REAL_F64, ST_CPLX, 128, 576, 14, 512, R233, R557),
    // 3125-point forward columns of the 10 MHz -> 48 kHz fast plan (FP64): 125 x 25 (4.70 ms per 31-task pass) and
    // 25 x 125 (5.11 ms; one pass: 6.55 ms); 125 points x 10 on 256 threads (250 butterflies per stage), 25 x 51 on
    // 256 (255)
    KB_LONG(POL_F64, LD_REAL_F64, ST_CPLX, 10, 256, 51, 256, R555, R55),
    KB_LONG(POL_F64, LD_REAL_F64, ST_CPLX, 51, 256, 10, 256, R55, R555),
};

}  // namespace

bool long_two_part(int policy, int n, int ld, int st) {
    for (const LongEntry& e : kLong)
        if (e.policy == policy && e.n == n && e.ld == ld && e.st == st) return true;
    return false;
}

hipError_t launch_long_two_part(int policy, int ld, int st, const PassArgs& a, double* scr, long long scr_ts,
                                hipStream_t s) {
    int skip = 0;   // KBSRC_LONG_VARIANT=k: the k-th entry of the shape, the first where it has fewer
    if (const char* v = std::getenv("KBSRC_LONG_VARIANT")) skip = std::max(0, std::atoi(v));
    const LongEntry* first = nullptr;
    for (const LongEntry& e : kLong)
        if (e.policy == policy && e.n == a.sd.n && e.ld == ld && e.st == st) {
            if (!first) first = &e;
            if (skip-- <= 0) return e.launch(a, scr, scr_ts, s);
        }
    return first ? first->launch(a, scr, scr_ts, s) : hipErrorInvalidValue;
}

}  // namespace kbsrc::gpu
