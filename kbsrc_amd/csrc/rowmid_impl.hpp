// The part of the fused forward row pass + split/multiply/fold kernels that follows the row sub-FFTs (shared by the
// plain form, kernels_rowmid.hip, and the split form, kernels_rowmid_split.hip): which rows a workgroup owns (units),
// the workgroup order, and the split/fold step over the rows in LDS.
#pragma once
#include "fixed_impl.hpp"
#include "mid_impl.hpp"

namespace kbsrc::gpu {
namespace {

template <int N1, int TT, int U> struct RowUnits {
    static constexpr int NQ = N1 / TT;   // rows per residue
    static constexpr int Q = 2 * NQ;     // rows per unit
    static constexpr int G = U * Q;      // rows per workgroup
    static constexpr int nunit = TT / 2 + 1;
    // row of local slot g in unit group grp, or -1
    static __device__ __forceinline__ int row(int g, int grp) {
        const int ul = g / Q, q = g - ul * Q, side = q / NQ, mm = q - side * NQ;
        const int u = grp * U + ul;
        if (2 * u > TT) return -1;
        if (side && (u == 0 || 2 * u == TT)) return -1;
        return (side ? TT - u : u) + TT * mm;
    }
    // local slot of row j (a row of unit group grp)
    static __device__ __forceinline__ int slot(int j, int grp) {
        int mm = 0, v = j;
        if constexpr (NQ > 1) {
            mm = (int)div_small<TT>((unsigned)j);
            v = j - mm * TT;
        }
        const int u = v <= TT - v ? v : TT - v, side = v == u ? 0 : 1;
        return (u - grp * U) * Q + side * NQ + mm;
    }
};

// Workgroup wg's task and unit group: tasks in groups of tgroup, each group's unit groups in turn with its tasks
// fastest (task_group_order), so that the tasks of one unit group, which read the same entries of the kernel spectrum
// and the split twiddles (the same residues mod t), run one after another on one XCD (xcd_wg) and find them in its L2.
__device__ __forceinline__ void rowmid_order(int wg, int ntask, int ngrp, int tgroup, int& task, int& grp) {
    task_group_order(wg, ntask, ngrp, tgroup, task, grp);
}

// After the row sub-FFTs of unit group grp (local sub-FFT g = RowUnits::row(g, grp)) are in the LDS tile L, in
// natural order: renormalise them once as the row pass's store does (a.renorm), then run the split/fold step of every
// inverse-input pair (k, hinv - k) with k in the group's units, writing the inverse FFT's input (m.zi, task's part);
// RES: its tables from the residue-major copies (MidArgs::Hr, wfr, wir), whose entries for one residue are adjacent.
template <class A, class RL, int N1, int TT, int U, int T, bool NTS, bool RES>
__device__ __forceinline__ void rowmid_tail(const Lds<A, RL::N * RowUnits<N1, TT, U>::G>& L, const PassArgs& a,
                                            const MidArgs& m, int task, int grp, int tid) {
    using RU = RowUnits<N1, TT, U>;
    using S = SplitPol<A>;
    static_assert(std::is_same_v<Val<S>, Val<A>>);
    constexpr int G = RU::G, CAP = RL::N * G;
    constexpr int LR = RU::NQ * RL::N + 1;   // residue-major entries per residue: h / t + 1 (MidArgs::Lr)
    auto jof = [&](int g) -> int { return RU::row(g, grp); };
    if (a.renorm) {   // the row pass's stores renormalise each element: once here, in place
        for (int w = tid; w < CAP; w += T)
            if (jof(w % G) >= 0) L.put(w, renorm_dd<A>(L.get(w)));
        __syncthreads();
    }

    // forward bin i = row j, element e, as the row pass would have stored it
    auto zat = [&](int i) -> Cx<S> {   // bin i = row j, element e (i <= kDivSmallMax: the launcher checks)
        const unsigned e = div_small<N1>((unsigned)i), j = (unsigned)i - __umul24(e, N1);
        return L.get((int)e * G + RU::slot((int)j, grp));
    };
    double* zi = m.zi + task * m.zi_tstride;
    auto put = [&](int i, Cx<S> v) {
        if constexpr (NTS) {
            typedef double d2 __attribute__((ext_vector_type(2)));
            d2* p = reinterpret_cast<d2*>(const_cast<double2*>(at32<S>(zi, (unsigned)i)));
            if constexpr (S::kDD) {
                __builtin_nontemporal_store(d2{v.re.hi, v.im.hi}, p);
                __builtin_nontemporal_store(d2{v.re.lo, v.im.lo}, p + 1);
            } else {
                __builtin_nontemporal_store(d2{v.re, v.im}, p);
            }
        } else {
            gst32<S>(zi, (unsigned)i, v);
        }
    };
    // pair (k, hinv - k) for k = +-u mod TT: unit fastest (neighbouring lanes write neighbouring bins), then side
    const int half = m.hinv / 2, mc = half / TT + 1, items = U * 2 * mc;
    for (int w = tid; w < items; w += T) {
        const int ul = w % U, rest = w / U, side = rest & 1, mm = rest >> 1;
        const int u = grp * U + ul;
        if (2 * u > TT || (side && (u == 0 || 2 * u == TT))) continue;
        const int k = (side ? TT - u : u) + TT * mm;
        if (k > half) continue;
        if constexpr (RU::NQ == 1) {   // row pairs: row u in slot 2 ul, row N1 - u in slot 2 ul + 1 (RowUnits::slot)
            auto zpair = [&](int i) -> Cx<S> {
                const unsigned e = div_small<N1>((unsigned)i), j = (unsigned)i - __umul24(e, N1);
                return L.get((int)e * G + 2 * ul + (j != (unsigned)u ? 1 : 0));
            };
            if constexpr (RES) mid_pair<S>(m, zpair, MidTabsRes<TT, LR>{m.Hr, m.wfr, m.wir}, k, put);
            else mid_pair<S>(m, zpair, MidTabs{m.H, m.wf, m.wi}, k, put);
        } else {
            if constexpr (RES) mid_pair<S>(m, zat, MidTabsRes<TT, LR>{m.Hr, m.wfr, m.wir}, k, put);
            else mid_pair<S>(m, zat, MidTabs{m.H, m.wf, m.wi}, k, put);
        }
    }
}

}  // namespace
}  // namespace kbsrc::gpu
