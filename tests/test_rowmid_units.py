"""The fused forward row pass + split/multiply/fold kernel (kbsrc_amd/csrc/kernels_rowmid.hip) on the CPU: its
workgroup owns the rows r + t m and (t - r) + t m of the four-step layout (t = gcd(N1, hinv)) and runs the
split/fold step of every inverse-input pair (k, hinv - k) with k = +-r mod t. Checked here:
- every forward bin that step reads for such a pair lies in those rows, by the reference's access pattern (the real
  split pairs bin i with h - i, src/dsp/fft.hpp:365-407; the bin map folds Mb - j onto j, src/resampler.cpp:64-78;
  the inverse pre-split pairs k with hinv - k, src/dsp/fft.hpp:412-458; upsampling extends the spectrum periodically,
  src/resampler.cpp:72-77), over the plans of many rate pairs, down and up;
- the kernel's slot and item maps (RowUnits::row / slot, the item loop) cover every row once and every pair once;
- the FP32 quotient its index arithmetic uses (div_small) is exact for every index up to its bound.
No GPU."""
from math import gcd

import numpy as np
import pytest

import helpers as H


def four_step_n1(n):
    """The reference's column length of an n-point complex FFT: the largest divisor <= sqrt(n) (src/dsp/fft.hpp:243),
    four-step only above 2048 points with N1 >= 8."""
    if n <= 2048:
        return None
    n1 = max(d for d in range(1, int(n ** 0.5) + 1) if n % d == 0)
    return n1 if n1 >= 8 else None


def bins_read(k, h, Mb, hinv, K, jmax, up=False):
    """Forward complex-FFT bins the split/fold step reads for the pair (k, hinv - k): downsampling folds bin Mb - idx
    onto idx; upsampling extends the spectrum periodically (bin idx mod Nb, mirrored above Nb / 2)."""
    acc = set()

    def x(i):   # X[i] from z[i'] and z[h - i'], i' = min(i, h - i); X[0] and X[h] from z[0]
        if i == 0 or i == h:
            acc.add(0)
            return
        lo = i if 2 * i <= h else h - i
        acc.update((lo, h - lo))

    def y(idx):
        if up:
            kk = idx % (2 * h)
            x(kk if 2 * kk <= 2 * h else 2 * h - kk)
            return
        x(idx)
        if idx != hinv and K <= Mb - idx <= jmax:
            x(Mb - idx)

    y(k)
    y(hinv - k)
    return acc


def plans():
    out = []
    pairs = ((96000, 44100), (48000, 44100), (96000, 48000), (192000, 44100), (88200, 48000), (44100, 32000),
             (48000, 22050), (176400, 96000), (32000, 22050), (192000, 88200))
    for rin, rout in pairs + tuple((b, a) for a, b in pairs):
        for mode in ("exact", "fast"):
            p = H.oracle_plan(2_000_000, 1, rin, rout, mode)
            if p["gen"]:
                continue
            n1 = four_step_n1(p["Nb"] // 2)
            if n1:
                out.append((rin, rout, mode, p["Nb"], p["Mb"], n1, p["up"]))
    return out


@pytest.mark.parametrize("rin,rout,mode,Nb,Mb,N1,up", plans())
def test_split_fold_reads_only_its_units_rows(rin, rout, mode, Nb, Mb, N1, up):
    h, hinv = Nb // 2, Mb // 2
    K, jmax = Mb // 2 + 1, min(Nb // 2, Mb - 1)
    t = gcd(N1, hinv)
    assert h % t == 0 and Mb % t == 0
    step = max(1, (hinv // 2) // 4000)   # every pair on small plans, a sample of ~4000 on large ones
    for k in list(range(0, hinv // 2 + 1, step)) + [hinv // 2]:
        rows = {(b % N1) % t for b in bins_read(k, h, Mb, hinv, K, jmax, bool(up))}
        assert rows <= {k % t, (-k) % t}, (k, rows)


@pytest.mark.parametrize("rin,rout,mode,n1,n2,hinv", [(96000, 44100, "exact", 315, 320, 46305),
                                                      (96000, 44100, "fast", 200, 200, 18375),
                                                      (88200, 44100, "exact", 280, 320, 44800),
                                                      (176400, 44100, "exact", 400, 448, 44800),
                                                      (192000, 48000, "exact", 392, 490, 48020),
                                                      (192000, 96000, "exact", 392, 490, 96040),
                                                      (176400, 48000, "exact", 405, 441, 48600)])
def test_registered_fused_shapes_are_these_plans(rin, rout, mode, n1, n2, hinv):
    """The fused kernel's registered shapes (kernels_rowmid.hip: row count, row length, hinv) are the plans of these
    rate pairs for long inputs (60 s stereo), and their units fit the kernel: t = gcd(N1, hinv) gives row pairs or
    units of 4 (exact) / 16 (fast) rows."""
    p = H.oracle_plan(60 * rin, 2, rin, rout, mode)
    assert four_step_n1(p["Nb"] // 2) == n1 and p["Nb"] // 2 == n1 * n2 and p["Mb"] // 2 == hinv
    assert not p["up"] and not p["gen"]
    assert 2 * n1 // gcd(n1, hinv) in ((2, 4) if mode == "exact" else (16,))


def test_registered_shapes_are_benchmark_plans():
    """The shapes the kernel table holds: stereo 96 kHz -> 44.1 kHz exact (315 x 320, t = 315: row pairs) and fast
    (200 x 200, t = 25: units of 16 rows); and stereo 44.1 kHz -> 48 kHz exact (189 x 245, t = 63: units of 6 rows),
    built, bit-exact and measured slower than the two kernels, so not registered."""
    for rin, rout, mode, n1, n2, t in ((96000, 44100, "exact", 315, 320, 315), (96000, 44100, "fast", 200, 200, 25),
                                       (44100, 48000, "exact", 189, 245, 63)):
        p = H.oracle_plan(180 * rin, 2, rin, rout, mode)
        assert four_step_n1(p["Nb"] // 2) == n1 and p["Nb"] // 2 == n1 * n2
        assert gcd(n1, p["Mb"] // 2) == t


def unit_maps(N1, TT, U, hinv):
    """RowUnits::row over every workgroup's slots, and the rows and pairs each workgroup's item loop handles."""
    NQ, nunit = N1 // TT, TT // 2 + 1
    Q = 2 * NQ
    G, ngrp, half = U * Q, (nunit + U - 1) // U, hinv // 2
    mc = half // TT + 1
    groups = []
    for grp in range(ngrp):
        rows = {}
        for g in range(G):
            ul, q = divmod(g, Q)
            side, mm = divmod(q, NQ)
            u = grp * U + ul
            if 2 * u > TT or (side and (u == 0 or 2 * u == TT)):
                continue
            r = (TT - u if side else u) + TT * mm
            m2, v = divmod(r, TT)   # RowUnits::slot
            uu = min(v, TT - v)
            assert (uu - grp * U) * Q + (0 if v == uu else 1) * NQ + m2 == g
            if NQ == 1:   # the row-pair shortcut of the item loop: row u in slot 2 ul, row N1 - u in 2 ul + 1
                assert 2 * ul + (1 if r != u else 0) == g
            rows[g] = r
        ks = []
        for w in range(U * 2 * mc):
            ul, rest = w % U, w // U
            side, mm = rest & 1, rest >> 1
            u = grp * U + ul
            if 2 * u > TT or (side and (u == 0 or 2 * u == TT)):
                continue
            k = (TT - u if side else u) + TT * mm
            if k <= half:
                ks.append(k)
        groups.append((rows, ks))
    return groups


@pytest.mark.parametrize("N1,TT,U,hinv", [(315, 315, 1, 46305), (315, 315, 2, 46305), (315, 315, 3, 46305),
                                          (315, 315, 4, 46305), (200, 25, 1, 18375), (200, 25, 2, 18375),
                                          (189, 63, 1, 50400), (280, 280, 2, 44800), (400, 400, 2, 44800),
                                          (280, 140, 1, 48020), (392, 392, 2, 96040), (392, 196, 1, 48020),
                                          (405, 405, 2, 48600), (405, 405, 2, 97200), (250, 125, 1, 91875),
                                          (24, 8, 1, 1000), (24, 12, 2, 1200), (40, 10, 3, 990)])
def test_workgroups_cover_rows_and_pairs_once(N1, TT, U, hinv):
    groups = unit_maps(N1, TT, U, hinv)
    rows = sorted(r for g, _ in groups for r in g.values())
    assert rows == list(range(N1))
    pairs = sorted(k for _, ks in groups for k in ks)
    assert pairs == list(range(hinv // 2 + 1))
    for g, ks in groups:   # each pair's rows (k and hinv - k, mod t) are the workgroup's own
        own = {r % TT for r in g.values()}
        for k in ks:
            assert k % TT in own and (hinv - k) % TT in own


def rowmid_order(wg, ntask, ngrp, tgroup):
    """The kernel's workgroup order: tasks in groups of tgroup, each group's unit groups in turn, tasks fastest."""
    per = tgroup * ngrp
    blk = wg // per
    t0 = blk * tgroup
    gs = min(tgroup, ntask - t0)
    r = wg - blk * per
    grp = r // gs
    return t0 + (r - grp * gs), grp


@pytest.mark.parametrize("ntask,ngrp,tgroup", [(214, 79, 4), (860, 13, 4), (5, 79, 4), (1, 13, 4), (7, 53, 3),
                                               (214, 79, 1), (6, 5, 6)])
def test_workgroup_order_is_a_bijection(ntask, ngrp, tgroup):
    seen = [rowmid_order(w, ntask, ngrp, min(tgroup, ntask)) for w in range(ntask * ngrp)]
    assert sorted(seen) == [(t, g) for t in range(ntask) for g in range(ngrp)]


def div_small(u, d):
    """The kernel's quotient floor(u / d) (div_small in kbsrc_amd/csrc/mid_impl.hpp): fl32((u + 1/2) fl32(1 / d)),
    one FP32 fma, truncated. The fma's single rounding is emulated exactly: u fl32(1/d) + fl32(1/d) / 2 needs at
    most 47 significant bits for u < 2^22, so the float64 sum is exact and one float32 rounding follows."""
    inv = np.float32(1.0) / np.float32(d)
    y = (u.astype(np.float64) * np.float64(inv) + np.float64(np.float32(0.5) * inv)).astype(np.float32)
    return y.astype(np.int64)


K_DIV_SMALL_MAX = (1 << 22) - 2


@pytest.mark.parametrize("d", [315, 200, 25, 280, 400, 392, 196, 405, 1, 3, 7, 4095])
def test_float_quotient_is_exact_below_its_bound(d):
    """Every u in [0, kDivSmallMax]: the row and residue divisors of the registered shapes, and a few others."""
    u = np.arange(0, K_DIV_SMALL_MAX + 1, dtype=np.int64)
    assert np.array_equal(div_small(u, d), u // d)


def test_float_quotient_fails_above_its_bound():
    """The bound is needed: above 2^22 the same product truncates to a wrong quotient for some u."""
    u = np.arange(K_DIV_SMALL_MAX + 1, 1 << 25, dtype=np.int64)
    assert np.count_nonzero(div_small(u, 315) != u // 315) > 0
