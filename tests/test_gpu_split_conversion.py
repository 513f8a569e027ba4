"""The split-precision form of the conversion (kbsrc_plan_execute_split): the stereo 96 kHz -> 44.1 kHz and
44.1 kHz -> 48 kHz conversions at the block size of their 180 s runs (630), whose sub-FFTs of 320 (forward rows)
and 224 (inverse columns) points run their first two radix-4 stages as one int8-MFMA radix-16 stage. Per-block
errors against the high-precision block restatement (tests/hp_block.py, pinned against the oracle in
tests/test_hp_block.py) for the plain exact output and the split form at 6, 7, 8 and 15 slices, and the split form
against the plain output over every block. Needs an MI355X."""
import math
import sys

import numpy as np
import pytest

import hp_block as hb

pytestmark = pytest.mark.gpu

SLICES = (6, 7, 8, 15)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu():
    import kbsrc_amd
    return kbsrc_amd


def run_forms(x, fi, fo, block=630):
    """(plan info, staged padded input, kernel spectrum, {0: plain output, S: split output})."""
    import torch
    k = gpu()
    frames, ch = x.shape
    plan = k.Plan(frames, ch, fi, fo, "exact", block=block)
    info = plan.info
    d_pad = torch.zeros(ch * info["xstride"], dtype=torch.float64, device="cuda")
    d_out = torch.empty(ch * info["nout"], dtype=torch.float64, device="cuda")
    plan.stage(x, d_pad.data_ptr())
    outs = {}
    plan.execute(d_pad.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    outs[0] = d_out.cpu().numpy().reshape(ch, info["nout"]).copy()
    for s in SLICES:
        plan.execute_split(d_pad.data_ptr(), d_out.data_ptr(), s)
        torch.cuda.synchronize()
        outs[s] = d_out.cpu().numpy().reshape(ch, info["nout"]).copy()
    staged = d_pad.cpu().numpy().reshape(ch, info["xstride"])
    fk = fo if info["up"] else fi
    gc = math.gcd(info["c2"], fk)
    taps = k.kernel_taps(info["Lk"], info["c2"] // gc, fk // gc, 40.0)
    spec = k.kernel_spectrum(taps, info["Mb"] if info["up"] else info["Nb"], info["Nb"])
    plan.close()
    return info, staged, spec, outs


@pytest.mark.parametrize("fi,fo", [(96000, 44100), (44100, 48000)])
def test_split_form_errors_per_block(fi, fo):
    rng = [np.random.default_rng(s) for s in (1, 2)]
    n = 3 * fi
    x = np.stack([r.uniform(-0.5, 0.5, n) for r in rng], 1)
    say(f"{fi} -> {fo}: plain and split {SLICES}")
    info, staged, spec, outs = run_forms(x, fi, fo)
    P, Q, hop = info["P"], info["Q"], info["hop"]
    peak = max(np.abs(outs[0]).max(), 1e-300)
    # the split form against the plain output, every block
    diff = {s: float(np.abs(outs[s] - outs[0]).max() / peak) for s in SLICES}
    say(f"  max |split - plain| / peak: " + ", ".join(f"S={s}: 2^{math.log2(d + 1e-300):.1f}" for s, d in diff.items()))
    # per-block errors against the high-precision block, first block (edge extension) and a middle one, channel 0
    worst = {f: 0.0 for f in (0,) + SLICES}
    for b in (0, info["nblk"] // 2):
        blk = staged[0, b * hop * Q: b * hop * Q + info["Nb"]]
        ref = hb.block_output(blk, spec, P, Q, info["nb"], info["marg"], bool(info["up"]))
        r = hb.to_double(ref)
        d0 = b * hop * P - info["o"]
        i0, i1 = max(0, -d0), min(len(r), info["nout"] - d0)
        for f, y in outs.items():
            e = hb.errors(y[0, d0 + i0: d0 + i1], ref[i0:i1])
            if f == 0:   # plain: the correctly rounded value up to the dd error
                ulp = np.spacing(np.abs(r[i0:i1]))
                assert np.all(e <= 0.5 * ulp + 2.0 ** -100 * peak), (b, (e / ulp).max())
            worst[f] = max(worst[f], float(e.max() / peak))
    say("  per-block max error vs 2^-256 block / peak: " +
        ", ".join(f"{'plain' if f == 0 else 'S=%d' % f}: 2^{math.log2(v + 1e-300):.1f}" for f, v in worst.items()))
    # the split stage's error falls with every slice (7 bits each) until the output rounding hides it
    assert worst[6] > worst[7] and (worst[7] > worst[8] or worst[8] <= 2.0 ** -52)
    assert worst[6] < 2.0 ** -30 and worst[7] < 2.0 ** -37
    assert worst[8] < 2.0 ** -44
    assert worst[15] <= 2.0 ** -52 and diff[15] <= 2.0 ** -52
    assert diff[6] > 2.0 ** -60   # the split form really ran: 6 slices cannot match dd


def test_split_form_rejections():
    """The split form needs an exact-mode plan and a built slice count; a rejected call names the reason."""
    import torch
    k = gpu()
    x = np.random.default_rng(3).uniform(-0.5, 0.5, 48000)
    for mode, slices in (("fast", 8), ("exact", 3), ("exact", 9), ("exact", 17), ("exact", -1)):
        plan = k.Plan(len(x), 1, 48000, 44100, mode)
        d_pad = torch.zeros(plan.info["xstride"], dtype=torch.float64, device="cuda")
        d_out = torch.empty(plan.info["nout"], dtype=torch.float64, device="cuda")
        plan.stage(x, d_pad.data_ptr())
        with pytest.raises(ValueError, match="option out of range"):
            plan.execute_split(d_pad.data_ptr(), d_out.data_ptr(), slices)
        plan.close()


@pytest.mark.parametrize("fi,fo,block", [(96000, 44100, 630), (44100, 48000, 630), (384000, 44100, 1536)])
def test_split_routes_give_the_same_bits(fi, fo, block, monkeypatch):
    """The split form's routes compute the same operations and give the same bits: the specialised split passes
    (320- and 224-point sub-FFTs; on 96 kHz -> 44.1 kHz the forward row pass fused with the split/fold step, and the
    two as separate kernels with KBSRC_ROWMID=0), the generic LDS pass with the matrix stage (KBSRC_SPLIT_GLOBAL=2),
    and the global-memory route (gather, stage, tail pass, scatter; KBSRC_SPLIT_GLOBAL=1)."""
    import torch
    k = gpu()
    x = np.random.default_rng(fi + fo).uniform(-0.5, 0.5, (2 * fi, 2))
    plan = k.Plan(x.shape[0], 2, fi, fo, "exact", length=0.5 if fi == 384000 else 0.2, block=block)
    d_pad = torch.zeros(2 * plan.info["xstride"], dtype=torch.float64, device="cuda")
    d_out = torch.empty(2 * plan.info["nout"], dtype=torch.float64, device="cuda")
    plan.stage(x, d_pad.data_ptr())
    for s in (6, 7, 8, 15) if fi == 96000 else (6, 15):
        outs = []
        for route in (None, "rowmid0", "2", "1"):
            say(f"{fi} -> {fo}: split S={s}, route {route or 'specialised'}")
            monkeypatch.delenv("KBSRC_SPLIT_GLOBAL", raising=False)
            monkeypatch.delenv("KBSRC_ROWMID", raising=False)
            if route == "rowmid0":
                monkeypatch.setenv("KBSRC_ROWMID", "0")
            elif route:
                monkeypatch.setenv("KBSRC_SPLIT_GLOBAL", route)
            d_out.zero_()
            plan.execute_split(d_pad.data_ptr(), d_out.data_ptr(), s)
            torch.cuda.synchronize()
            outs.append(d_out.cpu().numpy().copy())
        assert all(np.array_equal(outs[0], o) for o in outs[1:]), s
    plan.close()


def test_split_routes_agree_10mhz(monkeypatch):
    """The 10 MHz -> 48 kHz plan (--len 0.5: a 20,160,000-point forward transform, 3150 x 3200): its 3200-point
    forward rows run the matrix stage in the specialised split pass (one 768-thread workgroup per sub-FFT) and
    through the global-memory route (KBSRC_SPLIT_GLOBAL=1); both give the same bits, at 8 and 15 slices."""
    import torch
    k = gpu()
    x = np.random.default_rng(10).uniform(-0.5, 0.5, 1_000_000)
    plan = k.Plan(len(x), 1, 10_000_000, 48000, "exact", length=0.5)
    d_pad = torch.zeros(plan.info["xstride"], dtype=torch.float64, device="cuda")
    d_out = torch.empty(plan.info["nout"], dtype=torch.float64, device="cuda")
    plan.stage(x, d_pad.data_ptr())
    plan.execute(d_pad.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    plain = d_out.cpu().numpy().copy()
    peak = np.abs(plain).max()
    for s in (8, 15):
        outs = []
        for route in (None, "1"):
            say(f"10 MHz -> 48 kHz: split S={s}, route {route or 'specialised'}")
            if route:
                monkeypatch.setenv("KBSRC_SPLIT_GLOBAL", route)
            else:
                monkeypatch.delenv("KBSRC_SPLIT_GLOBAL", raising=False)
            d_out.zero_()
            plan.execute_split(d_pad.data_ptr(), d_out.data_ptr(), s)
            torch.cuda.synchronize()
            outs.append(d_out.cpu().numpy().copy())
        assert np.array_equal(outs[0], outs[1]), s
        # and the split output stays within its slicing error of the plain one
        assert np.abs(outs[0] - plain).max() <= (2.0 ** -40 if s == 8 else 2.0 ** -50) * peak, s
    plan.close()
