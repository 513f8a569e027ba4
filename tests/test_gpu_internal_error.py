"""Internal double-double error of whole conversions, plain and split-precision, against a 2^-256 restatement.

kbsrc_plan_execute_dd writes every output sample as its dd value before the final rounding to double (the rounding
src/resampler.cpp:80-84 applies). Each sample's error is stated exactly against the block restatement of
tests/hp_block.py (the block's forward real DFT, x H, the bin map and the unnormalised inverse in fixed point at
2^-256, pinned against the oracle in tests/test_hp_block.py): the plain exact form and the split form at 15 slices
must stay within 2^-100 of the block's peak, and hi + lo must round to the samples the ordinary executes write.
Needs an MI355X."""
import math
import sys

import numpy as np
import pytest

import hp_block as hb

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("fi,fo", [(96000, 44100), (44100, 48000)])
def test_dd_output_error_against_the_2e256_block(fi, fo):
    import torch

    import kbsrc_amd as k
    rng = [np.random.default_rng(s) for s in (21, 22)]
    x = np.stack([r.uniform(-0.5, 0.5, 3 * fi) for r in rng], 1)
    frames, ch = x.shape
    plan = k.Plan(frames, ch, fi, fo, "exact", block=630)
    info = plan.info
    nout = info["nout"]
    d_pad = torch.zeros(ch * info["xstride"], dtype=torch.float64, device="cuda")
    d_out = torch.empty(ch * nout, dtype=torch.float64, device="cuda")
    d_pairs = torch.empty(2 * ch * nout, dtype=torch.float64, device="cuda")
    plan.stage(x, d_pad.data_ptr())
    forms = {}
    for slices in (0, 8, 15):
        if slices:
            plan.execute_split(d_pad.data_ptr(), d_out.data_ptr(), slices)
        else:
            plan.execute(d_pad.data_ptr(), d_out.data_ptr())
        plan.execute_dd(d_pad.data_ptr(), d_pairs.data_ptr(), slices)
        torch.cuda.synchronize()
        y = d_out.cpu().numpy().reshape(ch, nout).copy()
        pairs = d_pairs.cpu().numpy().reshape(ch, nout, 2).copy()
        # hi + lo rounds to the sample the ordinary execute writes
        assert np.array_equal(pairs[..., 0] + pairs[..., 1], y), f"S={slices}: pairs do not round to the output"
        forms[slices] = pairs
    staged = d_pad.cpu().numpy().reshape(ch, info["xstride"])
    fk = fo if info["up"] else fi
    gc = math.gcd(info["c2"], fk)
    taps = k.kernel_taps(info["Lk"], info["c2"] // gc, fk // gc, 40.0)
    spec = k.kernel_spectrum(taps, info["Mb"] if info["up"] else info["Nb"], info["Nb"])
    plan.close()
    P, Q, hop = info["P"], info["Q"], info["hop"]
    worst = {s: 0.0 for s in forms}
    peak = 0.0
    for b in (1, info["nblk"] // 2):   # interior blocks of channel 1
        blk = staged[1, b * hop * Q: b * hop * Q + info["Nb"]]
        ref = hb.block_output(blk, spec, P, Q, info["nb"], info["marg"], bool(info["up"]))
        d0 = b * hop * P - info["o"]
        i0, i1 = max(0, -d0), min(len(ref), nout - d0)
        r = ref[i0:i1]
        peak = max(peak, float(np.abs(hb.to_double(r)).max()))
        for s, pairs in forms.items():
            hi, lo = pairs[1, d0 + i0: d0 + i1, 0], pairs[1, d0 + i0: d0 + i1, 1]
            diff = hb.fix(hi) + hb.fix(lo) - r
            worst[s] = max(worst[s], max(math.ldexp(float(abs(int(v))), -hb.F) for v in diff))
    rel = {s: worst[s] / peak for s in worst}
    say(f"{fi} -> {fo}: dd output error / peak: " +
        ", ".join(f"{'plain' if s == 0 else f'split S={s}'} 2^{math.log2(v + 1e-300):.1f}" for s, v in rel.items()))
    assert rel[0] <= 2.0 ** -100, "plain dd beyond 2^-100 of peak"
    assert rel[15] <= 2.0 ** -100, "split form at 15 slices beyond 2^-100 of peak"
    assert rel[8] > rel[15], "8 slices should be coarser than 15"


@pytest.mark.parametrize("fi,fo,frames,env", [(96000, 32001, 9000, None), (44100, 44101, 6000, None),
                                              (48000, 44100, 20000, "KBSRC_GENERIC_PASSES"),
                                              (48000, 44100, 20000, "KBSRC_GLOBAL_PASSES"),
                                              (22050, 48001, 5000, None)])
def test_dd_pairs_round_to_the_output_on_every_route(fi, fo, frames, env, monkeypatch):
    """Every kernel that writes output samples writes the pairs too: Bluestein post-multiplies (96k -> 32001,
    22.05k -> 48001), the general fold's generic passes (44.1k -> 44.101k), every pass through the generic kernel and
    through global scratch; hi + lo rounds to the ordinary execute's sample in each."""
    import torch

    import kbsrc_amd as k
    for v in ("KBSRC_GENERIC_PASSES", "KBSRC_GLOBAL_PASSES"):
        monkeypatch.delenv(v, raising=False)
    if env:
        monkeypatch.setenv(env, "1")
    x = np.random.default_rng(fi + fo).uniform(-0.5, 0.5, (frames, 2))
    plan = k.Plan(frames, 2, fi, fo, "exact", length=0.05)
    info = plan.info
    nout = info["nout"]
    d_pad = torch.zeros(2 * info["xstride"], dtype=torch.float64, device="cuda")
    d_out = torch.empty(2 * nout, dtype=torch.float64, device="cuda")
    d_pairs = torch.full((4 * nout,), float("nan"), dtype=torch.float64, device="cuda")
    plan.stage(x, d_pad.data_ptr())
    plan.execute(d_pad.data_ptr(), d_out.data_ptr())
    plan.execute_dd(d_pad.data_ptr(), d_pairs.data_ptr(), 0)
    torch.cuda.synchronize()
    plan.close()
    say(f"{fi} -> {fo} {env or 'default route'}: {nout} samples per channel")
    y = d_out.cpu().numpy().reshape(2, nout)
    p = d_pairs.cpu().numpy().reshape(2, nout, 2)
    assert np.isfinite(p).all(), "a pair left unwritten"
    assert np.array_equal(p[..., 0] + p[..., 1], y)
    assert (np.abs(p[..., 1]) <= np.abs(np.spacing(p[..., 0])) / 2).all(), "pairs are normalised"
