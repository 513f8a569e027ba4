"""Batches of jobs (kbsrc_batch_*): every job's output equals its own single-job conversion bit for bit, whatever
the mix of rate pairs, channel counts, lengths, equal-rate copies, Bluestein and general-fold shapes in the batch.
Needs an MI355X."""
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def gpu():
    import kbsrc_amd
    return kbsrc_amd


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


JOBS = [
    # frames, channels, rate_in, rate_out
    (30000, 2, 96000, 44100),
    (41000, 1, 96000, 44100),     # same rates: shares the first job's group when the block size agrees
    (5000, 6, 48000, 44100),
    (3000, 1, 44100, 48000),
    (0, 2, 48000, 44100),
    (7000, 2, 44100, 44100),      # equal rates: a copy
    (5000, 1, 96000, 32001),      # Bluestein lengths
    (3000, 1, 44100, 44101),      # general fold
    (120000, 2, 96000, 44100),
    (64, 1, 48000, 96000),
    (9000, 3, 22050, 16000),
]


def signals(jobs, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-0.5, 0.5, (f, c)) for f, c, _, _ in jobs]


# no Bluestein lengths (32001 and 44101 have prime factors above 7); 48000 -> 48020 (2401/2400) is a general fold
FUSABLE = [j for j in JOBS if j[3] not in (32001, 44101)] + [(7000, 2, 48000, 96000), (3000, 1, 48000, 48020)]


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("fuse", ["1", "0"])
def test_batch_one_launch_per_step(mode, fuse, monkeypatch):
    """The batch run as one launch per step over every group (opt-in: the generic pass kernel and the
    split/multiply/fold kernel over per-group descriptor tables) gives each job's own bits, as the per-group
    launches do."""
    import torch
    monkeypatch.setenv("KBSRC_BATCH_FUSE", fuse)
    say(f"batch {mode}, one launch per step = {fuse}")
    xs = signals(FUSABLE, 22)
    b = gpu().Batch([(x.shape[0], x.shape[1], ri, ro) for x, (_, _, ri, ro) in zip(xs, FUSABLE)], mode, length=0.05)
    assert b.fused == (fuse == "1")
    d_in = torch.zeros(b.in_doubles, dtype=torch.float64, device="cuda")
    d_out = torch.zeros(b.out_doubles, dtype=torch.float64, device="cuda")
    for j, x in enumerate(xs):
        b.stage(j, x, d_in.data_ptr())
    for rep in range(2):   # the descriptor tables are made on the first execute and reused
        d_out.zero_()
        b.execute(d_in.data_ptr(), d_out.data_ptr())
        torch.cuda.synchronize()
        out = d_out.cpu().numpy()
        for (f, c, ri, ro), x, L in zip(FUSABLE, xs, b.layout):
            y = out[L["out_offset"]:L["out_offset"] + L["out_doubles"]].reshape(c, L["nout"])
            want = gpu().resample(x, ri, ro, mode, length=0.05)
            assert y.tobytes() == want.tobytes(), (rep, f, c, ri, ro)
    g = b.capture(d_in.data_ptr(), d_out.data_ptr())   # and as a graph
    d_out.zero_()
    torch.cuda.synchronize()
    g.launch()
    torch.cuda.synchronize()
    assert d_out.cpu().numpy().tobytes() == out.tobytes()
    g.close()
    b.close()


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_batch_equals_single_jobs(mode):
    say(f"batch {mode}: {len(JOBS)} jobs")
    xs = signals(JOBS, 21)
    got = gpu().resample_batch(xs, [j[2] for j in JOBS], [j[3] for j in JOBS], mode, length=0.05)
    for (f, c, ri, ro), x, y in zip(JOBS, xs, got):
        want = gpu().resample(x, ri, ro, mode, length=0.05)
        assert y.shape == want.shape, (f, c, ri, ro)
        assert y.tobytes() == want.tobytes(), (f, c, ri, ro)


def test_batch_matches_oracle_and_groups():
    import torch
    jobs = [(30000, 2, 96000, 44100), (30000, 1, 96000, 44100), (4000, 1, 8000, 44100), (3000, 2, 44100, 48000)]
    xs = signals(jobs, 5)
    b = gpu().Batch(jobs, "exact", length=0.1)
    assert b.layout[0]["group"] == b.layout[1]["group"]   # same rates and block size: one engine
    assert b.groups == 3
    d_in = torch.zeros(b.in_doubles, dtype=torch.float64, device="cuda")
    d_out = torch.zeros(b.out_doubles, dtype=torch.float64, device="cuda")
    for j, x in enumerate(xs):
        if j % 2:   # half the jobs staged from device memory
            dx = torch.from_numpy(np.ascontiguousarray(x)).cuda()
            b.stage_device(j, dx.data_ptr(), d_in.data_ptr())
            torch.cuda.synchronize()
        else:
            b.stage(j, x, d_in.data_ptr())
    b.execute(d_in.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    for (f, c, ri, ro), x, L in zip(jobs, xs, b.layout):
        y = out[L["out_offset"]:L["out_offset"] + L["out_doubles"]].reshape(c, L["nout"])
        want = H.oracle_resample(x.reshape(-1), c, ri, ro, "exact", length=0.1)
        assert y.tobytes() == want.tobytes(), (f, c, ri, ro)
    # a second execute over the same buffers gives the same bits
    b.execute(d_in.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    assert d_out.cpu().numpy().tobytes() == out.tobytes()
    b.close()


def test_batch_rejections():
    with pytest.raises(ValueError, match="sample rates must be positive"):
        gpu().Batch([(1000, 1, 48000, 0)])
    with pytest.raises(ValueError, match="bad input buffer"):
        gpu().Batch([(1000, 0, 48000, 44100)])


def test_batch_graph_keeps_its_step_tables(monkeypatch):
    """A graph captured from a batch that runs one launch per step keeps the step tables it was captured with,
    even after executes on four other pairs of buffers have pushed them out of the batch's own cache."""
    import torch
    monkeypatch.setenv("KBSRC_BATCH_FUSE", "1")
    say("batch graph after step-table eviction")
    xs = signals(FUSABLE, 23)
    b = gpu().Batch([(x.shape[0], x.shape[1], ri, ro) for x, (_, _, ri, ro) in zip(xs, FUSABLE)], "exact", length=0.05)
    assert b.fused
    d_in = torch.zeros(b.in_doubles, dtype=torch.float64, device="cuda")
    for j, x in enumerate(xs):
        b.stage(j, x, d_in.data_ptr())
    outs = [torch.zeros(b.out_doubles, dtype=torch.float64, device="cuda") for _ in range(5)]
    b.execute(d_in.data_ptr(), outs[0].data_ptr())
    torch.cuda.synchronize()
    want = outs[0].clone()
    g = b.capture(d_in.data_ptr(), outs[0].data_ptr())
    for o in outs[1:]:   # four other buffer pairs: the captured pair's tables leave the batch's cache
        b.execute(d_in.data_ptr(), o.data_ptr())
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(o, want)
    outs[0].zero_()
    torch.cuda.synchronize()
    g.launch()
    torch.cuda.synchronize()
    assert torch.equal(outs[0], want)
    g.close()
    b.close()
