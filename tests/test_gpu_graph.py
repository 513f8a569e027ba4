"""Captured executes (kbsrc_plan_capture / kbsrc_batch_capture, kbsrc_graph_launch): a plan's and a batch's execute
captured once as a HIP graph, replayed several times with the output buffer cleared by a fill kernel and other eager
work and synchronisations between replays; every replay must give the eager execute's bits. Needs an MI355X."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_plan_graph_replays_match_eager(mode):
    import torch

    import kbsrc_amd as k
    rng = np.random.default_rng(7)
    x = rng.uniform(-0.5, 0.5, (3 * 96000, 2))
    plan = k.Plan(x.shape[0], 2, 96000, 44100, mode)
    other = k.Plan(48000, 1, 48000, 44100, mode)
    info, oinfo = plan.info, other.info
    d_pad = torch.zeros(2 * info["xstride"], dtype=torch.float64, device="cuda")
    d_ref = torch.empty(2 * info["nout"], dtype=torch.float64, device="cuda")
    d_out = torch.empty_like(d_ref)
    o_pad = torch.zeros(oinfo["xstride"], dtype=torch.float64, device="cuda")
    o_out = torch.empty(oinfo["nout"], dtype=torch.float64, device="cuda")
    plan.stage(x, d_pad.data_ptr())
    other.stage(rng.uniform(-0.5, 0.5, 48000), o_pad.data_ptr())
    plan.execute(d_pad.data_ptr(), d_ref.data_ptr())
    torch.cuda.synchronize()
    ref = d_ref.cpu().numpy()
    g = plan.capture(d_pad.data_ptr(), d_out.data_ptr())
    say(f"{mode}: captured; replaying")
    for rep in range(4):
        d_out.zero_()
        other.execute(o_pad.data_ptr(), o_out.data_ptr())   # other work on the device between replays
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(d_out.cpu().numpy(), ref), (mode, rep)
    # an eager execute of the same plan after the replays still gives the same bits
    plan.execute(d_pad.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(d_out.cpu().numpy(), ref)
    g.close()
    plan.close()
    other.close()


def test_batch_graph_replays_match_eager():
    import torch

    import kbsrc_amd as k
    rng = np.random.default_rng(8)
    jobs = [(96000, 2, 96000, 44100), (44100, 1, 44100, 48000), (30000, 2, 48000, 48000), (88200, 6, 88200, 96000)]
    b = k.Batch(jobs, "exact")
    d_in = torch.zeros(b.in_doubles, dtype=torch.float64, device="cuda")
    d_ref = torch.zeros(b.out_doubles, dtype=torch.float64, device="cuda")
    d_out = torch.zeros_like(d_ref)
    for j, (fr, ch, _, _) in enumerate(jobs):
        b.stage(j, rng.uniform(-0.5, 0.5, (fr, ch)), d_in.data_ptr())
    b.execute(d_in.data_ptr(), d_ref.data_ptr())
    torch.cuda.synchronize()
    ref = d_ref.cpu().numpy()
    g = b.capture(d_in.data_ptr(), d_out.data_ptr())
    say("batch: captured; replaying")
    for rep in range(3):
        d_out.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(d_out.cpu().numpy(), ref), rep
        b.execute(d_in.data_ptr(), d_ref.data_ptr())   # eager work between replays
        torch.cuda.synchronize()
    g.close()
    b.close()
