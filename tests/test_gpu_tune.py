"""Measured kernel shapes (kbsrc_plan_tune, kbsrc_batch_tune): tuning runs many shapes of every pass, keeps the
fastest, and never changes the output bits. Needs an MI355X."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def gpu():
    import kbsrc_amd
    return kbsrc_amd


@pytest.mark.parametrize("rin,rout,mode", [(96000, 44100, "exact"), (88200, 192000, "exact"), (176400, 96000, "fast"),
                                           (44100, 44101, "exact")])
def test_tuned_plan_gives_the_same_bits(rin, rout, mode):
    import torch
    frames, ch = 40000, 2
    x = np.random.default_rng(rin + rout).uniform(-0.5, 0.5, (frames, ch))
    plan = gpu().Plan(frames, ch, rin, rout, mode, length=0.05)
    info = plan.info
    xx = torch.empty(ch * info["xstride"], dtype=torch.float64, device="cuda")
    plan.stage(x, xx.data_ptr())
    outs = []
    for step in ("before", "tune", "after"):
        out = torch.full((ch * info["nout"],), np.nan, dtype=torch.float64, device="cuda")
        if step == "tune":
            plan.tune(xx.data_ptr(), out.data_ptr())
        else:
            plan.execute(xx.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy().tobytes())
    assert outs[0] == outs[1] == outs[2]
    assert gpu().tuned_shapes() > 0
    want = H.oracle_resample(x.reshape(-1), ch, rin, rout, mode, length=0.05)
    assert outs[2] == want.tobytes()
    plan.close()


def test_tuned_batch_gives_the_same_bits():
    import torch
    jobs = [(30000, 2, 96000, 44100), (20000, 1, 48000, 96000), (9000, 3, 22050, 16000), (5000, 1, 96000, 32001)]
    xs = [np.random.default_rng(j).uniform(-0.5, 0.5, (f, c)) for j, (f, c, _, _) in enumerate(jobs)]
    b = gpu().Batch(jobs, "exact", length=0.05)
    d_in = torch.zeros(b.in_doubles, dtype=torch.float64, device="cuda")
    for j, x in enumerate(xs):
        b.stage(j, x, d_in.data_ptr())
    res = []
    for step in ("execute", "tune", "execute"):
        d_out = torch.full((b.out_doubles,), np.nan, dtype=torch.float64, device="cuda")
        getattr(b, step)(d_in.data_ptr(), d_out.data_ptr())
        torch.cuda.synchronize()
        res.append(d_out.cpu().numpy())
    for L in b.layout:
        sl = slice(L["out_offset"], L["out_offset"] + L["out_doubles"])
        assert res[0][sl].tobytes() == res[1][sl].tobytes() == res[2][sl].tobytes()
    b.close()
