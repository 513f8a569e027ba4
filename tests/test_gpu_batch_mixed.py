"""The 64-job mixed batch at full size: lengths 0.5-300 s, 1/2/6/8 channels, rate pairs among 44.1/48/88.2/96/
176.4/192 kHz both ways, half noise and half sums of 8 sinusoids, exact mode (4.69 G input samples, 28 plan shapes).
The batch splits its larger groups into several launches over task tables. Every job's output must equal its own
single-plan execute bit for bit, and the same rate pairs at every channel count must equal the oracle. The task loop
the batch generalises is the reference's src/resampler.cpp:41-88. Needs an MI355X."""
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(900)
def test_mixed_batch_full_size_equals_single_plans():
    import torch

    import bench
    import kbsrc_amd
    jobs = bench.batch_jobs()
    say(f"mixed batch: {len(jobs)} jobs, {sum(j['frames'] * j['channels'] for j in jobs) / 1e9:.2f} G samples")
    b = kbsrc_amd.Batch([(j["frames"], j["channels"], j["rate_in"], j["rate_out"]) for j in jobs], "exact")
    assert b.groups >= 20
    d_in = torch.empty(b.in_doubles, dtype=torch.float64, device="cuda")
    d_out = torch.full((b.out_doubles,), float("nan"), dtype=torch.float64, device="cuda")
    for k, job in enumerate(jobs):
        x = bench.batch_signal(job)
        torch.cuda.synchronize()
        b.stage_device(k, x.data_ptr(), d_in.data_ptr())
        del x
    torch.cuda.synchronize()
    say("staged; executing the batch")
    b.execute(d_in.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    ref = torch.full_like(d_out, float("nan"))
    for k, (job, L) in enumerate(zip(jobs, b.layout)):
        plan = kbsrc_amd.Plan(job["frames"], job["channels"], job["rate_in"], job["rate_out"], "exact")
        assert plan.info["xstride"] * job["channels"] == L["in_doubles"]
        plan.execute(d_in.data_ptr() + 8 * L["in_offset"], ref.data_ptr() + 8 * L["out_offset"])
        torch.cuda.synchronize()
        plan.close()
        lo, hi = L["out_offset"], L["out_offset"] + L["out_doubles"]
        got, want = d_out[lo:hi], ref[lo:hi]
        assert bool(torch.isfinite(got).all()), f"job {k}: unwritten output"
        assert torch.equal(got, want), f"job {k} ({job['rate_in']}->{job['rate_out']}, {job['channels']} ch) differs"
    say("every job equals its own plan")
    b.close()


@pytest.mark.parametrize("channels", [1, 2, 6, 8])
def test_mixed_batch_rates_match_oracle(channels):
    """For each channel count, the first batch job with that count, cut to a 24,000-frame prefix and run as its own
    conversion (batch of one and single plan), equals the oracle bit for bit."""
    import torch

    import bench
    import kbsrc_amd
    job = next(j for j in bench.batch_jobs() if j["channels"] == channels)
    x = bench.batch_signal(job, max_frames=24000).cpu().numpy()
    say(f"{channels} ch {job['rate_in']}->{job['rate_out']} ({job['kind']}): 24000-frame prefix")
    want = H.oracle_resample(x.reshape(-1), channels, job["rate_in"], job["rate_out"], "exact")
    got = kbsrc_amd.resample(x, job["rate_in"], job["rate_out"], "exact")
    assert got.tobytes() == want.tobytes()
    got_b = kbsrc_amd.resample_batch([x], [job["rate_in"]], [job["rate_out"]], "exact")[0]
    assert got_b.tobytes() == want.tobytes()
    torch.cuda.synchronize()
