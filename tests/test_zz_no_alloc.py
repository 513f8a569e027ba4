"""kbsrc_plan_execute_noalloc, the execute whose work buffers live in the library's static device array (no device
allocation), called through the C ABI: its output equals kbsrc_plan_execute's bit for bit on two calls in a row,
the second on a smaller plan (the array keeps the first call's contents); a plan whose one-task work buffers exceed
the array is rejected with a clear error. Needs an MI355X."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def lib():
    import kbsrc_amd
    L = ctypes.CDLL(kbsrc_amd.library_path())
    P = ctypes.c_void_p
    L.kbsrc_plan_execute_noalloc.argtypes = [P, P, P, P]
    L.kbsrc_plan_execute_noalloc.restype = ctypes.c_int
    L.kbsrc_plan_execute.argtypes = [P, P, P, P]
    L.kbsrc_plan_execute.restype = ctypes.c_int
    L.kbsrc_last_error.restype = ctypes.c_char_p
    return L


def run_both(L, frames, ch, rin, rout, precision, seed, length=0.2, block=0):
    import torch

    import kbsrc_amd
    say(f"noalloc vs execute: {ch} ch {rin}->{rout} {frames} frames {precision}")
    x = np.random.default_rng(seed).uniform(-0.5, 0.5, (frames, ch))
    plan = kbsrc_amd.Plan(frames, ch, rin, rout, precision, length=length, block=block)
    info = plan.info
    xx = torch.empty(ch * info["xstride"], dtype=torch.float64, device="cuda")
    plan.stage(x, xx.data_ptr())
    want = torch.full((ch * info["nout"],), float("nan"), dtype=torch.float64, device="cuda")
    got = torch.full_like(want, float("nan"))
    torch.cuda.synchronize()
    assert L.kbsrc_plan_execute(plan._h, xx.data_ptr(), want.data_ptr(), None) == 0
    rc = L.kbsrc_plan_execute_noalloc(plan._h, xx.data_ptr(), got.data_ptr(), None)
    assert rc == 0, L.kbsrc_last_error()
    torch.cuda.synchronize()
    assert bool(torch.isfinite(got).all())
    assert torch.equal(got, want)
    plan.close()
    return info


def test_noalloc_equals_execute_twice_second_smaller():
    L = lib()
    # 20 s stereo at the benchmark's rates: 50 tasks of 6.4 MB of work buffers, in several launches of the array
    big = run_both(L, 96000 * 20, 2, 96000, 44100, "exact", 1)
    assert big["tasks"] > 20
    # then a smaller plan: the array holds the previous call's data, which must not leak into this output
    small = run_both(L, 48000, 1, 48000, 44100, "exact", 2)
    assert small["tasks"] < big["tasks"]


def test_noalloc_fast_bluestein_and_general_fold():
    L = lib()
    run_both(L, 96000 * 5, 2, 96000, 44100, "fast", 3)
    # the benchmark's block: the forward rows and the split/fold as one kernel, 6.4 MB per task, several launches
    run_both(L, 96000 * 20, 2, 96000, 44100, "exact", 6, block=630)
    run_both(L, 96000 * 20, 2, 96000, 44100, "fast", 7, block=250)
    run_both(L, 20000, 1, 96000, 32001, "exact", 4, length=0.05)    # Bluestein lengths: four work buffers
    run_both(L, 20000, 2, 44100, 44101, "exact", 5, length=0.05)    # general fold


def test_noalloc_rejects_plans_beyond_the_array():
    import torch

    import kbsrc_amd
    # 10 MHz -> 48 kHz at --len 0.5: one task needs 2 x 322 MB of work buffers, beyond the 128 MiB array
    plan = kbsrc_amd.Plan(200000, 1, 10000000, 48000, "exact", length=0.5)
    xx = torch.zeros(plan.info["xstride"], dtype=torch.float64, device="cuda")
    out = torch.zeros(plan.info["nout"], dtype=torch.float64, device="cuda")
    L = lib()
    rc = L.kbsrc_plan_execute_noalloc(plan._h, xx.data_ptr(), out.data_ptr(), None)
    assert rc == 1
    assert b"too large for the static scratch" in L.kbsrc_last_error()
    plan.close()
