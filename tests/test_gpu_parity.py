"""GPU engine parity: bit-identical to the reference's pins and to the oracle (CPU restatement) on the same
inputs, through the C ABI. Needs an MI355X."""
import os
import struct
import subprocess
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


@pytest.mark.parametrize("name,rate,mode,want", H.CLI_PINS)
def test_reference_cli_pins(name, rate, mode, want):
    x = H.primary97() if name == "primary" else H.subnormal97()
    say(f"pin {name} 48000->{rate} {mode}")
    y = gpu().resample(x.reshape(-1, 2), 48000, rate, mode)
    assert H.wav_sha_planar(y, rate) == want


def test_frame1_and_subnormal_planar_pins():
    y = gpu().resample(np.array([H.FRAME1]), 8000, 16000, "exact", length=0.01)
    assert [struct.pack(">d", v).hex() for v in y.reshape(-1)] == H.FRAME1_OUT
    y = gpu().resample(H.subnormal97().reshape(-1, 2), 48000, 96000, "exact", length=0.1)
    assert H.sha(y.astype("<f8").tobytes()) == H.SUBNORMAL_96K_PLANAR


@pytest.mark.parametrize("mode,lpc", [("exact", True), ("fast", True), ("exact", False)])
def test_48k_to_44k1_pins(mode, lpc):
    say(f"48k->44.1k mono 1 s {mode} lpc={lpc}")
    y, st = gpu().resample(H.input_48k_1s(), 48000, 44100, mode, lpc_edges=lpc, return_stats=True)
    assert st.block == 250 and st.p == 147 and st.q == 160 and st.isa == b"gfx950"
    assert H.wav_sha_planar(y, 44100) == H.PINS_48K_1S[(mode, lpc)]


CASES = [
    # rin, rout, ch, frames, mode, length, block, lpc
    (48000, 96000, 2, 3000, "exact", 0.2, 0, True),
    (48000, 44100, 1, 20000, "fast", 0.2, 0, True),
    (96000, 44100, 2, 30000, "exact", 0.05, 0, True),
    (96000, 44100, 2, 30000, "fast", 0.05, 0, True),
    (44100, 48000, 2, 9000, "exact", 0.1, 0, False),
    (44100, 96000, 1, 5000, "fast", 0.2, 0, True),
    (8000, 44100, 1, 4000, "exact", 0.1, 0, True),
    (44100, 8000, 3, 12000, "exact", 0.1, 0, True),
    (192000, 44100, 1, 40000, "exact", 0.02, 0, True),
    (22050, 16000, 2, 7000, "fast", 0.1, 0, True),
    (48000, 96000, 2, 64, "exact", 0.1, 9605, True),
    (48000, 96000, 2, 64, "fast", 0.1, 20000, True),
    (48000, 44100, 1, 10, "exact", 0.2, 0, True),
    (48000, 44100, 1, 1, "exact", 0.2, 0, True),
    (48000, 44100, 1, 0, "exact", 0.2, 0, True),
    (48000, 44100, 6, 2000, "fast", 0.01, 0, True),
    (16000, 48000, 1, 3000, "exact", 0.01, 0, True),
    (48000, 16000, 2, 1000, "exact", 0.002, 0, True),
    (44100, 88200, 1, 20000, "exact", 0.05, 6000, True),
    # lengths with prime factors above 7 (Bluestein) and near-unity ratios (general fold)
    (96000, 32001, 1, 5000, "exact", 0.02, 0, True),
    (44100, 44101, 1, 3000, "exact", 0.02, 0, True),
    (48000, 48001, 1, 3000, "fast", 0.02, 0, False),
    (22050, 48001, 1, 3000, "exact", 0.02, 0, True),
    (1001, 1000, 1, 4000, "exact", 0.2, 0, True),
    (44100, 44200, 2, 3000, "fast", 0.02, 0, True),
    (48000, 44101, 1, 3000, "exact", 0.02, 0, True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}-{c[2]}ch-{c[3]}-{c[4]}-L{c[5]}-b{c[6]}")
def test_gpu_equals_oracle(case):
    rin, rout, ch, frames, mode, length, block, lpc = case
    say(f"case {case}")
    x = np.random.default_rng(frames + rin + rout).uniform(-0.5, 0.5, frames * ch)
    want = H.oracle_resample(x, ch, rin, rout, mode, length=length, block=block, lpc=lpc)
    got = gpu().resample(x.reshape(frames, ch), rin, rout, mode, length=length, block=block, lpc_edges=lpc)
    assert got.shape == want.shape
    if got.tobytes() != want.tobytes():
        bad = np.flatnonzero(got.reshape(-1) != want.reshape(-1))
        pytest.fail(f"{len(bad)} of {got.size} samples differ, first at {bad[:5]}: "
                    f"{got.reshape(-1)[bad[:3]]} vs {want.reshape(-1)[bad[:3]]}")


def test_equal_rates_copy_bit_exact():
    x = np.random.default_rng(5).standard_normal((4410, 2)) * 0.1
    y = gpu().resample(x, 44100, 44100)
    assert np.array_equal(y, x.T)


def test_gpu_cli_pins(tmp_path):
    """The kbsrc command line over the GPU engine writes the pinned WAV bytes (src/io/wav.cpp layout)."""
    src = tmp_path / "primary.wav"
    src.write_bytes(H.wav_f64(H.primary97(), 2, 48000))
    for name, rate, mode, want in H.CLI_PINS[:3]:
        out = tmp_path / "out.wav"
        r = subprocess.run([H.GPU_CLI, "--rate", str(rate), "--mode", mode, "-v", str(src), str(out)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert "engine=gfx950" in r.stderr
        assert H.sha(out.read_bytes()) == want


def _run_96k_180s(mode):
    """Full benchmark workload: stereo 96 kHz -> 44.1 kHz, 180 s, through plan/stage/execute on device buffers."""
    import torch
    x = H.input_96k_180s()
    assert H.sha(H.wav_f64(x, 2, 96000)) == H.IN_96K_180S
    with gpu().Plan(len(x), 2, 96000, 44100, mode) as plan:
        info = plan.info
        xx = torch.empty(2 * info["xstride"], dtype=torch.float64, device="cuda")
        out = torch.empty(2 * info["nout"], dtype=torch.float64, device="cuda")
        plan.stage(x, xx.data_ptr())
        plan.execute(xx.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        return out.cpu().numpy().reshape(2, info["nout"]), info


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_96k_to_44k1_180s_pins(mode):
    say(f"96k->44.1k stereo 180 s {mode}")
    y, info = _run_96k_180s(mode)
    assert y.shape == (2, 7_938_000)
    assert H.wav_sha_planar(y, 44100) == H.PINS_96K_180S[f"{mode}_f64"]


def test_96k_to_44k1_cli_files_bit_exact(tmp_path):
    """The benchmark workload through the streamed CLI (decode and upload behind the GPU design, output packed and
    quantised on the GPU, no dither): f64, s24, s16 and s32 files match the reference's bytes."""
    say("96k->44.1k stereo 180 s CLI f64/s24/s16/s32")
    src = tmp_path / "in96k.wav"
    with open(src, "wb") as f:
        f.write(H.wav_f64(H.input_96k_180s(), 2, 96000))
    for fmt in ("f64", "s24", "s16", "s32"):
        out = tmp_path / f"out_{fmt}.wav"
        r = subprocess.run([H.GPU_CLI, "--rate", "44100", "--format", fmt, "-v", str(src), str(out)],
                           capture_output=True, text=True, timeout=110)
        assert r.returncode == 0, r.stderr
        say("  " + fmt + ": " + r.stderr.strip().splitlines()[-1])
        assert H.sha(out.read_bytes()) == H.PINS_96K_180S[f"exact_{fmt}"], fmt
        os.remove(out)


def test_44k1_to_48k_180s_pin():
    say("44.1k->48k stereo 180 s")
    x = H.input_44k_180s()
    y = gpu().resample(x, 44100, 48000, "exact")
    assert H.wav_sha_planar(y, 48000) == H.PIN_44K_TO_48K_180S


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_long_subfft_extreme_ratio_pin(mode):
    """10 MHz -> 48 kHz at --len 0.5: the exact engine's 3150 x 3200 four-step (100 KB of LDS per sub-FFT, one
    per workgroup in the specialised passes); both modes give the reference's bits."""
    import hashlib
    say(f"10M->48k len 0.5 {mode}")
    y = gpu().resample(H.input_10m_short().reshape(-1, 1), 10000000, 48000, mode, length=0.5)
    assert y.shape == (1, 960)
    assert hashlib.sha256(y.tobytes()).hexdigest() == H.PINS_10M_SHORT[mode]


@pytest.mark.parametrize("mode,var,val", [(m, v, x) for m in ("exact", "fast")
                                          for v, x in (("KBSRC_LONG_PARTS", "0"), ("KBSRC_LONG_VARIANT", "0"),
                                                       ("KBSRC_LONG_VARIANT", "1"), ("KBSRC_LONG_ORDER", "0"))])
def test_long_column_parts_give_the_one_pass_bits(mode, var, val, monkeypatch):
    """The forward columns of the 10 MHz -> 48 kHz plans run as two passes through the row pass's buffer (exact:
    3150 points as an 18-point head on 175 columns and a 175-point tail on 18 classes; fast: 3125 points as 25 x 125
    or 125 x 25): each shape and workgroup order of that route, and the one-pass kernel (KBSRC_LONG_PARTS=0), give
    the reference's bits; on two channels of several blocks each, the same bits as the one-pass kernel."""
    import hashlib
    for v in ("KBSRC_LONG_PARTS", "KBSRC_LONG_VARIANT", "KBSRC_LONG_ORDER"):
        monkeypatch.delenv(v, raising=False)
    x2 = np.random.default_rng(11).uniform(-0.5, 0.5, (700000, 2))
    monkeypatch.setenv("KBSRC_LONG_PARTS", "0")
    ref2 = gpu().resample(x2, 10000000, 48000, mode, length=0.5)
    monkeypatch.delenv("KBSRC_LONG_PARTS")
    monkeypatch.setenv(var, val)
    say(f"10M->48k two-part columns {mode} {var}={val}")
    y = gpu().resample(H.input_10m_short().reshape(-1, 1), 10000000, 48000, mode, length=0.5)
    assert hashlib.sha256(y.tobytes()).hexdigest() == H.PINS_10M_SHORT[mode]
    y2 = gpu().resample(x2, 10000000, 48000, mode, length=0.5)
    assert y2.tobytes() == ref2.tobytes()


def _rowmid_run(x, rin, rout, mode, block, lpc=True):
    """One conversion through plan/stage/execute_timed: (output, pass times, plan)."""
    import torch
    with gpu().Plan(len(x), x.shape[1], rin, rout, mode, block=block, lpc_edges=lpc) as plan:
        info = plan.info
        xx = torch.empty(x.shape[1] * info["xstride"], dtype=torch.float64, device="cuda")
        out = torch.empty(x.shape[1] * info["nout"], dtype=torch.float64, device="cuda")
        plan.stage(x, xx.data_ptr())
        t = plan.execute_timed(xx.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        return out.cpu().numpy(), t, info


@pytest.mark.parametrize("rin,rout,mode,block,nvar", [(96000, 44100, "exact", 630, 7), (96000, 44100, "fast", 250, 6),
                                                      (88200, 44100, "exact", 89600, 1), (176400, 44100, "exact", 89600, 1),
                                                      (192000, 48000, "exact", 96040, 1), (192000, 96000, "exact", 192080, 1),
                                                      (176400, 48000, "exact", 2430, 1)])
def test_row_pass_with_split_fold_gives_the_separate_bits(rin, rout, mode, block, nvar, monkeypatch):
    """The forward row pass fused with the split/multiply/fold step (one launch, the forward spectrum kept in LDS)
    against the row pass and the split/fold kernel it replaces (KBSRC_ROWMID=0), every registered workgroup shape, on
    the registered transform shapes (96 kHz -> 44.1 kHz exact: 315 x 320 forward, row pairs; fast: 200 x 200, units of
    16 rows; the other downsampling pairs' shapes of long inputs, at their block sizes) with 1, 2 and 3 channels of
    several blocks, LPC edges on and off."""
    monkeypatch.delenv("KBSRC_ROWMID", raising=False)
    monkeypatch.delenv("KBSRC_ROWMID_VARIANT", raising=False)
    rng = np.random.default_rng(23)
    cases = [(rng.uniform(-0.5, 0.5, (n, c)), lpc) for n, c, lpc in ((700001, 1, True), (1500000, 2, True),
                                                                      (400000, 3, False))]
    for x, lpc in cases:
        monkeypatch.setenv("KBSRC_ROWMID", "0")
        ref, t0, info = _rowmid_run(x, rin, rout, mode, block, lpc)
        assert t0["mid"] > 0
        monkeypatch.delenv("KBSRC_ROWMID")
        for v in range(nvar):
            monkeypatch.setenv("KBSRC_ROWMID_VARIANT", str(v))
            say(f"{rin}->{rout} fused row/split-fold {mode} {x.shape} lpc={lpc} variant {v} (Nb {info['Nb']})")
            y, t, _ = _rowmid_run(x, rin, rout, mode, block, lpc)
            assert t["mid"] == 0 and t["fwd_row"] > 0   # the fused launch ran
            assert y.tobytes() == ref.tobytes()
        monkeypatch.delenv("KBSRC_ROWMID_VARIANT")
        monkeypatch.setenv("KBSRC_BATCH_TASKS", "3")   # several launches per execute (tasks from a later offset)
        y, t, _ = _rowmid_run(x, rin, rout, mode, block, lpc)
        assert t["mid"] == 0 and y.tobytes() == ref.tobytes()
        monkeypatch.delenv("KBSRC_BATCH_TASKS")


def test_extreme_ratio_full_length_pins():
    """10 MHz -> 48 kHz, 30 s mono, --len 0.5 (the 5,000,001-tap design on the GPU), both modes."""
    import hashlib
    x = H.input_10m_30s().reshape(-1, 1)
    for mode in ("exact", "fast"):
        say(f"10M->48k 30 s len 0.5 {mode}")
        y = gpu().resample(x, 10000000, 48000, mode, length=0.5)
        assert y.shape == (1, 1440000)
        assert hashlib.sha256(y.tobytes()).hexdigest() == H.PINS_10M_30S[mode]


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_384k_8ch_pin(mode):
    """8 channels 384 kHz -> 44.1 kHz at --len 0.5 and the full-length plan's block (1536: a 1,966,080-point
    forward transform, 960 x 1024 four-step, inverse 336 x 336), on 2 s: the reference's bits."""
    import hashlib
    say(f"384k->44.1k 8 ch len 0.5 block 1536 {mode}")
    y = gpu().resample(H.input_384k(H.FRAMES_384K_SHORT), 384000, 44100, mode, length=0.5, block=1536)
    assert y.shape == (8, 88200)
    assert hashlib.sha256(y.tobytes()).hexdigest() == H.PINS_384K_SHORT[mode]


def test_384k_8ch_full_length_pin():
    """8 channels 384 kHz -> 44.1 kHz, 600 s (230,400,000 frames per channel, 14.7 GB of input), exact, --len 0.5:
    the reference's bits over all 1176 blocks."""
    import hashlib
    say("384k->44.1k 8 ch 600 s len 0.5 exact")
    x = H.input_384k(H.FRAMES_384K_FULL)
    y = gpu().resample(x, 384000, 44100, "exact", length=0.5)
    del x
    assert y.shape == (8, 26_460_000)
    assert hashlib.sha256(y.tobytes()).hexdigest() == H.PIN_384K_600S


@pytest.mark.parametrize("rin,rout,mode,length", [(10000000, 48000, "exact", 0.5), (192000, 8000, "exact", 0.2),
                                                   (48000, 8000, "fast", 0.2), (44100, 44101, "exact", 0.05),
                                                   (88200, 96000, "exact", 0.2)])
def test_unstored_bins_change_nothing(rin, rout, mode, length, monkeypatch):
    """The forward pass leaves unstored the spectrum bins the bin map never reads (above its last bin, and their
    split partners): same bits as storing every bin (KBSRC_STORE_ALL=1)."""
    x = np.random.default_rng(rin + rout).uniform(-0.5, 0.5, (200000 if rin == 10000000 else 3 * rin, 2))
    say(f"unstored bins {rin} -> {rout} {mode}")
    monkeypatch.delenv("KBSRC_STORE_ALL", raising=False)
    y = gpu().resample(x, rin, rout, mode, length=length)
    monkeypatch.setenv("KBSRC_STORE_ALL", "1")
    z = gpu().resample(x, rin, rout, mode, length=length)
    assert y.tobytes() == z.tobytes()


@pytest.mark.parametrize("rin,rout,mode", [(96000, 44100, "exact"), (96000, 44100, "fast"), (48000, 44100, "exact"),
                                           (192000, 44100, "exact")])
def test_split_fold_block_order_changes_nothing(rin, rout, mode, monkeypatch):
    """Down conversions run the split/fold kernel's bin blocks abs(h - Mb) apart back to back (the threads there
    share their z reads), in groups of 16 tasks; the launch order changes no value: same bits as the plain block
    order (KBSRC_MID_ORDER=0), as tasks fastest (KBSRC_MID_TGROUP=0) and as groups of 3 tasks (several groups, the
    last one partial)."""
    x = np.random.default_rng(rin + 7 * rout).uniform(-0.5, 0.5, (3 * rin, 2))
    say(f"split/fold block order {rin} -> {rout} {mode}")
    for v in ("KBSRC_MID_ORDER", "KBSRC_MID_TGROUP"):
        monkeypatch.delenv(v, raising=False)
    y = gpu().resample(x, rin, rout, mode)
    for var, val in (("KBSRC_MID_ORDER", "0"), ("KBSRC_MID_TGROUP", "0"), ("KBSRC_MID_TGROUP", "3")):
        monkeypatch.setenv(var, val)
        z = gpu().resample(x, rin, rout, mode)
        monkeypatch.delenv(var)
        assert y.tobytes() == z.tobytes(), (var, val)


GLOBAL_CASES = [c for c in CASES if c[3] >= 1000 and c[2] <= 2][::2]


@pytest.mark.parametrize("level", ["1", "2"], ids=["two-lds-passes", "per-stage"])
@pytest.mark.parametrize("case", GLOBAL_CASES, ids=lambda c: f"{c[0]}-{c[1]}-{c[2]}ch-{c[3]}-{c[4]}")
def test_global_passes_equal_oracle(case, level, monkeypatch):
    """Every sub-FFT pass forced through the global-memory forms (split into two LDS passes, or one launch per
    stage) gives the oracle's bits."""
    monkeypatch.setenv("KBSRC_GLOBAL_PASSES", level)
    rin, rout, ch, frames, mode, length, block, lpc = case
    say(f"global passes {case}")
    x = np.random.default_rng(frames + rin + rout).uniform(-0.5, 0.5, frames * ch)
    want = H.oracle_resample(x, ch, rin, rout, mode, length=length, block=block, lpc=lpc)
    got = gpu().resample(x.reshape(frames, ch), rin, rout, mode, length=length, block=block, lpc_edges=lpc)
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("env", [{"KBSRC_FUSED": "1"}, {"KBSRC_GENERIC_PASSES": "1"}, {"KBSRC_GLOBAL_PASSES": "1"}],
                         ids=["fused", "generic", "global"])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_engine_variants_agree(env, mode, monkeypatch):
    """The fused (split and bin map inside the passes) and the generic-pass engines give the pinned bits too."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    say(f"variant {env} {mode}")
    y = gpu().resample(H.input_48k_1s(), 48000, 44100, mode)
    assert H.wav_sha_planar(y, 44100) == H.PINS_48K_1S[(mode, True)]
    x = np.random.default_rng(11).uniform(-0.5, 0.5, 2 * 30000)
    want = H.oracle_resample(x, 2, 96000, 44100, mode, length=0.05)
    got = gpu().resample(x.reshape(-1, 2), 96000, 44100, mode, length=0.05)
    assert got.tobytes() == want.tobytes()
