"""Streamed I/O around the GPU engine: output packing on the GPU (src/io/wav.cpp:52-86,235-250), staging from a
device-resident input, and the command line's streamed read -> design/upload overlap -> packed write. Each is
checked against the library's own results (which the parity tests pin to the reference). Needs an MI355X."""
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


BYTES = {"f64": 8, "f32": 4, "s32": 4, "s24": 3, "s16": 2}


def quantise(y, fmt):
    """Interleaved sample bytes of planar y as the reference's writer makes them without dither (src/io/wav.cpp:
    52-86): scale by 2^(8B-1), round to nearest even, clip counting, NaN -> 0, little-endian. Returns (bytes,
    clipped)."""
    x = np.ascontiguousarray(np.asarray(y, dtype=np.float64).T).reshape(-1)
    if fmt == "f64":
        return x.astype("<f8").tobytes(), 0
    if fmt == "f32":
        return x.astype("<f4").tobytes(), 0
    B = BYTES[fmt]
    sc = float(2 ** (8 * B - 1))
    with np.errstate(invalid="ignore"):
        v = np.rint(x * sc)
        hi, lo = v > sc - 1, v < -sc
    v = np.where(hi, sc - 1, np.where(lo, -sc, v))
    v = np.where(np.isnan(v), 0.0, v)
    iv = v.astype(np.int64)
    b = ((iv[:, None] >> (8 * np.arange(B))) & 255).astype(np.uint8)
    return b.tobytes(), int(hi.sum() + lo.sum())


def wav_data(blob):
    """(channels, data bytes) of a WAV file."""
    pos, ch = 12, None
    while pos + 8 <= len(blob):
        tag, size = blob[pos:pos + 4], struct.unpack("<I", blob[pos + 4:pos + 8])[0]
        if tag == b"fmt ":
            ch = struct.unpack("<H", blob[pos + 10:pos + 12])[0]
        if tag == b"data":
            return ch, blob[pos + 8:pos + 8 + size]
        pos += 8 + size + (size & 1)
    raise AssertionError("no data chunk")


def pcm_wav(x_interleaved, ch, rate, B):
    """A plain PCM WAV (16-byte fmt) of B-byte integer samples x * 2^(8B-1), already integral."""
    iv = np.asarray(x_interleaved, dtype=np.int64)
    data = ((iv[:, None] >> (8 * np.arange(B))) & 255).astype(np.uint8).tobytes()
    ext = ch > 2 or B > 2
    if ext:   # WAVE_FORMAT_EXTENSIBLE with the PCM sub-format
        fmt = struct.pack("<HHIIHHHHI", 0xFFFE, ch, rate, rate * ch * B, ch * B, 8 * B, 22, 8 * B, 0)
        fmt += struct.pack("<H", 1) + bytes([0, 0, 0, 0, 0x10, 0, 0x80, 0, 0, 0xAA, 0, 0x38, 0x9B, 0x71])
    else:
        fmt = struct.pack("<HHIIHH", 1, ch, rate, rate * ch * B, ch * B, 8 * B)
    body = b"WAVE" + b"fmt " + struct.pack("<I", len(fmt)) + fmt + b"data" + struct.pack("<I", len(data)) + data
    if len(data) & 1:
        body += b"\0"
    return b"RIFF" + struct.pack("<I", len(body)) + body


@pytest.mark.parametrize("fmt", ["f64", "f32", "s32", "s24", "s16"])
def test_pack_matches_reference_quantiser(fmt):
    import torch
    say(f"pack {fmt}")
    ch, nout, f0, nf = 3, 5000, 123, 4000
    rng = np.random.default_rng(3)
    y = rng.uniform(-1.3, 1.3, (ch, nout))
    sc = float(2 ** (8 * BYTES.get(fmt, 4) - 1))
    y[0, 200:210] = (np.arange(10) - 5 + 0.5) / sc       # ties: round half to even
    y[1, 300:306] = [np.nan, np.inf, -np.inf, 1.0, -1.0, 1 - 2.0 ** -40]
    y[2, 400:403] = [5e-324, -0.0, 2.0 ** -1074 * 3]
    d_y = torch.from_numpy(y).cuda()
    d_b = torch.zeros(nf * ch * BYTES[fmt] + 8, dtype=torch.uint8, device="cuda")
    d_c = torch.zeros(1, dtype=torch.int64, device="cuda")
    gpu().pack(d_y.data_ptr(), ch, nout, f0, nf, fmt, d_b.data_ptr(), d_c.data_ptr())
    torch.cuda.synchronize()
    want, clipped = quantise(y[:, f0:f0 + nf], fmt)
    assert d_b.cpu().numpy()[:len(want)].tobytes() == want
    assert int(d_c.item()) == clipped


def test_pack_rejects_bad_ranges():
    with pytest.raises(ValueError):
        gpu().pack(0, 1, 10, 5, 6, "s16", 0, 0)   # past the end
    with pytest.raises(ValueError):
        gpu().pack(0, 0, 10, 0, 1, "s16", 0, 0)   # no channels


@pytest.mark.parametrize("frames,ch,lpc", [(30000, 2, True), (3000, 1, True), (2000, 3, False), (1, 1, True)])
def test_stage_device_equals_stage(frames, ch, lpc):
    import torch
    say(f"stage_device {frames}x{ch} lpc={lpc}")
    x = np.random.default_rng(frames).uniform(-0.5, 0.5, (frames, ch))
    plan = gpu().Plan(frames, ch, 96000, 44100, "exact", length=0.05, lpc_edges=lpc)
    n = ch * plan.info["xstride"]
    a = torch.full((n,), 7.0, dtype=torch.float64, device="cuda")
    b = torch.full((n,), -7.0, dtype=torch.float64, device="cuda")
    plan.stage(x, a.data_ptr())
    d_x = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    plan.stage_device(d_x.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    tot = plan.info["tot"]
    for c in range(ch):
        s = c * plan.info["xstride"]
        assert a[s:s + tot].cpu().numpy().tobytes() == b[s:s + tot].cpu().numpy().tobytes()
    plan.close()


def run_cli(tmp_path, src, args):
    out = tmp_path / "out.wav"
    r = subprocess.run([H.GPU_CLI, *args, str(src), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out.read_bytes(), r.stderr


CLI_CASES = [
    # frames, channels, rate_in, rate_out, mode, format, amplitude
    (0, 2, 48000, 44100, "exact", "f64", 0.5),
    (1, 1, 48000, 44100, "exact", "s16", 0.5),
    (3000, 5, 44100, 48000, "exact", "s24", 0.5),
    (20000, 1, 96000, 44100, "fast", "s16", 1.6),    # clipping
    (7000, 2, 44100, 44100, "exact", "f32", 0.5),    # equal rates: a copy
    (9000, 2, 22050, 16000, "exact", "s32", 0.9),
]


@pytest.mark.parametrize("case", CLI_CASES, ids=lambda c: f"{c[0]}x{c[1]}-{c[2]}-{c[3]}-{c[4]}-{c[5]}")
def test_cli_streamed_equals_library(case, tmp_path):
    frames, ch, rin, rout, mode, fmt, amp = case
    say(f"cli {case}")
    x = np.random.default_rng(frames + ch).uniform(-amp, amp, (frames, ch))
    src = tmp_path / "in.wav"
    src.write_bytes(H.wav_f64(x, ch, rin))
    blob, err = run_cli(tmp_path, src, ["--rate", str(rout), "--mode", mode, "--format", fmt, "--len", "0.05"])
    y = gpu().resample(x, rin, rout, mode, length=0.05)
    want, clipped = quantise(y, fmt)
    got_ch, data = wav_data(blob)
    assert got_ch == ch
    assert data == want
    if clipped:
        assert f"{clipped} samples clipped" in err
    else:
        assert "clipped" not in err


@pytest.mark.parametrize("B", [2, 3])
def test_cli_reads_pcm_input(B, tmp_path):
    """s16 / s24 input decoded in streamed chunks (src/io/wav.cpp:150-170)."""
    say(f"cli pcm{8 * B} input")
    sc = 2 ** (8 * B - 1)
    iv = np.random.default_rng(B).integers(-sc, sc, 2 * 6000)
    src = tmp_path / "in.wav"
    src.write_bytes(pcm_wav(iv, 2, 48000, B))
    blob, _ = run_cli(tmp_path, src, ["--rate", "44100"])
    y = gpu().resample((iv / sc).reshape(-1, 2), 48000, 44100)
    assert wav_data(blob)[1] == quantise(y, "f64")[0]


def test_cli_reports_errors_in_reference_order(tmp_path):
    """A rejected conversion exits 1 without an output file, with the reference's messages."""
    src = tmp_path / "in.wav"
    src.write_bytes(H.wav_f64(np.zeros(2000), 1, 8000))
    out = tmp_path / "out.wav"
    r = subprocess.run([H.GPU_CLI, "--rate", "44100", "--len", "0.0001", str(src), str(out)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "length too short for these sample rates" in r.stderr
    assert not out.exists()
    r = subprocess.run([H.GPU_CLI, "--rate", "44100", str(tmp_path / "missing.wav"), str(out)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "cannot open input" in r.stderr


def test_host_buffer_conversions_from_two_threads():
    """kbsrc_resample from two threads at once, each input larger than the pinned staging pair (2 x 64 MB), gives
    each thread the bits it gets alone."""
    from concurrent.futures import ThreadPoolExecutor

    import kbsrc_amd
    frames = 96000 * 90   # stereo: 138 MB of input per conversion
    xs = [np.random.default_rng(40 + k).uniform(-0.5, 0.5, (frames, 2)) for k in range(2)]
    print("two-thread host conversions: 2 x 138 MB", flush=True)
    alone = [kbsrc_amd.resample(x, 96000, 44100, "exact") for x in xs]
    with ThreadPoolExecutor(2) as ex:
        both = list(ex.map(lambda x: kbsrc_amd.resample(x, 96000, 44100, "exact"), xs))
    for a, b in zip(alone, both):
        assert a.tobytes() == b.tobytes()


def test_cached_device_blocks_across_conversions():
    """kbsrc_resample keeps its released device buffers for later calls (each reused only after the work that last
    used it): conversions of growing, shrinking and other-rate sizes in turn, from one thread and then from three
    at once, give the bits of the oracle (short ones) or of the same conversion run first (long ones)."""
    from concurrent.futures import ThreadPoolExecutor

    import kbsrc_amd
    cases = [(96000, 44100, 2, 96000 * 30), (44100, 48000, 1, 5000), (96000, 44100, 2, 96000 * 60),
             (48000, 96000, 3, 3000), (96000, 44100, 2, 96000 * 10), (44100, 48000, 2, 44100 * 40)]
    xs = [np.random.default_rng(70 + k).uniform(-0.5, 0.5, (f, ch)) for k, (_, _, ch, f) in enumerate(cases)]
    print("cached device blocks: six conversions in turn, then three threads", flush=True)
    first = []
    for (rin, rout, ch, f), x in zip(cases, xs):
        y = kbsrc_amd.resample(x, rin, rout, "exact")
        if f <= 5000:
            assert y.tobytes() == H.oracle_resample(x.reshape(-1), ch, rin, rout, "exact").tobytes()
        first.append(y)
    for (rin, rout, _, _), x, y in zip(cases, xs, first):
        assert kbsrc_amd.resample(x, rin, rout, "exact").tobytes() == y.tobytes()
    with ThreadPoolExecutor(3) as ex:
        again = list(ex.map(lambda k: kbsrc_amd.resample(xs[k], cases[k][0], cases[k][1], "exact"),
                            [2, 0, 5, 1, 4, 3]))
    for k, y in zip([2, 0, 5, 1, 4, 3], again):
        assert y.tobytes() == first[k].tobytes(), k
