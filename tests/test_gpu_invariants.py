"""Accuracy invariants the reference's own test scripts check, run on the GPU engine: sine accuracy across common,
odd and near-unity rate pairs (tests/test_rates.py: below -200 dB), no aliasing or imaging in the continuous
impulse response (tests/test_alias.py: stopband below -280 dB, passband flat to 1e-12), and the command line's
formats, edge cases and exit codes (tests/test_formats.py). Written here from those checks; needs an MI355X."""
import math
import struct
import subprocess
import sys
import wave

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


RATE_PAIRS = [(96000, 44100), (44100, 48000), (48000, 44100), (96000, 48000), (44100, 96000), (192000, 44100),
              (44100, 88200), (8000, 44100), (44100, 8000), (48000, 192000),
              (44100, 44101), (44101, 44100), (48000, 44101), (96000, 32001), (22050, 48001), (48000, 48001),
              (44100, 44200)]


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_sine_accuracy_across_rate_pairs(mode):
    """2 s stereo (x, -x) of three sines (100 Hz, 1 kHz, 0.4 x the lower rate): the middle half of the output is
    within -200 dB of the same sines sampled at the output rate, and the output has ceil(n P / Q) frames."""
    rng = np.random.default_rng(1)
    worst = []
    for fi, fo in RATE_PAIRS:
        n = 2 * fi
        freqs, phases = [100.0, 1000.0, 0.4 * min(fi, fo)], rng.uniform(0, 6, 3)

        def sig(t):
            return sum(0.3 * np.sin(2 * np.pi * f * t + p) for f, p in zip(freqs, phases))

        x = sig(np.arange(n) / fi)
        y = gpu().resample(np.stack([x, -x], 1), fi, fo, mode)
        assert y.shape == (2, -(-n * fo // fi)), (fi, fo)
        want = sig(np.arange(y.shape[1]) / fo)
        mid = slice(y.shape[1] // 4, 3 * y.shape[1] // 4)
        err = max(np.abs(y[0, mid] - want[mid]).max(), np.abs(y[1, mid] + want[mid]).max())
        db = 20 * np.log10(err + 1e-300)
        worst.append((db, fi, fo))
        assert db < -200.0, f"{fi} -> {fo} {mode}: {db:.1f} dB"
    say(f"{mode}: worst {max(worst)[0]:.1f} dB")


@pytest.mark.parametrize("fi,fo", [(96000, 44100), (44100, 96000), (1001, 1000), (1000, 1001)])
@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_no_aliasing_or_imaging(fi, fo, mode):
    """Q unit impulses, one at each input phase modulo Q, interleave into the continuous response on the
    lcm(fi, fo) grid; its spectrum is below -280 dB from the lower Nyquist frequency up and flat to 1e-12 in the
    passband the converter reports."""
    length = 0.2
    g = math.gcd(fi, fo)
    P, Q, R = fo // g, fi // g, fi * (fo // g)
    reach = math.ceil(length * fi) + 1
    D = Q * math.ceil((2 * reach + 64) / Q)
    m = D * np.arange(1, Q + 1) + np.arange(Q)
    x = np.zeros(D * (Q + 2))
    x[m] = 1.0
    W = reach * P
    y, st = gpu().resample(x, fi, fo, mode, length=length, lpc_edges=False, return_stats=True)
    y = y[0]
    res = np.zeros(2 * W + 1)
    for mi in m:
        k = np.arange(-(-(mi * P - W) // Q), (mi * P + W) // Q + 1)
        res[k * Q - mi * P + W] += y[k]
    Hm = np.abs(np.fft.rfft(res))
    Hm /= Hm[0]
    f = np.arange(len(Hm)) * R / len(res)
    assert 0 < st.passband and st.stopband <= min(fi, fo) / 2
    stop_db = 20 * np.log10(Hm[f >= min(fi, fo) / 2].max() + 1e-300)
    flat = np.abs(Hm[f <= st.passband] - 1).max()
    say(f"{fi} -> {fo} {mode}: stopband {stop_db:.1f} dB, passband flat {flat:.1e}")
    assert stop_db < -280.0
    assert flat < 1e-12


def run_cli(*args, check=True):
    r = subprocess.run([H.GPU_CLI, *map(str, args)], capture_output=True, text=True, timeout=120)
    if check:
        assert r.returncode == 0, r.stderr
    return r


def read_wav(path):
    """(samples (frames, ch) float64, rate, bits) of a PCM or float WAV."""
    blob = open(path, "rb").read()
    pos, fmt = 12, None
    while pos + 8 <= len(blob):
        tag, size = blob[pos:pos + 4], struct.unpack("<I", blob[pos + 4:pos + 8])[0]
        if tag == b"fmt ":
            fmt = struct.unpack("<HHIIHH", blob[pos + 8:pos + 24])
            if fmt[0] == 0xFFFE:
                fmt = (struct.unpack("<H", blob[pos + 32:pos + 34])[0],) + fmt[1:]
        if tag == b"data":
            data = blob[pos + 8:pos + 8 + size]
            break
        pos += 8 + size + (size & 1)
    tagv, ch, rate, _, align, bits = fmt
    if tagv == 3:
        v = np.frombuffer(data, dtype="<f8" if bits == 64 else "<f4").astype(np.float64)
    else:
        B = bits // 8
        raw = np.frombuffer(data, dtype=np.uint8).reshape(-1, B).astype(np.int64)
        iv = sum(raw[:, i] << (8 * i) for i in range(B))
        iv = np.where(iv >= 1 << (8 * B - 1), iv - (1 << (8 * B)), iv)
        v = iv / float(1 << (8 * B - 1))
    return v.reshape(-1, ch), rate, bits


def test_cli_formats_and_edge_cases(tmp_path):
    t = np.arange(48000) / 48000
    s = 0.5 * np.sin(2 * np.pi * 440 * t)
    # PCM input at the identity rate: the output is the decoded input
    for width in [1, 2, 3, 4]:
        src = tmp_path / "pcm.wav"
        with wave.open(str(src), "wb") as w:
            w.setnchannels(2), w.setsampwidth(width), w.setframerate(48000)
            if width == 1:
                d = np.round(s * 127 + 128).astype(np.uint8).reshape(-1, 1)
            else:
                v = np.round(s * (2 ** (8 * width - 1) - 1)).astype(np.int64)
                d = np.stack([(v >> (8 * i)) & 255 for i in range(width)], 1).astype(np.uint8)
            w.writeframes(np.stack([d, d], 1).tobytes())
        run_cli(src, tmp_path / "out.wav", "--rate", 48000)
        y = read_wav(tmp_path / "out.wav")[0]
        assert np.abs(y[:, 0] - s).max() <= 1.5 * 2.0 ** -(8 * width - 1) and np.array_equal(y[:, 0], y[:, 1])
    # output formats against the f64 output
    src = tmp_path / "in.wav"
    src.write_bytes(H.wav_f64(s, 1, 48000))
    run_cli(src, tmp_path / "ref.wav", "--rate", 44100)
    ref = read_wav(tmp_path / "ref.wav")[0]
    for fmt, bits, tol in [("f32", 32, 2 ** -24), ("s32", 32, 2 ** -31), ("s24", 24, 2 ** -23), ("s16", 16, 2 ** -15)]:
        run_cli(src, tmp_path / "out.wav", "--rate", 44100, "--format", fmt)
        y, sr, b = read_wav(tmp_path / "out.wav")
        assert b == bits and sr == 44100 and np.abs(y - ref).max() <= tol, fmt
    # six channels, tiny and empty inputs
    src.write_bytes(H.wav_f64(np.stack([s * (i + 1) / 7 for i in range(6)], 1), 6, 48000))
    run_cli(src, tmp_path / "out.wav", "--rate", 96000, "--mode", "fast")
    assert read_wav(tmp_path / "out.wav")[0].shape == (96000, 6)
    for n in [1, 10]:
        src.write_bytes(H.wav_f64(s[:n], 1, 48000))
        run_cli(src, tmp_path / "out.wav", "--rate", 44100)
        assert read_wav(tmp_path / "out.wav")[0].shape == (-(-n * 44100 // 48000), 1)
    hdr = b"WAVEfmt " + struct.pack("<IHHIIHH", 16, 1, 1, 48000, 96000, 2, 16) + b"data" + struct.pack("<I", 0)
    (tmp_path / "empty.wav").write_bytes(b"RIFF" + struct.pack("<I", len(hdr)) + hdr)
    run_cli(tmp_path / "empty.wav", tmp_path / "out.wav", "--rate", 44100)
    assert read_wav(tmp_path / "out.wav")[0].shape == (0, 1)
    # errors: usage exits 2, bad files exit 1, always with a "kbsrc: " message
    (tmp_path / "trunc.wav").write_bytes((tmp_path / "ref.wav").read_bytes()[:30])
    ref_wav, o = tmp_path / "ref.wav", tmp_path / "o.wav"
    for argv, code in [([ref_wav, o, "--rate", "0"], 2), ([ref_wav, o, "--rate", "abc"], 2),
                       ([ref_wav, o, "--mode", "nope"], 2), ([ref_wav, o, "--len", "-1"], 1), ([ref_wav], 2),
                       ([tmp_path / "missing.wav", o], 1), ([tmp_path / "trunc.wav", o], 1)]:
        r = run_cli(*argv, check=False)
        assert r.returncode == code and r.stderr.startswith("kbsrc: "), (argv, r.returncode, r.stderr)
