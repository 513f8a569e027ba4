"""The command line's WAV reading and writing, and its argument handling, against the compiled reference (oracle/_ref,
built from the reference's sources): equal-rate conversions take no GPU work, so every output format (with and
without dither, clipping and NaN samples, plain and WAVE_FORMAT_EXTENSIBLE headers) and every input encoding (PCM 8,
16, 24, 32 bits and IEEE float) must give the reference's file byte for byte, and malformed headers and bad arguments
the reference's message and exit status (src/io/wav.cpp:109-253, app/main.cpp:84-165). CPU only."""
import os
import struct
import subprocess

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.skipif(not H.ref_available(), reason="oracle/_ref (the compiled reference) is not built")


def run(binary, *args):
    r = subprocess.run([binary, *map(str, args)], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout, r.stderr


def both(tmp_path, *args, out_name="o.wav"):
    """(status, stderr, file bytes) of this CLI and of the reference on the same arguments."""
    got = []
    for k, binary in enumerate((H.GPU_CLI, H.REF_BIN)):
        out = tmp_path / f"{k}_{out_name}"
        if out.exists():
            out.unlink()
        rc, _, err = run(binary, *[str(out) if a == "OUT" else a for a in args])
        got.append((rc, err.replace(H.REF_BIN, "").strip(), out.read_bytes() if out.exists() else None))
    return got


def f64_input(tmp_path, ch=3, frames=5001, seed=3):
    x = np.random.default_rng(seed).uniform(-1.2, 1.2, (frames, ch))   # some samples clip
    x[7, ch - 1] = np.nan
    path = tmp_path / "in.wav"
    path.write_bytes(H.wav_f64(x, ch, 44100))
    return path


@pytest.mark.parametrize("ch", [1, 2, 3])
@pytest.mark.parametrize("fmt", ["f64", "f32", "s32", "s24", "s16"])
@pytest.mark.parametrize("dither", [False, True])
def test_written_files_equal_the_reference(tmp_path, ch, fmt, dither):
    src = f64_input(tmp_path, ch)
    ours, ref = both(tmp_path, src, "OUT", "--rate", 44100, "-f", fmt, *(["--dither"] if dither else []))
    assert ours[0] == ref[0] == 0
    assert ours[1] == ref[1]   # the clipping warning
    assert ours[2] == ref[2]


@pytest.mark.parametrize("fmt", ["s16", "s24", "s32", "f32"])
def test_read_encodings_equal_the_reference(tmp_path, fmt):
    src = f64_input(tmp_path, 2)
    enc = tmp_path / f"enc_{fmt}.wav"
    assert run(H.REF_BIN, src, enc, "--rate", 44100, "-f", fmt)[0] == 0
    ours, ref = both(tmp_path, enc, "OUT", "--rate", 44100)
    assert ours[0] == ref[0] == 0 and ours[2] == ref[2]


def pcm8(tmp_path, fmt_len=16, tag=1, align=2, bits=8, ch=2, rate=8000, data=None, riff=b"RIFF", extra=b""):
    d = np.random.default_rng(1).integers(0, 256, 2000).astype(np.uint8).tobytes() if data is None else data
    fmt = struct.pack("<HHIIHH", tag, ch, rate, rate * align, align, bits) + b"\0" * (fmt_len - 16)
    body = b"WAVE" + extra + b"fmt " + struct.pack("<I", fmt_len) + fmt + b"data" + struct.pack("<I", len(d)) + d
    path = tmp_path / "pcm.wav"
    path.write_bytes(riff + struct.pack("<I", len(body)) + body)
    return path


def test_pcm8_and_odd_chunks(tmp_path):
    src = pcm8(tmp_path, extra=b"LIST" + struct.pack("<I", 3) + b"abc\0")   # odd chunk, padded
    for fmt in ("f64", "s16", "s24"):
        ours, ref = both(tmp_path, src, "OUT", "--rate", 8000, "-f", fmt)
        assert ours[0] == ref[0] == 0 and ours[2] == ref[2]


@pytest.mark.parametrize("case", [
    dict(riff=b"RIFX"), dict(fmt_len=14), dict(ch=0), dict(rate=0), dict(align=3), dict(tag=2),
    dict(tag=3, bits=32, align=4), dict(bits=20, align=2), dict(bits=0), dict(align=10, bits=40),
])
def test_malformed_headers_give_the_reference_message(tmp_path, case):
    src = pcm8(tmp_path, **case)
    ours, ref = both(tmp_path, src, "OUT", "--rate", 8000)
    assert ours[0] == ref[0] == 1
    assert ours[1] == ref[1]
    assert ours[2] is None and ref[2] is None


@pytest.mark.parametrize("argv", [
    ["a.wav", "b.wav", "--rate", "0"], ["a.wav", "b.wav", "--rate", "abc"], ["a.wav", "b.wav", "--rate", "12x"],
    ["a.wav", "b.wav", "--rate", "99999999999"], ["a.wav", "b.wav", "--len", "1e400"], ["a.wav", "b.wav", "--len", "nan"],
    ["a.wav", "b.wav", "--beta"], ["a.wav", "b.wav", "--mode", "nope"], ["a.wav"], ["--bogus"], ["a", "b", "c"],
    ["a.wav", "b.wav", "-f", "s8"], ["--help", "--bogus"], ["--version"], ["/nonexistent.wav", "o.wav"],
])
def test_arguments_are_handled_as_the_reference_does(tmp_path, argv):
    rc, out, err = run(H.GPU_CLI, *argv)
    rrc, rout, rerr = run(H.REF_BIN, *argv)
    assert rc == rrc, (argv, err, rerr)
    if rc:
        assert err == rerr.replace(H.REF_BIN, "")
    elif argv[0] == "--version":
        assert out.split()[0] == rout.split()[0] == "kbsrc"
    else:
        assert out.startswith("kbsrc ") and "usage: kbsrc [options] input.wav output.wav" in out
