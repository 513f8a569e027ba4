"""Rejected options and malformed WAV headers fail cleanly through the GPU command line, and their valid neighbours
convert: the checks of the reference's tests/test_robust.py, written here from them (exact and fast modes; the
x87 long-double mode is not built). A rejection exits 1 with "kbsrc: <reason>" and writes no file. Needs an
MI355X (the accepted neighbours run on it)."""
import math
import os
import random
import re
import struct
import subprocess

import pytest

import helpers as H

pytestmark = pytest.mark.gpu

OUT_OF_RANGE = "option out of range"
TOO_LARGE = "block, rates or length too large"
TOO_SHORT = "length too short for these sample rates"


def wav(tag, ch, rate, bits, align, data):
    """RIFF/WAVE bytes with a 16-byte fmt chunk, header fields written as given."""
    fmt = struct.pack("<HHIIHH", tag, ch, rate, rate * align & 0xFFFFFFFF, align, bits)
    body = b"WAVE" + b"fmt " + struct.pack("<I", len(fmt)) + fmt + b"data" + struct.pack("<I", len(data)) + data
    return b"RIFF" + struct.pack("<I", len(body)) + body


def pcm(width, values):
    if width == 1:
        return bytes(v + 128 for v in values)
    return b"".join((v & (1 << 8 * width) - 1).to_bytes(width, "little") for v in values)


def f64_samples(path):
    blob = open(path, "rb").read()
    pos = 12
    while pos + 8 <= len(blob):
        tag, size = blob[pos:pos + 4], struct.unpack("<I", blob[pos + 4:pos + 8])[0]
        if tag == b"data":
            data = blob[pos + 8:pos + 8 + size]
            return struct.unpack("<%dd" % (len(data) // 8), data)
        pos += 8 + size + (size & 1)
    return None


class Cli:
    def __init__(self, tmp):
        self.tmp = tmp
        self.out = os.path.join(tmp, "out.wav")

    def write(self, name, blob):
        path = os.path.join(self.tmp, name + ".wav")
        with open(path, "wb") as f:
            f.write(blob)
        return path

    def run(self, src, *opts):
        if os.path.exists(self.out):
            os.remove(self.out)
        r = subprocess.run([H.GPU_CLI, *map(str, opts), src, self.out], capture_output=True, text=True, timeout=120)
        return r.returncode, r.stderr

    def rejected(self, reason, src, *opts):
        rc, err = self.run(src, *opts)
        assert rc == 1 and err.startswith("kbsrc: ") and reason in err and not os.path.exists(self.out), \
            (opts, rc, err)

    def converts(self, src, *opts, plan=()):
        rc, err = self.run(src, "-v", *opts)
        assert rc == 0, (opts, err)
        y = f64_samples(self.out)
        assert y is not None and len(y) > 0 and all(map(math.isfinite, y)), opts
        for p in plan:
            assert re.search(r"(^|\s)" + re.escape(p) + r"(\s|$)", err), (opts, p, err)
        return y


@pytest.fixture
def cli(tmp_path):
    return Cli(str(tmp_path))


@pytest.fixture
def src(cli):
    rng = random.Random(0x524F42)
    x = [rng.uniform(-0.5, 0.5) for _ in range(2 * 64)]
    return cli.write("f64_48k", wav(3, 2, 48000, 64, 16, struct.pack("<%dd" % len(x), *x)))


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_block_limits(cli, src, mode):
    """At 48 kHz -> 96 kHz, --len 0.1 the kernel margin is 4802 frames: blocks up to 9604 leave no hop; 9605 is
    accepted and rounded up to the 7-smooth 9720."""
    for block in (1, 1000, 9604):
        cli.rejected(OUT_OF_RANGE, src, "--rate", 96000, "--mode", mode, "--len", "0.1", "--block", block)
    cli.converts(src, "--rate", 96000, "--mode", mode, "--len", "0.1", "--block", 9605, plan=["block=9720"])


def test_block_sizes_and_overflow(cli, src):
    cli.converts(src, "--rate", 96000)
    cli.converts(src, "--rate", 96000, "--len", "0.1", "--block", 20000, plan=["block=20000"])
    for block in (536870912, 2147483647):   # transform sizes block x P, block x Q would overflow int
        cli.rejected(TOO_LARGE, src, "--rate", 96000, "--block", block)


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_kernel_length_limits(cli, src, mode):
    cli.rejected(OUT_OF_RANGE, src, "--rate", 96000, "--mode", mode, "--len", "1e-5")      # < 1 kernel sample
    cli.rejected(TOO_SHORT, src, "--rate", 96000, "--mode", mode, "--len", "5.3e-4")      # no room for a passband
    cli.converts(src, "--rate", 96000, "--mode", mode, "--len", "5.33e-4", plan=["taps=103"])


def test_kernel_length_edges(cli, src):
    for length in ("1e-8", "0", "-0.1", "10.5"):
        cli.rejected(OUT_OF_RANGE, src, "--rate", 96000, "--len", length)
    cli.rejected(OUT_OF_RANGE, src, "--rate", 44100, "--len", "2e-5")
    cli.rejected(TOO_SHORT, src, "--rate", 44100, "--len", "5.79e-4")
    cli.converts(src, "--rate", 44100, "--len", "5.8e-4", plan=["taps=55"])
    cli.rejected(TOO_SHORT, src, "--rate", 44100, "--len", "5.8e-4", "--beta", 80)
    cli.converts(src, "--rate", 16000, "--len", 10, plan=["taps=960001"])


def test_input_rate_limit_and_equal_rates(cli, src):
    rng = random.Random(1)
    x = [rng.uniform(-0.5, 0.5) for _ in range(2 * 64)]
    data = struct.pack("<%dd" % len(x), *x)
    for rate in (10000001, 2000000000, 2147483647):
        cli.rejected("input sample rate above 10000000 Hz", cli.write("hi", wav(3, 2, rate, 64, 16, data)),
                     "--rate", 48000)
    cli.converts(cli.write("top", wav(3, 2, 10000000, 64, 16, data)), "--rate", 48000, "--len", "0.001",
                 "--mode", "fast")
    cli.converts(src, "--rate", 48000, plan=["P/Q=1/1", "threads=1"])


def test_wav_header_rejections(cli):
    for tag, bits, align, reason in ((1, 0, 2, "unsupported PCM width"), (1, 8, 2, "unsupported PCM width"),
                                     (1, 17, 2, "unsupported PCM width"), (1, 33, 4, "unsupported PCM width"),
                                     (1, 40, 5, "unsupported PCM width"),
                                     (3, 64, 4, "IEEE float sample size does not match the block alignment"),
                                     (3, 32, 8, "IEEE float sample size does not match the block alignment"),
                                     (3, 32, 2, "IEEE float sample size does not match the block alignment"),
                                     (3, 16, 2, "unsupported WAV format")):
        cli.rejected(reason, cli.write("bad", wav(tag, 1, 48000, bits, align, bytes(align * 4096))), "--rate", 96000)
    f64 = struct.pack("<4d", 0.1, 0.2, 0.3, 0.4)
    good = wav(3, 2, 48000, 64, 16, f64)
    for blob, reason in ((wav(3, 0, 48000, 64, 16, f64), "invalid WAV channel count"),
                         (wav(3, 2, 0, 64, 16, f64), "invalid WAV sample rate"),
                         (wav(3, 2, 1 << 31, 64, 16, f64), "invalid WAV sample rate"),
                         (wav(3, 2, (1 << 32) - 1, 64, 16, f64), "invalid WAV sample rate"),
                         (good[:12] + good[36:], "missing fmt/data chunk")):
        cli.rejected(reason, cli.write("bad", blob), "--rate", 96000)


def test_output_header_limits(cli):
    wide = cli.write("wide", wav(1, 8192, 192000, 16, 2 * 8192, bytes(2 * 8192 * 4)))
    cli.rejected("too many channels for a WAV file in this format", wide, "--rate", 192000, "--format", "f64")
    cli.rejected("sample rate too high for a WAV file with this many channels", wide, "--rate", 192000,
                 "--format", "f32")
    rc, err = cli.run(wide, "--rate", 192000, "--format", "s16")
    assert rc == 0, err
    with open(cli.out, "rb") as f:
        assert struct.unpack("<HIIH", f.read(34)[22:34]) == (8192, 192000, 192000 * 16384, 16384)


def test_pcm_containers(cli):
    rng = random.Random(0x524F42)
    ints = [rng.randint(-(1 << 31), (1 << 31) - 1) for _ in range(2 * 64)]
    x = [rng.uniform(-0.5, 0.5) for _ in range(2 * 64)]
    for tag, bits, width in ((1, 8, 1), (1, 16, 2), (1, 24, 3), (1, 32, 4), (3, 32, 4), (3, 64, 8)):
        if tag == 3:
            data = struct.pack("<%d%s" % (len(x), "f" if width == 4 else "d"), *x)
        else:
            data = pcm(width, [v >> 32 - bits for v in ints])
        cli.converts(cli.write("c", wav(tag, 2, 48000, bits, 2 * width, data)), "--rate", 96000)
    # fewer significant bits than the container decode like the full width (low bits zero)
    for bits, width in ((12, 2), (20, 3)):
        data = pcm(width, [(v >> 32 - bits) << 8 * width - bits for v in ints])
        y_full = cli.converts(cli.write("full", wav(1, 2, 48000, 8 * width, 2 * width, data)), "--rate", 96000)
        y_part = cli.converts(cli.write("part", wav(1, 2, 48000, bits, 2 * width, data)), "--rate", 96000)
        assert y_part == y_full
