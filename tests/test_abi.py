"""The C-ABI library: it loads, exports every entry point include/kbsrc_gpu.h declares, and rejects bad arguments
with the reference's messages before touching a device. CPU only."""
import ctypes
import os
import re
import struct
import subprocess

import pytest

import helpers as H

HEADER = os.path.join(H.ROOT, "include", "kbsrc_gpu.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kbsrc_\w+)\s*\(", src)))


def test_library_exports_every_declared_entry_point():
    import kbsrc_amd
    lib = ctypes.CDLL(kbsrc_amd.library_path())
    names = declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", kbsrc_amd.library_path()], capture_output=True, text=True)
    exported = set(re.findall(r" T (kbsrc_\w+)", out.stdout))
    assert set(names) <= exported


def test_version_and_output_frames():
    import kbsrc_amd
    assert kbsrc_amd.version().startswith("1.3.0")
    assert kbsrc_amd.output_frames(17_280_000, 96000, 44100) == 7_938_000
    assert kbsrc_amd.output_frames(1, 48000, 44100) == 1
    assert kbsrc_amd.output_frames(10, 48000, 44100) == 10
    assert kbsrc_amd.output_frames(0, 48000, 44100) == 0
    with pytest.raises(ValueError, match="sample rates must be positive"):
        kbsrc_amd.output_frames(10, 0, 44100)


@pytest.mark.parametrize("kw,msg", [
    (dict(rate_in=48000, rate_out=96000, length=0.1, block=9604), "option out of range"),
    (dict(rate_in=48000, rate_out=96000, length=1e-5), "option out of range"),
    (dict(rate_in=48000, rate_out=96000, length=5.3e-4), "length too short for these sample rates"),
    (dict(rate_in=48000, rate_out=44100, length=5.79e-4), "length too short for these sample rates"),
    (dict(rate_in=48000, rate_out=96000, block=536870912), "block, rates or length too large"),
    (dict(rate_in=48000, rate_out=96000, length=10.5), "option out of range"),
    (dict(rate_in=48000, rate_out=96000, beta=0.0), "option out of range"),
    (dict(rate_in=48000, rate_out=96000, precision="ld"), "option out of range"),
    (dict(rate_in=0, rate_out=96000), "sample rates must be positive"),
])
def test_plan_rejections_before_device(kw, msg):
    import kbsrc_amd
    kw = dict(kw)
    with pytest.raises(ValueError, match=msg):
        kbsrc_amd.Plan(64, 2, kw.pop("rate_in"), kw.pop("rate_out"), **kw)


def test_cli_usage_errors():
    """Usage errors exit 2 with "kbsrc: ..."; a missing input exits 1 (app/main.cpp:152-165)."""
    for argv, code in [(["a.wav", "b.wav", "--rate", "0"], 2), (["a.wav", "b.wav", "--rate", "abc"], 2),
                       (["a.wav", "b.wav", "--mode", "nope"], 2), (["a.wav"], 2), (["--bogus"], 2),
                       (["/nonexistent.wav", "/tmp/o.wav"], 1)]:
        r = subprocess.run([H.GPU_CLI, *argv], capture_output=True, text=True)
        assert r.returncode == code and r.stderr.startswith("kbsrc: "), (argv, r.returncode, r.stderr)
    r = subprocess.run([H.GPU_CLI, "--version"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("kbsrc 1.3.0")


@pytest.mark.parametrize("mode", ["nearest", "down", "up", "zero"])
def test_rounding_driver_sees_a_changed_mode(mode, tmp_path):
    """The rounding-mode driver of the GPU tests (tests/fenv_driver.cpp) on calls that need no GPU (equal rates, a
    rejected rate): it reports the caller's mode as restored, and not restored when the SSE rounding field alone is
    changed behind the library's back."""
    driver = os.path.join(H.ROOT, "kbsrc_amd", "bin", "kbsrc_fenv_driver")
    src, dst = tmp_path / "in.f64", tmp_path / "out.f64"
    src.write_bytes(struct.pack("<4d", 0.1, 0.2, 0.3, 0.4))
    for rate_in, code in ((8000, 0), (0, 3)):
        for tamper, restored in (("0", "1"), ("1", "0")):
            r = subprocess.run([driver, f"round={mode}", "precision=exact", f"from={rate_in}", "to=8000", "channels=2",
                                "len=0.01", "beta=40", f"src={src}", f"dst={dst}", f"tamper={tamper}"],
                               capture_output=True, text=True, timeout=60)
            info = dict(kv.split("=", 1) for kv in r.stdout.split() if "=" in kv)
            assert r.returncode == code and info.get("restored") == restored, (rate_in, tamper, r.stdout)
