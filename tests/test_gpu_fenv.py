"""Exact mode does not depend on the caller's rounding mode, and the caller's mode survives every call (also a
rejected one, and fast mode): the reference's tests/test_engines.py rounding-mode checks, through the C++ API of
the GPU library and a driver of our own (tests/fenv_driver.cpp, built as kbsrc_amd/bin/kbsrc_fenv_driver). Needs an
MI355X."""
import os
import struct
import subprocess

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(H.ROOT, "kbsrc_amd", "bin", "kbsrc_fenv_driver")
MODES = ["nearest", "down", "up", "zero"]


def drive(tmp_path, mode, x, rate_in, rate_out, prec="exact", length="0.1"):
    src, dst = tmp_path / "in.f64", tmp_path / "out.f64"
    src.write_bytes(np.asarray(x, dtype="<f8").tobytes())
    if dst.exists():
        dst.unlink()
    r = subprocess.run([DRIVER, f"round={mode}", f"precision={prec}", f"from={rate_in}", f"to={rate_out}", "channels=2",
                        f"len={length}", "beta=40", f"src={src}", f"dst={dst}"], capture_output=True, text=True,
                       timeout=120)
    info = dict(kv.split("=", 1) for kv in r.stdout.split() if "=" in kv)
    data = dst.read_bytes() if dst.exists() else b""
    return r.returncode, info, data


@pytest.mark.parametrize("mode", MODES)
def test_exact_is_independent_of_the_callers_mode(mode, tmp_path):
    # one stereo frame 8 kHz -> 16 kHz, sensitive to the rounding mode: the pinned samples in every mode
    rc, info, data = drive(tmp_path, mode, H.FRAME1, 8000, 16000, length="0.01")
    assert rc == 0 and info.get("restored") == "1"
    got = [struct.pack(">d", v).hex() for v in struct.unpack("<%dd" % (len(data) // 8), data)]
    assert got == H.FRAME1_OUT
    # subnormal input 48k -> 96k: the pinned round-to-nearest output
    rc, info, data = drive(tmp_path, mode, H.subnormal97(), 48000, 96000)
    assert rc == 0 and info.get("restored") == "1" and H.sha(data) == H.SUBNORMAL_96K_PLANAR
    # normal input, up- and downsampling: the round-to-nearest output
    for rate_out in (96000, 44100):
        rc0, _, ref = drive(tmp_path, "nearest", H.primary97(), 48000, rate_out)
        rc, info, data = drive(tmp_path, mode, H.primary97(), 48000, rate_out)
        assert rc0 == 0 and rc == 0 and info.get("restored") == "1" and data == ref


def test_callers_mode_survives_rejection_and_fast(tmp_path):
    rc, info, _ = drive(tmp_path, "down", H.FRAME1, 0, 16000)
    assert rc == 3 and info.get("rejected") == "1" and info.get("restored") == "1"
    rc, info, _ = drive(tmp_path, "up", H.FRAME1, 8000, 16000, prec="fast", length="0.01")
    assert rc == 0 and info.get("restored") == "1"
