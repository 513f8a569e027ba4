"""The host copy crew that stages host buffers for the device (kbsrc_amd/csrc/copy_crew.hpp): concurrent copies of
random sizes equal their sources, in a plain build and under ThreadSanitizer (no data race reported). CPU only."""
import os
import subprocess

import pytest

import helpers as H

SRC = os.path.join(H.ROOT, "tests", "copy_crew_stress.cpp")
INC = os.path.join(H.ROOT, "kbsrc_amd", "csrc")


@pytest.mark.parametrize("flags", [["-O2"], ["-O1", "-g", "-fsanitize=thread"]], ids=["plain", "tsan"])
def test_copies_equal_their_sources(flags, tmp_path):
    exe = tmp_path / "crew"
    build = subprocess.run(["g++", "-std=c++17", *flags, "-pthread", "-I", INC, SRC, "-o", str(exe)],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([str(exe), "12"], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr[-3000:])
    assert "mismatches 0" in r.stdout
