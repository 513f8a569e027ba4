"""The odd-radix butterflies start their sine sums as zero plus the first product in two operations
(zero_add_product in kbsrc_amd/csrc/dd.hpp) where the reference adds the product to a dd zero: on the host, over
millions of random operands with subnormals, signed zeros and cancellations, both dd adds of zero and a product
give exactly that (tests/dd_zero_add.cpp). CPU only."""
import os
import subprocess

import helpers as H

SRC = os.path.join(H.ROOT, "tests", "dd_zero_add.cpp")
INC = os.path.join(H.ROOT, "kbsrc_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def test_zero_plus_product_equals_both_dd_adds(tmp_path):
    exe = tmp_path / "dd_zero_add"
    build = subprocess.run([os.path.join(ROCM, "bin", "hipcc"), "-std=c++20", "-O2", "-ffp-contract=off", "-mfma",
                            "-D__HIP_PLATFORM_AMD__", "-isystem", os.path.join(ROCM, "include"), "-I", INC, "-x", "c++",
                            SRC, "-o", str(exe)], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-3000:]
    r = subprocess.run([str(exe), "3000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert int(r.stdout.split()[-1]) > 5_000_000
