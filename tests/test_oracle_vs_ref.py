"""The oracle against the compiled reference (oracle/_ref, built from the reference's sources by oracle/ref.mk):
bit-identical output over rate pairs, modes, block sizes, edge modes and lengths. Skipped where the reference was
not built (it is never shipped; the committed pins in test_oracle.py cover that case)."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.skipif(not H.ref_available(), reason="oracle/_ref not built (needs the reference sources)")

CASES = [
    # rin, rout, ch, frames, mode, length, block, lpc
    (48000, 96000, 2, 3000, "exact", 0.2, 0, True),
    (48000, 44100, 1, 20000, "exact", 0.2, 0, True),
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
    (48000, 44100, 6, 2000, "fast", 0.01, 0, True),
    (16000, 48000, 1, 3000, "exact", 0.01, 0, True),
    (96000, 32001, 1, 5000, "exact", 0.02, 0, True),    # Bluestein lengths
    (44100, 44101, 1, 3000, "exact", 0.02, 0, True),    # near-unity: general fold
    (48000, 48001, 1, 3000, "fast", 0.02, 0, False),
    (22050, 48001, 1, 3000, "exact", 0.02, 0, True),
    (1001, 1000, 1, 4000, "exact", 0.2, 0, True),
    (44100, 44200, 2, 3000, "fast", 0.02, 0, True),
    (48000, 44101, 1, 3000, "exact", 0.02, 0, True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}-{c[2]}ch-{c[3]}-{c[4]}-L{c[5]}-b{c[6]}")
def test_oracle_equals_reference(case, tmp_path):
    rin, rout, ch, frames, mode, length, block, lpc = case
    x = np.random.default_rng(frames + rin + rout).uniform(-0.5, 0.5, frames * ch)
    got = H.oracle_resample(x, ch, rin, rout, mode, length=length, block=block, lpc=lpc)
    want = H.ref_resample(str(tmp_path), x, ch, rin, rout, mode, length=length, block=block, lpc=lpc)
    assert got.shape == want.shape
    assert got.tobytes() == want.tobytes()
