"""Seeded random conversions against the oracle, bit for bit: rate pairs drawn from the common rates, multiples of
100 Hz, near-unity pairs and rates with a prime factor above 7 (so the down fold, the up extension, the general
fold and Bluestein transforms all occur), 1-4 channels, 0-6000 frames, kernel lengths 0.005-0.1 s, automatic or explicit blocks, LPC on and
off, both modes. A plan the oracle rejects must be rejected by the GPU library with the same message. Needs an
MI355X."""
import math
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

SMOOTH = [8000, 11025, 16000, 22050, 24000, 32000, 44100, 48000, 64000, 88200, 96000, 176400, 192000]


def draw(seed):
    """One case; rates are drawn so that the plans' transform lengths (Nb = nb Q) stay small enough for the oracle
    to finish in seconds."""
    rng = np.random.default_rng(seed)
    kind = seed % 4
    if kind == 0:      # common rates
        rin, rout = (int(v) for v in rng.choice(SMOOTH, 2, replace=False))
    elif kind == 1:    # multiples of 100 Hz, 4-200 kHz
        rin, rout = (100 * int(v) for v in rng.integers(40, 2001, 2))
    elif kind == 2:    # near unity below 20 kHz: the general fold
        rin = int(rng.integers(4000, 20000))
        rout = rin + int(rng.choice([-3, -2, -1, 1, 2, 3]))
    else:              # a common rate against 1 kHz times a prime above 7: Bluestein transforms
        rin = int(rng.choice(SMOOTH))
        rout = 1000 * int(rng.choice([11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79]))
        if rng.random() < 0.5:
            rin, rout = rout, rin
    ch = int(rng.integers(1, 5))
    frames = int(rng.choice([0, 1, 7, int(rng.integers(50, 6000))]))
    length = float(np.round(rng.uniform(0.005, 0.05 if kind == 2 else 0.1), 4))
    mode = "exact" if rng.random() < 0.6 else "fast"
    lpc = bool(rng.random() < 0.7)
    block = 0 if rng.random() < 0.7 else int(rng.integers(50, 4000))
    g = math.gcd(rin, rout)
    if block * max(rin, rout) // g > 400_000:   # explicit blocks only where the transforms stay small
        block = 0
    return rin, rout, ch, frames, mode, length, block, lpc


@pytest.mark.parametrize("seed", range(40))
def test_random_conversion_equals_oracle(seed):
    import kbsrc_amd
    rin, rout, ch, frames, mode, length, block, lpc = draw(seed)
    print(f"seed {seed}: {rin} -> {rout}, {ch} ch, {frames} frames, {mode}, len {length}, block {block}, lpc {lpc}",
          flush=True)
    sys.stdout.flush()
    x = np.random.default_rng(1000 + seed).uniform(-0.5, 0.5, frames * ch)
    try:
        want = H.oracle_resample(x, ch, rin, rout, mode, length=length, block=block, lpc=lpc)
    except ValueError as e:
        with pytest.raises(ValueError, match=str(e).replace("(", r"\(").replace(")", r"\)")):
            kbsrc_amd.resample(x.reshape(frames, ch), rin, rout, mode, length=length, block=block, lpc_edges=lpc)
        return
    got = kbsrc_amd.resample(x.reshape(frames, ch), rin, rout, mode, length=length, block=block, lpc_edges=lpc)
    g = math.gcd(rin, rout)
    assert got.shape == (ch, -(-frames * (rout // g) // (rin // g)))
    assert got.tobytes() == want.tobytes()
