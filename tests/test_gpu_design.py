"""The kernel design on the GPU: Kaiser-sinc taps (src/design.cpp:17-37) and the kernel spectrum
(src/design.cpp:40-49), bit-identical to the oracle through the C ABI. Needs an MI355X."""
import sys
import time

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


TAP_CASES = [
    # L, a, b, beta
    (1, 11009, 24000, 40.0),
    (2, 1, 2, 40.0),
    (9600, 11009, 12000, 40.0),        # 48 kHz -> 44.1 kHz, len 0.2
    (19200, 11009, 24000, 40.0),       # 96 kHz -> 44.1 kHz, len 0.2
    (8820, 2, 3, 5.0),
    (4000, 999999937, 1000000007, 200.0),   # n a mod 2b beyond 64 bits
    (30000, 1, 1000, 0.5),
]


@pytest.mark.parametrize("case", TAP_CASES, ids=lambda c: f"L{c[0]}-{c[1]}/{c[2]}-b{c[3]}")
def test_taps_equal_oracle(case):
    L, a, b, beta = case
    say(f"taps {case}")
    got = gpu().kernel_taps(L, a, b, beta)
    want = H.oracle_taps(L, a, b, beta)
    assert got.tobytes() == want.tobytes()


def test_taps_extreme_ratio_sampled():
    """10 MHz -> 48 kHz at --len 0.5: 5,000,001 taps on the GPU against the oracle at 4000 spread indices."""
    L, a, b, beta = 5000000, 23987, 5000000, 40.0
    say("taps 10M->48k len 0.5")
    t0 = time.perf_counter()
    got = gpu().kernel_taps(L, a, b, beta)
    say(f"  {L + 1} taps in {time.perf_counter() - t0:.3f} s")
    rng = np.random.default_rng(7)
    idx = np.unique(np.concatenate([np.arange(0, 600), np.arange(L - 600, L + 1), rng.integers(0, L + 1, 2800)]))
    want = H.oracle_taps(L, a, b, beta, idx)
    assert got[idx].tobytes() == want.tobytes()


SPEC_CASES = [
    # Lk, n (even), scale: four-step and single-pass lengths, and halves with a prime factor above 7 (Bluestein)
    (9600, 40000, 40000.0),
    (19200, 201600, 201600.0),
    (300, 2 * 1009, 2018.0),
    (500, 2 * 3 * 11 * 13, 858.0),
    (100, 256, 256.0),
    (1, 64, 64.0),
]


@pytest.mark.parametrize("case", SPEC_CASES, ids=lambda c: f"Lk{c[0]}-n{c[1]}")
def test_spectrum_equals_oracle(case):
    Lk, n, scale = case
    say(f"spectrum {case}")
    taps = H.oracle_taps(Lk, 11009, 24000, 40.0)
    got = gpu().kernel_spectrum(taps, n, scale)
    want = H.oracle_spectrum(taps, n, scale)
    assert got.tobytes() == want.tobytes()


def test_design_rejections():
    with pytest.raises(ValueError):
        gpu().kernel_taps(10, 0, 2, 40.0)
    with pytest.raises(ValueError):
        gpu().kernel_spectrum(np.zeros((11, 2)), 21, 1.0)   # odd length
    with pytest.raises(ValueError):
        gpu().kernel_spectrum(np.zeros((11, 2)), 10, 1.0)   # kernel longer than the transform
