"""Every conversion between the common audio rates (8 kHz to 192 kHz), both modes: the specialised pass kernels
(measured and generated tables) give the same bits as the generic runtime-radix kernel, whose arithmetic the
oracle tests pin on many shapes. Inputs of 8 channels x 3 s and 2 channels x 20 s reach the block sizes of long and
short files. Needs an MI355X."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RATES = [8000, 11025, 16000, 22050, 24000, 32000, 44100, 48000, 88200, 96000, 176400, 192000]


def gpu():
    import kbsrc_amd
    return kbsrc_amd


@pytest.mark.parametrize("ch,secs", [(8, 3), (2, 20)], ids=["8ch-3s", "2ch-20s"])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_specialised_equals_generic_over_rate_pairs(mode, ch, secs, monkeypatch):
    done = 0
    for ri in RATES:
        x = np.random.default_rng(ri).uniform(-0.5, 0.5, (secs * ri, ch))
        for ro in RATES:
            if ro == ri:
                continue
            monkeypatch.delenv("KBSRC_GENERIC_PASSES", raising=False)
            a = gpu().resample(x, ri, ro, mode)
            monkeypatch.setenv("KBSRC_GENERIC_PASSES", "1")
            b = gpu().resample(x, ri, ro, mode)
            assert a.tobytes() == b.tobytes(), (ri, ro, mode, ch, secs)
            done += 1
    print(f"{mode} {ch}ch {secs}s: {done} rate pairs identical", flush=True)
    sys.stdout.flush()
