"""The last generated table of specialised passes (kbsrc_amd/csrc/kernels_fixed_auto4.hip: lengths that conversions
between the common rates, 12, 24 and 64 kHz included, need for inputs of 0.5 s to 600 s): one conversion for each of
its entries, found on the CPU from the oracle's plans, gives the same bits through the specialised passes as through
the generic runtime-radix kernel (KBSRC_GENERIC_PASSES=1), whose arithmetic the oracle tests pin on many shapes.

The pass shapes of a plan follow the engine's rule (CFFTPlan::init in kbsrc_amd/csrc/engine.cpp, after the
reference's src/dsp/fft.hpp:226-313): a complex FFT of n points runs as one sub-FFT pass when n <= 2048, else as the
four-step pair N1 = the largest divisor of n with N1^2 <= n (column pass, from the forward input or the
inverse's) and N2 = n / N1 (row pass, to the spectrum or the output), unless N1 < 8; the forward input is the
real-input packing (LD_REAL_F64), the inverse ends in the output store (ST_OUT). The first shape list is checked on
the CPU; the bit comparison needs an MI355X."""
import os
import re
import sys

import numpy as np
import pytest

import helpers as H

RATES = [8000, 11025, 12000, 16000, 22050, 24000, 32000, 44100, 48000, 64000, 88200, 96000, 176400, 192000]
SECONDS = (0.5, 0.6, 0.75, 0.9, 1, 1.25, 1.5, 1.75, 2, 2.5, 3, 4, 5, 7, 10, 15, 20, 30, 45, 60, 90, 120, 180, 300,
           600)
TABLE = os.path.join(H.ROOT, "kbsrc_amd", "csrc", "kernels_fixed_auto4.hip")


def table_entries():
    src = open(TABLE).read()
    return {(m.group(1), int(m.group(2)), m.group(3), m.group(4))
            for m in re.finditer(r"KB_AUTO\((POL_\w+), (\d+), (LD_\w+), (ST_\w+)\)", src)}


def pass_shapes(n, first_load, last_store, dd):
    if n > 2048:
        n1 = max(d for d in range(1, int(n ** 0.5) + 1) if n % d == 0)
        if n1 >= 8:
            pol = "POL_DDS" if dd else "POL_F64"
            return [(pol, n1, first_load, "ST_CPLX"), (pol, n // n1, "LD_CPLX", last_store)]
    return [("POL_DDA" if dd else "POL_F64", n, first_load, last_store)]


def covering_conversions():
    """For each table entry, the first conversion (stereo; shortest input, then mode and rates in list order) whose
    plan runs it; returns {conversion: [entries]} and the entries no conversion reaches."""
    want = table_entries()
    first = {}
    for secs in SECONDS:
        for mode in ("exact", "fast"):
            for ri in RATES:
                for ro in RATES:
                    if ri == ro:
                        continue
                    try:
                        p = H.oracle_plan(int(secs * ri), 2, ri, ro, mode)
                    except ValueError:
                        continue
                    if p["equal"] or p["gen"]:
                        continue
                    dd = mode == "exact"
                    for sh in (pass_shapes(p["Nb"] // 2, "LD_REAL_F64", "ST_CPLX", dd)
                               + pass_shapes(p["Mb"] // 2, "LD_CPLX", "ST_OUT", dd)):
                        if sh in want and sh not in first:
                            first[sh] = (ri, ro, secs, mode)
    convs = {}
    for sh, c in first.items():
        convs.setdefault(c, []).append(sh)
    return convs, want - set(first)


CONVERSIONS, UNREACHED = covering_conversions()


def test_every_entry_is_reached_by_a_conversion():
    assert len(table_entries()) == 41
    assert not UNREACHED, sorted(UNREACHED)


@pytest.mark.gpu
@pytest.mark.parametrize("ri,ro,secs,mode", sorted(CONVERSIONS, key=lambda c: (c[2] * c[0], c)),
                         ids=lambda v: str(v))
def test_generated_passes_equal_generic(ri, ro, secs, mode, monkeypatch):
    import kbsrc_amd
    x = np.random.default_rng(ri * 7 + ro).uniform(-0.5, 0.5, (int(secs * ri), 2))
    monkeypatch.delenv("KBSRC_GENERIC_PASSES", raising=False)
    a = kbsrc_amd.resample(x, ri, ro, mode)
    monkeypatch.setenv("KBSRC_GENERIC_PASSES", "1")
    b = kbsrc_amd.resample(x, ri, ro, mode)
    assert a.tobytes() == b.tobytes(), (ri, ro, secs, mode, CONVERSIONS[(ri, ro, secs, mode)])
    print(f"{ri} -> {ro} {secs} s {mode}: identical ({len(CONVERSIONS[(ri, ro, secs, mode)])} entries)", flush=True)
    sys.stdout.flush()
