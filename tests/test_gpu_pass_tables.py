"""The tables of specialised passes (kbsrc_amd/csrc/kernels_fixed*.hip, measured and generated): for every entry,
the shortest conversion between the common rates (8 kHz to 192 kHz, 12, 24 and 64 kHz included; stereo inputs of
0.5 s to 600 s; both modes) whose plan runs it, found on the CPU from the oracle's plans, gives the same bits
through the specialised passes as through the generic runtime-radix kernel (KBSRC_GENERIC_PASSES=1), whose
arithmetic the oracle tests pin on many shapes.

The pass shapes of a plan follow the engine's rule (CFFTPlan::init in kbsrc_amd/csrc/engine.cpp, after the
reference's src/dsp/fft.hpp:226-313): a complex FFT of n points runs as one sub-FFT pass when n <= 2048, else as the
four-step pair N1 = the largest divisor of n with N1^2 <= n (column pass, from the forward input or the
inverse's) and N2 = n / N1 (row pass, to the spectrum or the output), unless N1 < 8; the forward input is the
real-input packing (LD_REAL_F64), the inverse ends in the output store (ST_OUT). Entries no such conversion runs
are the opt-in fused forms' (LD_FOLD, ST_SPLIT), the extreme-ratio plans' (kernels_fixed_extreme.hip, compared
with the reference on their own inputs by the extreme-ratio tests) and the 16-point baseline of the split-precision
comparison. The shape lists are checked on the CPU; the bit comparison needs an MI355X."""
import glob
import os
import re
import sys

import numpy as np
import pytest

import helpers as H

RATES = [8000, 11025, 12000, 16000, 22050, 24000, 32000, 44100, 48000, 64000, 88200, 96000, 176400, 192000]
SECONDS = (0.5, 0.6, 0.75, 0.9, 1, 1.25, 1.5, 1.75, 2, 2.5, 3, 4, 5, 7, 10, 15, 20, 30, 45, 60, 90, 120, 180, 300,
           600)
CSRC = os.path.join(H.ROOT, "kbsrc_amd", "csrc")


def table_entries():
    """{(policy, length, first load, last store): table file} over every specialised-pass table."""
    out = {}
    for path in sorted(glob.glob(os.path.join(CSRC, "kernels_fixed*.hip"))):
        src, name = open(path).read(), os.path.basename(path)
        for m in re.finditer(r"KB_FIXED\((POL_\w+), \d+, \d+, (LD_\w+), (ST_\w+), ([\d, ]+)\)", src):
            n = 1
            for r in m.group(4).split(","):
                n *= int(r)
            out[(m.group(1), n, m.group(2), m.group(3))] = name
        for m in re.finditer(r"KB_AUTO\((POL_\w+), (\d+), (LD_\w+), (ST_\w+)\)", src):
            out[(m.group(1), int(m.group(2)), m.group(3), m.group(4))] = name
    return out


def pass_shapes(n, first_load, last_store, dd):
    if n > 2048:
        n1 = max(d for d in range(1, int(n ** 0.5) + 1) if n % d == 0)
        if n1 >= 8:
            pol = "POL_DDS" if dd else "POL_F64"
            return [(pol, n1, first_load, "ST_CPLX"), (pol, n // n1, "LD_CPLX", last_store)]
    return [("POL_DDA" if dd else "POL_F64", n, first_load, last_store)]


def covering_conversions():
    """For each table entry, the first conversion (shortest input, then mode and rates in list order) whose plan
    runs it; returns {conversion: [entries]} and the entries no conversion reaches."""
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
    return convs, {sh: f for sh, f in want.items() if sh not in first}


CONVERSIONS, UNREACHED = covering_conversions()
# conversions grouped by mode and input rate, one test each
GROUPS = {}
for _c in sorted(CONVERSIONS, key=lambda c: (c[2] * c[0], c)):
    GROUPS.setdefault((_c[3], _c[0]), []).append(_c)


def test_generated_table_is_reached():
    """The last generated table: its 41 lengths were added for these conversions, and each is reached."""
    auto4 = {sh for sh, f in table_entries().items() if f == "kernels_fixed_auto4.hip"}
    assert len(auto4) == 41
    assert not auto4 & set(UNREACHED)


def test_unreached_entries_are_the_other_forms():
    assert len(CONVERSIONS) > 200 and len(table_entries()) - len(UNREACHED) > 300
    for (pol, n, ld, st), f in UNREACHED.items():
        assert (ld == "LD_FOLD" or st == "ST_SPLIT" or f == "kernels_fixed_extreme.hip"
                or (pol, n) == ("POL_DDS", 16)), (pol, n, ld, st, f)


@pytest.mark.gpu
@pytest.mark.parametrize("mode,ri", sorted(GROUPS), ids=lambda v: str(v))
def test_specialised_passes_equal_generic(mode, ri, monkeypatch):
    for (_, ro, secs, _) in GROUPS[(mode, ri)]:
        x = np.random.default_rng(ri * 7 + ro).uniform(-0.5, 0.5, (int(secs * ri), 2))
        monkeypatch.delenv("KBSRC_GENERIC_PASSES", raising=False)
        a = kbsrc().resample(x, ri, ro, mode)
        monkeypatch.setenv("KBSRC_GENERIC_PASSES", "1")
        b = kbsrc().resample(x, ri, ro, mode)
        entries = CONVERSIONS[(ri, ro, secs, mode)]
        assert a.tobytes() == b.tobytes(), (ri, ro, secs, mode, entries)
    print(f"{mode} from {ri}: {len(GROUPS[(mode, ri)])} conversions identical", flush=True)
    sys.stdout.flush()


def kbsrc():
    import kbsrc_amd
    return kbsrc_amd
