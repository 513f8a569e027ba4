"""The high-precision block restatement (tests/hp_block.py) pinned before it is used to state errors: its DFT
against a direct 300-bit DFT (mpmath) on small mixed-radix lengths, and whole blocks against the oracle's exact
conversion, which rounds the dd block result once to double, so on every sample it must be within half an ulp
(plus the dd error, 2^-100 of the peak) of the exact value. A wrong bin map, sign, scale or discard offset moves
samples by far more than that."""
import math

import numpy as np
import pytest

import helpers as H
import hp_block as hb

mpmath = pytest.importorskip("mpmath")


@pytest.mark.parametrize("n", [8, 12, 30, 49, 60, 64, 105])
def test_dft_against_direct_300_bit(n):
    mpmath.mp.prec = 300
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    re, im = hb.dft(hb.fix(x.real), hb.fix(x.imag), n)
    xs = [mpmath.mpc(float(v.real), float(v.imag)) for v in x]
    worst = mpmath.mpf(0)
    for k in range(n):
        want = mpmath.fsum(xs[t] * mpmath.expjpi(-mpmath.mpf(2 * t * k) / n) for t in range(n))
        got = mpmath.mpc(mpmath.mpf(int(re[k])) / 2 ** hb.F, mpmath.mpf(int(im[k])) / 2 ** hb.F)
        worst = max(worst, abs(got - want))
    assert worst < mpmath.mpf(2) ** -240


def _interior_block(x, p):
    """First block whose Nb input samples all lie inside x (no padding or edge extension)."""
    off = p["base"] + p["pre"]
    for b in range(p["nblk"]):
        lo = b * p["hop"] * p["Q"] - off
        if lo >= 0 and lo + p["Nb"] <= len(x):
            return b, lo
    raise AssertionError("no interior block")


@pytest.mark.parametrize("fi,fo", [(48000, 44100), (44100, 48000)])
def test_block_against_oracle_exact(fi, fo):
    n = 4 * fi
    x = np.random.default_rng(1).uniform(-0.5, 0.5, n)
    p = H.oracle_plan(n, 1, fi, fo)
    assert not p["gen"]
    y = H.oracle_resample(x, 1, fi, fo)[0]
    fk = fo if p["up"] else fi
    gc = math.gcd(p["c2"], fk)
    taps = H.oracle_taps(p["Lk"], p["c2"] // gc, fk // gc, 40.0)
    spec = H.oracle_spectrum(taps, p["Mb"] if p["up"] else p["Nb"], p["Nb"])
    b, lo = _interior_block(x, p)
    ref = hb.block_output(x[lo:lo + p["Nb"]], spec, p["P"], p["Q"], p["nb"], p["marg"], bool(p["up"]))
    d0 = b * p["hop"] * p["P"] - p["o"]
    got = y[d0:d0 + p["hop"] * p["P"]]
    err = hb.errors(got, ref)
    r = hb.to_double(ref)
    peak = np.abs(r).max()
    assert peak > 0.1
    assert np.all(err <= 0.5 * np.spacing(np.abs(r)) + 2.0 ** -100 * peak), (err / np.spacing(np.abs(r))).max()
    # and the rounding is not trivially loose: most samples are off by a visible fraction of an ulp
    assert (err / np.spacing(np.abs(r))).max() > 0.4
