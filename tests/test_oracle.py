"""The oracle (oracle/kbsrc_oracle.cpp, a CPU restatement of the reference) pinned against the reference's own
golden values, the compiled reference under oracle/_ref, and closed forms. CPU only."""
import math
import struct

import numpy as np
import pytest

import helpers as H


def test_inputs_match_reference_pins():
    assert H.sha(H.wav_f64(H.primary97(), 2, 48000)) == H.PRIMARY_WAV
    assert H.sha(H.wav_f64(H.subnormal97(), 2, 48000)) == H.SUBNORMAL_WAV
    assert H.sha(H.wav_f64(H.input_48k_1s(), 1, 48000)) == H.IN_48K_1S


@pytest.mark.parametrize("name,rate,mode,want", H.CLI_PINS)
def test_oracle_cli_pins(name, rate, mode, want):
    x = H.primary97() if name == "primary" else H.subnormal97()
    y = H.oracle_resample(x, 2, 48000, rate, mode)
    assert H.wav_sha_planar(y, rate) == want


def test_oracle_frame1_and_planar_pins():
    y = H.oracle_resample(np.array(H.FRAME1), 2, 8000, 16000, "exact", length=0.01)
    assert [struct.pack(">d", v).hex() for v in y.reshape(-1)] == H.FRAME1_OUT
    y = H.oracle_resample(H.subnormal97(), 2, 48000, 96000, "exact", length=0.1)
    assert H.sha(y.astype("<f8").tobytes()) == H.SUBNORMAL_96K_PLANAR


@pytest.mark.parametrize("mode,lpc", [("exact", True), ("fast", True), ("exact", False)])
def test_oracle_c1_pins(mode, lpc):
    y = H.oracle_resample(H.input_48k_1s(), 1, 48000, 44100, mode, lpc=lpc)
    assert H.wav_sha_planar(y, 44100) == H.PINS_48K_1S[(mode, lpc)]
    if mode == "exact" and lpc:   # first samples and peak quoted with the pins
        assert y[0, :3].tolist() == [float.fromhex("0x1.4323a1f5ecdb9p-4"), float.fromhex("0x1.50fb438506410p-2"),
                                     float.fromhex("-0x1.07f0f20da41efp-2")]


# Plans quoted from the reference's `-v` line and tests/test_robust.py (block=, taps=).
@pytest.mark.parametrize("args,want", [
    ((48000, 1, 48000, 44100, "exact"), dict(P=147, Q=160, nb=250, Nb=40000, Mb=36750, hop=126, nblk=4, o=9114)),
    ((17_280_000, 2, 96000, 44100, "exact"), dict(P=147, Q=320, nb=630, Nb=201600, Mb=92610, hop=506, nblk=107,
                                                   marg=62, E=19521, pre=19840, o=9114, Lk=19200, c2=44036)),
    ((17_280_000, 2, 96000, 44100, "fast"), dict(nb=250, Nb=80000, Mb=36750)),
    ((44100 * 180, 2, 44100, 48000, "exact"), dict(P=160, Q=147, nb=630)),
    ((300_000_000, 1, 10_000_000, 48000, "exact", 0.5), dict(P=3, Q=625, nb=32256, Nb=20_160_000, Mb=96768, nblk=31)),
    ((230_400_000, 8, 384000, 44100, "exact", 0.5), dict(P=147, Q=1280, nb=1536, Nb=1_966_080, Mb=225_792, nblk=147)),
    ((64, 2, 48000, 96000, "exact", 0.1, 40.0, 9605), dict(nb=9720)),
    ((64, 2, 48000, 96000, "fast", 0.1, 40.0, 20000), dict(nb=20000)),
    ((64, 2, 48000, 96000, "exact", 5.33e-4), dict(Lk=51)),
    ((64, 2, 48000, 44100, "exact", 5.8e-4), dict(Lk=27)),
    ((64, 2, 48000, 16000, "exact", 10.0), dict(Lk=480000)),
])
def test_oracle_plans(args, want):
    p = H.oracle_plan(*args)
    for k, v in want.items():
        assert p[k] == v, (k, p[k], v)


@pytest.mark.parametrize("args,msg", [
    ((64, 2, 48000, 96000, "exact", 0.1, 40.0, 9604), "option out of range"),
    ((64, 2, 48000, 96000, "exact", 0.1, 40.0, 1), "option out of range"),
    ((64, 2, 48000, 96000, "exact", 0.2, 40.0, 536870912), "block, rates or length too large"),
    ((64, 2, 48000, 96000, "exact", 1e-5), "option out of range"),
    ((64, 2, 48000, 96000, "exact", 5.3e-4), "length too short for these sample rates"),
    ((64, 2, 48000, 44100, "exact", 5.79e-4), "length too short for these sample rates"),
    ((64, 2, 48000, 44100, "exact", 5.8e-4, 80.0), "length too short for these sample rates"),
    ((64, 2, 48000, 96000, "exact", 10.5), "option out of range"),
    ((64, 2, 0, 96000, "exact"), "sample rates must be positive"),
    ((64, 0, 48000, 96000, "exact"), "bad input buffer"),
])
def test_oracle_rejections(args, msg):
    with pytest.raises(ValueError, match=msg):
        H.oracle_plan(*args)


# ---- closed forms: what the mathematics fixes independently of the reference --------------------------------
mpmath = pytest.importorskip("mpmath")


def _dd(v):
    return mpmath.mpf(v[0]) + mpmath.mpf(v[1])


@pytest.mark.parametrize("k,N", [(0, 7), (1, 3), (2, 5), (3, 7), (1, 64), (17, 64), (12345, 100800), (99999, 201600),
                                 (5, 8), (201599, 201600), (7, 28)])
def test_oracle_cis_against_256bit(k, N):
    mpmath.mp.prec = 256
    v = np.zeros(4)
    H.oracle().oracle_cis(k, N, v.ctypes.data)
    ang = 2 * mpmath.pi * k / N
    assert abs(_dd(v[:2]) - mpmath.cos(ang)) < mpmath.mpf(2) ** -104
    assert abs(_dd(v[2:]) - mpmath.sin(ang)) < mpmath.mpf(2) ** -104


def test_oracle_cis_table_against_256bit():
    mpmath.mp.prec = 256
    N, k0, cnt = 100800, -50, 400
    v = np.zeros(4 * cnt)
    H.oracle().oracle_cis_table(N, k0, cnt, v.ctypes.data)
    for i in range(0, cnt, 7):
        ang = 2 * mpmath.pi * (k0 + i) / N
        assert abs(_dd(v[4 * i:4 * i + 2]) - mpmath.cos(ang)) < mpmath.mpf(2) ** -102
        assert abs(_dd(v[4 * i + 2:4 * i + 4]) - mpmath.sin(ang)) < mpmath.mpf(2) ** -102


def test_oracle_kernel_taps_closed_form():
    # h[n] = (a/b) sinc(pi n a/b) I0(beta sqrt(1-(n/L)^2)) / I0(beta); h[0] = a/b
    mpmath.mp.prec = 256
    L, a, b, beta = 40, 11009, 24000, 40.0
    v = np.zeros(2 * (L + 1))
    assert H.oracle().oracle_kernel_taps(L, a, b, beta, v.ctypes.data) == 0
    r = mpmath.mpf(a) / b
    for n in range(L + 1):
        x = mpmath.pi * n * r
        s = mpmath.mpf(1) if n == 0 else mpmath.sin(x) / x
        w = mpmath.besseli(0, beta * mpmath.sqrt(1 - (mpmath.mpf(n) / L) ** 2)) / mpmath.besseli(0, beta)
        want = r * s * w
        assert abs(_dd(v[2 * n:2 * n + 2]) - want) <= mpmath.mpf(2) ** -98 * max(abs(want), mpmath.mpf(2) ** -60)
    assert v[0] == float(mpmath.mpf(a) / b) or abs(_dd(v[:2]) - r) < mpmath.mpf(2) ** -104


def test_oracle_sampled_taps_equal_full_taps():
    """The spot-check export (single taps) gives the same bits as the full tap vector."""
    for L, a, b, beta in [(3000, 11009, 24000, 40.0), (2000, 999999937, 1000000007, 200.0), (7, 2, 3, 5.0)]:
        full = H.oracle_taps(L, a, b, beta)
        idx = np.unique(np.concatenate([[0, 1, L], np.random.default_rng(L).integers(0, L + 1, 200)]))
        assert H.oracle_taps(L, a, b, beta, idx).tobytes() == full[idx].tobytes()


@pytest.mark.parametrize("n,kind", [(60, 0), (105, 0), (128, 0), (2100, 0), (2100, 2), (2401, 0), (96, 1),
                                    (2100, 1), (98, 2)])
def test_oracle_cfft_against_dft(n, kind):
    """Forward DFT (sign -1) against numpy's (FP64: 1e-13) and, for dd, against a 256-bit DFT of the same input."""
    rng = np.random.default_rng(n + kind)
    x = rng.uniform(-1, 1, (n, 2))
    inp = np.zeros(4 * n)
    inp[0::4], inp[2::4] = x[:, 0], x[:, 1]
    out = np.zeros(4 * n)
    assert H.oracle().oracle_cfft(n, kind, inp.ctypes.data, out.ctypes.data) == 0
    ref = np.fft.fft(x[:, 0] + 1j * x[:, 1])
    got = (out[0::4] + out[1::4]) + 1j * (out[2::4] + out[3::4])
    assert np.abs(got - ref).max() < 1e-12 * math.sqrt(n)
    if kind != 1 and n <= 128:   # dd accuracy, far below FP64: exact DFT at 256 bits
        mpmath.mp.prec = 256
        xs = [mpmath.mpc(a, b) for a, b in x]
        for k in range(0, n, max(1, n // 9)):
            s = mpmath.fsum(xs[j] * mpmath.expjpi(-2 * mpmath.mpf(j * k % n) / n) for j in range(n))
            g = mpmath.mpc(_dd(out[4 * k:4 * k + 2]), _dd(out[4 * k + 2:4 * k + 4]))
            assert abs(g - s) < mpmath.mpf(2) ** -96 * n


def test_oracle_burg_predicts_sinusoid():
    """A sum of two sinusoids is an AR(4) process: the order-32 Burg model continues it (to the conditioning of
    the over-fitted recursion, ~1e-6 here), and the order-2 part of the recursion alone would be exact."""
    t = np.arange(4096)

    def sig(u):
        return 0.5 * np.sin(2 * np.pi * 0.1 * u + 0.3) + 0.2 * np.sin(2 * np.pi * 0.0371 * u)
    x = sig(t)
    n = 200
    out = np.zeros(n)
    H.oracle().oracle_burg(x.ctypes.data, len(x), n, out.ctypes.data)
    assert np.abs(out - sig(4096 + np.arange(n))).max() < 1e-5
    z = np.zeros(n)
    H.oracle().oracle_burg(np.zeros(100).ctypes.data, 100, n, z.ctypes.data)
    assert not z.any()   # a silent segment predicts zeros
