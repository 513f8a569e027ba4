"""The split-precision (Ozaki int8 MFMA) radix-16 DFT beside the plain dd pass: errors against a 256-bit DFT
(mpmath) of the same dd inputs, normwise per vector. Not on the conversion path. Needs an MI355X."""
import mpmath
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def gpu():
    import kbsrc_amd
    return kbsrc_amd


def dd_vectors(nvec, seed):
    """(nvec, 16, 4) doubles {re.hi, im.hi, re.lo, im.lo}, lo at most half an ulp of hi, varied magnitudes."""
    rng = np.random.default_rng(seed)
    hi = rng.uniform(-1, 1, (nvec, 16, 2)) * np.exp2(rng.integers(-20, 20, (nvec, 1, 1)))
    lo = hi * rng.uniform(-1, 1, hi.shape) * 2.0 ** -54
    lo = (hi + lo) - hi   # exactly representable remainder
    return np.ascontiguousarray(np.concatenate([hi, lo], axis=2))


def errors(x, y, count):
    """Normwise relative error of y against the exact DFT of x, per vector, for the first count vectors."""
    mpmath.mp.prec = 256
    out = []
    w = [mpmath.expjpi(-mpmath.mpf(2 * k) / 16) for k in range(16)]
    for v in range(count):
        xs = [mpmath.mpc(mpmath.mpf(x[v, n, 0]) + mpmath.mpf(x[v, n, 2]), mpmath.mpf(x[v, n, 1]) + mpmath.mpf(x[v, n, 3]))
              for n in range(16)]
        ref = [mpmath.fsum(xs[n] * w[(k * n) % 16] for n in range(16)) for k in range(16)]
        got = [mpmath.mpc(mpmath.mpf(y[v, k, 0]) + mpmath.mpf(y[v, k, 2]), mpmath.mpf(y[v, k, 1]) + mpmath.mpf(y[v, k, 3]))
               for k in range(16)]
        norm = max(abs(r) for r in ref)
        out.append(float(max(abs(g - r) for g, r in zip(got, ref)) / norm))
    return out


def run(fn, x, *args):
    import torch
    nvec = x.shape[0]
    dx = torch.from_numpy(x.reshape(-1)).cuda()
    dy = torch.zeros_like(dx)
    fn(dx.data_ptr(), nvec, *args, dy.data_ptr())
    torch.cuda.synchronize()
    return dy.cpu().numpy().reshape(nvec, 16, 4)


def test_plain_and_split_errors_against_256_bit():
    x = dd_vectors(100, 3)
    plain = run(gpu().plain_dft16, x)
    e_plain = max(errors(x, plain, 12))
    assert e_plain < 2.0 ** -100
    prev = 1.0
    worst = {}
    for s in (6, 7, 8, 10, 12, 14, 15, 16):
        y = run(gpu().split_dft16, x, s)
        e = max(errors(x, y, 12))
        worst[s] = e
        assert e < prev or e < 2.0 ** -104, (s, e, prev)
        assert e < 2.0 ** (-7 * s + 14), (s, e)   # the digits kept, times the transform's gain
        prev = e
    print({s: round(float(np.log2(e)), 1) for s, e in worst.items()}, "plain", round(float(np.log2(e_plain)), 1))
    assert worst[16] < 2.0 ** -100


@pytest.mark.parametrize("nvec", [1, 31, 32, 33, 128, 1000, 4099])
def test_inwave_exchange_gives_the_lds_pass_bits(nvec):
    """The plain radix-16 stage with its stage exchange inside the wave (v_permlane32_swap) computes the same
    operations as the LDS pass: the same bits on every vector, for counts that fill waves partly (32 vectors per
    wave, 128 per workgroup) and wholly, and the DFT within 2^-100 of a 256-bit one."""
    x = dd_vectors(nvec, 11 + nvec)
    want = run(gpu().plain_dft16, x)
    got = run(gpu().plain_dft16_inwave, x)
    assert got.tobytes() == want.tobytes()
    assert max(errors(x, got, min(nvec, 4))) < 2.0 ** -100
