"""High-precision restatement of one overlap-save block, for error statements against the engines: the block's
forward real DFT, the multiply by the kernel spectrum H, the bin map (fold or periodic extension) and the
unnormalised inverse real DFT (src/resampler.cpp:34-88, the non-general fold; the engine's structure as the oracle
restates it in oracle/kbsrc_oracle.cpp `engine`), evaluated with no rounding but that of fixed-point products at
2^-256 (Python integers). H is taken as given (its dd values are exact inputs), so the result is the exact value
of the block the engines compute, to ~2^-240 relative.

Test infrastructure only: used by tests/ and bench.py's split measurement as a checker.

The DFTs are mixed-radix Stockham passes over numpy object arrays of Python integers; a 201,600-point block takes
a few seconds.
"""
import math

import numpy as np

F = 256               # fractional bits of the fixed-point values
_ONE = 1 << F


def _fix_scalar(x: float) -> int:
    """x * 2^F rounded to an integer (exact for every double at or above 2^-F)."""
    if x == 0.0:
        return 0
    m, e = math.frexp(x)
    mi = int(math.ldexp(m, 53))
    sh = e - 53 + F
    if sh >= 0:
        return mi << sh
    q, r = divmod(mi, 1 << -sh)
    half = 1 << (-sh - 1)
    return q + (1 if r > half or (r == half and q & 1) else 0)


_fix_vec = np.frompyfunc(_fix_scalar, 1, 1)


def fix(a) -> np.ndarray:
    """Doubles -> object array of fixed-point integers."""
    return _fix_vec(np.asarray(a, dtype=np.float64)).astype(object)


def _cis_fixed(k: int, n: int, mp):
    """(cos, sin) of -2 pi k / n, fixed point (mpmath at 320 bits)."""
    a = -2 * mp.pi * mp.mpf(k) / n
    return int(mp.nint(mp.cos(a) * _ONE)), int(mp.nint(mp.sin(a) * _ONE))


def _roots(n: int):
    """w_n^j = e^{-2 pi i j / n}, j = 0..n-1, as two object arrays: a two-level product of ~2 sqrt(n) values."""
    import mpmath
    mp = mpmath.mp
    saved = mp.prec
    mp.prec = 320
    try:
        B = max(1, math.isqrt(n))
        lo = [_cis_fixed(b, n, mp) for b in range(B)]
        hi = [_cis_fixed(a * B, n, mp) for a in range((n + B - 1) // B)]
    finally:
        mp.prec = saved
    j = np.arange(n)
    lr = np.array([c for c, _ in lo], dtype=object)[j % B]
    li = np.array([s for _, s in lo], dtype=object)[j % B]
    hr = np.array([c for c, _ in hi], dtype=object)[j // B]
    hi_ = np.array([s for _, s in hi], dtype=object)[j // B]
    return (hr * lr - hi_ * li) >> F, (hr * li + hi_ * lr) >> F


def _radices(n: int):
    rs, t = [], n
    for f in (4, 2, 3, 5, 7, 11, 13):
        while t % f == 0:
            rs.append(f)
            t //= f
    if t != 1:
        raise ValueError("length with a prime factor above 13")
    return rs


def dft(re, im, n: int, roots=None):
    """Forward DFT X[k] = sum_t x[t] w_n^{t k} of fixed-point complex data (object arrays), Stockham passes."""
    wr, wi = roots if roots is not None else _roots(n)
    a_re, a_im = np.asarray(re, dtype=object).copy(), np.asarray(im, dtype=object).copy()
    rem, s = n, 1
    for r in _radices(n):
        m = rem // r
        xr = a_re.reshape(r, m, s)
        xi = a_im.reshape(r, m, s)
        yr = np.empty((m, r, s), dtype=object)
        yi = np.empty((m, r, s), dtype=object)
        p = np.arange(m)
        for u in range(r):
            if r == 4 or r == 2:   # w_r^{t u} is 1, -i, -1 or i: exact
                ar, ai = xr[0].copy(), xi[0].copy()
                for t in range(1, r):
                    e = (t * u * (4 // r)) % 4
                    if e == 0:
                        ar, ai = ar + xr[t], ai + xi[t]
                    elif e == 1:   # * -i
                        ar, ai = ar + xi[t], ai - xr[t]
                    elif e == 2:
                        ar, ai = ar - xr[t], ai - xi[t]
                    else:          # * i
                        ar, ai = ar - xi[t], ai + xr[t]
            else:
                ar, ai = xr[0] << F, xi[0] << F
                for t in range(1, r):
                    c, sn = wr[(t * u % r) * (n // r)], wi[(t * u % r) * (n // r)]
                    ar = ar + xr[t] * c - xi[t] * sn
                    ai = ai + xr[t] * sn + xi[t] * c
                ar, ai = ar >> F, ai >> F
            if u:   # twiddle w_n^{s p u}
                idx = (s * p * u) % n
                cr, ci = wr[idx][:, None], wi[idx][:, None]
                ar, ai = (ar * cr - ai * ci) >> F, (ar * ci + ai * cr) >> F
            yr[:, u, :] = ar
            yi[:, u, :] = ai
        a_re, a_im = yr.reshape(n), yi.reshape(n)
        rem //= r
        s *= r
    return a_re, a_im


def block_output(staged_block, H, P: int, Q: int, nb: int, marg: int, up: bool):
    """Fixed-point output samples r[marg P + i], i < (nb - 2 marg) P, of one block: staged_block = the Nb = nb Q
    padded input samples the block reads (doubles), H = the kernel spectrum as (hi, lo) rows (Nb / 2 + 1 of them
    downsampling, Mb / 2 + 1 upsampling)."""
    Nb, Mb = nb * Q, nb * P
    K, h = Mb // 2 + 1, Nb // 2
    x = fix(staged_block)
    assert len(x) == Nb
    Hf = fix(np.asarray(H, dtype=np.float64)[:, 0]) + fix(np.asarray(H, dtype=np.float64)[:, 1])
    zero = np.zeros(Nb, dtype=object)
    Xr, Xi = dft(x, zero, Nb)
    Xr, Xi = Xr[:h + 1], Xi[:h + 1]
    Yr = np.zeros(K, dtype=object)
    Yi = np.zeros(K, dtype=object)
    if not up:
        Yr[:] = (Xr[:K] * Hf[:K]) >> F
        Yi[:] = (Xi[:K] * Hf[:K]) >> F
        jmax = min(h, Mb - 1)
        j = np.arange(K, jmax + 1)
        if len(j):
            vr, vi = (Xr[j] * Hf[j]) >> F, (Xi[j] * Hf[j]) >> F
            np.add.at(Yr, Mb - j, vr)
            np.add.at(Yi, Mb - j, -vi)
        Yr[Mb // 2] = (2 * Xr[Mb // 2] * Hf[Mb // 2]) >> F
        Yi[Mb // 2] = 0
    else:
        k = np.arange(K)
        kk = k % Nb
        lo = kk <= h
        vr = np.where(lo, Xr[np.minimum(kk, h)], Xr[np.minimum(Nb - kk, h)])
        vi = np.where(lo, Xi[np.minimum(kk, h)], -Xi[np.minimum(Nb - kk, h)])
        Yr[:] = (vr * Hf[:K]) >> F
        Yi[:] = (vi * Hf[:K]) >> F
        Yi[Mb // 2] = 0
    # unnormalised inverse: r[t] = sum over the Hermitian spectrum of Y[k] e^{+2 pi i k t / Mb}
    #                          = conj(DFT(conj(Yfull)))[t]
    Fr = np.zeros(Mb, dtype=object)
    Fi = np.zeros(Mb, dtype=object)
    Fr[:K] = Yr
    Fi[:K] = -Yi
    k = np.arange(1, Mb - K + 1)   # Yfull[Mb - k] = conj(Y[k])
    Fr[Mb - k] = Yr[k]
    Fi[Mb - k] = Yi[k]
    rr, _ = dft(Fr, Fi, Mb)
    hopP = (nb - 2 * marg) * P
    return rr[marg * P: marg * P + hopP]


def errors(y, ref) -> np.ndarray:
    """|y - ref| as doubles, y doubles, ref fixed point."""
    d = fix(y) - ref
    return np.array([math.ldexp(float(abs(int(v))), -F) for v in d])


def to_double(ref) -> np.ndarray:
    return np.array([math.ldexp(float(int(v) >> (F - 120)), -120) for v in ref])
