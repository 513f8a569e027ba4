"""The two-part route of long sub-FFTs, checked on the CPU by running both orders of the same arithmetic.

A sub-FFT of n = R M points whose stage list splits after stage k (R = the product of the first k radices) runs
as a head, an R-point sub-FFT on each column c + M l whose butterfly p takes the long transform's twiddles of
butterfly c + M p, stored at c R + e, and a tail, an M-point sub-FFT on each class q + R j with the long strides
divided by R and the same butterfly indices. Here a Stockham stage in the reference's indexing (src/dsp/fft.hpp:
123-145: butterfly b = p s + q reads b + t n / r, writes s r p + q + u s, twiddle row p) runs on random complex data
with random twiddle rows, once as the long transform and once as head and tail: the two must agree exactly (the
same operations on the same values in the same order), and a wrong twiddle index or a wrong split must not.
"""
import numpy as np
import pytest

CASES = [([2, 3, 3, 5, 5, 7], 3), ([2, 3, 3, 5, 5, 7], 4), ([5, 5, 5, 5, 5], 3), ([5, 5, 5, 5, 5], 2),
         ([4, 4, 4, 2, 5, 5], 3), ([4, 2, 3, 3, 3, 3, 5], 4), ([4, 4, 4, 4, 4], 2)]


def strides(radices):
    out, s = [], 1
    for r in radices:
        out.append(s)
        s *= r
    return out


def stage(x, n, r, s, tw_row):
    """One Stockham stage: every butterfly a radix-r DFT of its inputs, outputs u >= 1 times tw_row(p)[u - 1] (row 0
    of a long transform's table is ones: butterfly 0 has no twiddle)."""
    y = np.empty_like(x)
    nbf = n // r
    F = np.exp(-2j * np.pi * np.outer(np.arange(r), np.arange(r)) / r)
    for b in range(nbf):
        p, q = divmod(b, s)
        v = F @ x[[b + t * nbf for t in range(r)]]
        v[1:] = v[1:] * tw_row(p)
        y[[s * r * p + q + u * s for u in range(r)]] = v
    return y


def run(radices, k, head_tw_index=None):
    rng = np.random.default_rng(len(radices) * 100 + k)
    n = int(np.prod(radices))
    R = int(np.prod(radices[:k]))
    M = n // R
    st = strides(radices)
    # twiddle rows of the long transform: stage i, butterfly p -> r_i - 1 values (row 0 unused: no twiddle)
    W = [rng.standard_normal((n // (st[i] * r), r - 1)) + 1j * rng.standard_normal((n // (st[i] * r), r - 1))
         for i, r in enumerate(radices)]
    for w in W:
        w[0] = 1.0
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)

    mid = x.copy()
    for i in range(k):
        mid = stage(mid, n, radices[i], st[i], lambda p, i=i: W[i][p])
    full = mid.copy()
    for i in range(k, len(radices)):
        full = stage(full, n, radices[i], st[i], lambda p, i=i: W[i][p])

    idx = head_tw_index or (lambda c, p: c + M * p)
    head_out = np.empty_like(x)
    for c in range(M):
        v = x[[c + M * l for l in range(R)]]
        for i in range(k):
            v = stage(v, R, radices[i], st[i], lambda p, i=i, c=c: W[i][idx(c, p)])
        head_out[[c * R + e for e in range(R)]] = v
    tail_out = np.empty_like(x)
    for q in range(R):
        v = head_out[[j * R + q for j in range(M)]]
        for i in range(k, len(radices)):
            v = stage(v, M, radices[i], st[i] // R, lambda p, i=i: W[i][p])
        tail_out[[e * R + q for e in range(M)]] = v
    return mid, head_out, full, tail_out


@pytest.mark.parametrize("radices,k", CASES)
def test_head_then_tail_is_the_long_transform(radices, k):
    mid, head_out, full, tail_out = run(radices, k)
    assert np.array_equal(head_out, mid), "the head leaves the long transform's intermediate at c R + e"
    assert np.array_equal(tail_out, full), "the tail finishes it"


def test_a_wrong_twiddle_index_is_caught():
    """Taking the head's own butterfly index p (instead of c + M p) changes the result."""
    radices, k = [2, 3, 3, 5, 5, 7], 3
    mid, head_out, _, _ = run(radices, k, head_tw_index=lambda c, p: p)
    assert not np.allclose(head_out, mid)
