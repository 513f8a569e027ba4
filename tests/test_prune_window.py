"""The stage pruning of the specialised passes, checked on the CPU by brute force.

A pass's sub-FFT outputs e are needed where some element i = e ste + j stg (j < nsub) is read: i <= zkeep or
i >= zh - zkeep (the forward transform's bins up to the bin map's last one, and their real-split partners). The
passes take the union over j as one cyclic window of e and run, in a Stockham stage of stride s, only the butterflies
(p, q) with q in the window reduced mod s. Here the butterflies whose outputs actually reach a needed output are
found by following the Stockham index flow backwards (the reference's stage indexing, src/dsp/fft.hpp:123-145:
butterfly b = p s + q reads b + t N / r and writes s r p + q + u s), and compared with that rule: the rule must keep
every such butterfly and, for the stages it prunes, no other.
"""
import random

import pytest

RADIX_LISTS = [[4, 4, 4, 2, 5, 5], [2, 3, 3, 5, 5, 7], [4, 4, 4, 5], [3, 3, 5, 7], [5, 7, 7], [4, 4, 4, 4, 4],
               [4, 2, 3, 3, 3, 3, 5], [5, 5, 5, 5, 5], [4, 4, 2, 7], [3, 7, 7]]


def window(n, zkeep, zh, ste, stg, nsub):
    """(start, length) of the needed outputs e as a cyclic window mod n; None for all of them."""
    if zkeep < 0 or ste <= 0 or n < 2:
        return None
    elo = zkeep // ste
    top = zh - zkeep - (nsub - 1) * stg
    ehi = 0 if top <= 0 else -(-top // ste)
    if elo >= n - 1 or ehi <= elo + 1:
        return None
    if ehi >= n:
        return 0, elo + 1
    return ehi, n - ehi + elo + 1


def needed_butterflies(radices, needed_out):
    """Per stage, the set of butterflies (p, q) one of whose outputs reaches a needed output (brute force)."""
    n = 1
    for r in radices:
        n *= r
    strides = []
    s = 1
    for r in radices:
        strides.append(s)
        s *= r
    per_stage = [None] * len(radices)
    need = set(needed_out)   # needed outputs of the current stage (element indices)
    for k in range(len(radices) - 1, -1, -1):
        r, s = radices[k], strides[k]
        nbf = n // r
        bfs, inputs = set(), set()
        for b in range(nbf):
            p, q = divmod(b, s)
            if any(s * r * p + q + u * s in need for u in range(r)):
                bfs.add((p, q))
                inputs.update(b + t * nbf for t in range(r))
        per_stage[k] = bfs
        need = inputs
    return per_stage, strides


def window_butterflies(radices, strides, win):
    n = 1
    for r in radices:
        n *= r
    out = []
    for r, s in zip(radices, strides):
        m = n // (s * r)
        start, length = win
        if length >= s:
            out.append({(p, q) for p in range(m) for q in range(s)})
        else:
            out.append({(p, (start + k) % s) for p in range(m) for k in range(length)})
    return out


def cases():
    rng = random.Random(20261204)
    out = []
    for radices in RADIX_LISTS:
        n = 1
        for r in radices:
            n *= r
        for _ in range(6):
            if rng.random() < 0.5:   # a four-step row pass: N1 rows of n, element e of row j at e N1 + j
                n1 = rng.choice([3, 7, 16, 45, 128])
                zh, ste, stg, nsub = n * n1, n1, 1, n1
            else:                    # a single pass: element e at e
                zh, ste, stg, nsub = n, 1, 0, 1
            zkeep = rng.randrange(0, zh // 2)
            out.append((radices, zkeep, zh, ste, stg, nsub))
    return out


@pytest.mark.parametrize("radices,zkeep,zh,ste,stg,nsub", cases())
def test_window_keeps_exactly_the_needed_butterflies(radices, zkeep, zh, ste, stg, nsub):
    n = zh // nsub if nsub > 1 else zh
    needed_out = {e for e in range(n) for j in range(nsub)
                  if e * ste + j * stg <= zkeep or e * ste + j * stg >= zh - zkeep}
    win = window(n, zkeep, zh, ste, stg, nsub)
    if win is None:
        # every output e is needed by some j: no stage may be pruned
        assert needed_out == set(range(n))
        return
    start, length = win
    assert needed_out == {(start + k) % n for k in range(length)}, "the window is the union over the sub-FFTs"
    brute, strides = needed_butterflies(radices, needed_out)
    rule = window_butterflies(radices, strides, win)
    for k in range(len(radices)):
        assert brute[k] <= rule[k], f"stage {k}: a needed butterfly is skipped"
        if win[1] < strides[k]:
            assert brute[k] == rule[k], f"stage {k}: the pruned stage runs a butterfly no reader needs"


def test_the_10mhz_row_pass_prunes_its_last_three_stages():
    """The 10 MHz -> 48 kHz (--len 0.5) plan's forward rows: Nb = 20,160,000, Mb = 96,768, so the bin map reads the
    spectrum up to bin min(Nb / 2, Mb - 1) = 96,767 (src/resampler.cpp:66); 3150 rows of 3200 points."""
    zkeep = min(20160000 // 2, 96768 - 1)
    win = window(3200, zkeep, 10080000, 3150, 1, 3150)
    assert win == (3169, 62)
    radices = [4, 4, 4, 2, 5, 5]
    _, strides = needed_butterflies(radices, set())
    kept = [min(win[1], s) / s for s in strides]
    assert kept == [1.0, 1.0, 1.0, 62 / 64, 62 / 128, 62 / 640]
