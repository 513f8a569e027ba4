"""The moduli (CRT) form of the split-precision radix-16 DFT, emulated exactly on the CPU (tests/split_schemes.py):
the residue products reconstruct the exact integer product, and with 110-bit operands the scaled result is within
2^-100 of a 256-bit DFT; the vector-ALU work it needs per point is set beside the slice form's and the plain dd
stage's. CPU only."""
import math
import random

import numpy as np

import split_schemes as SS


def dd_vectors(nvec, seed):
    rng = np.random.default_rng(seed)
    hi = rng.uniform(-1, 1, (nvec, 16, 2)) * np.exp2(rng.integers(-20, 20, (nvec, 1, 1)))
    lo = hi * rng.uniform(-1, 1, hi.shape) * 2.0 ** -54
    lo = (hi + lo) - hi
    return np.concatenate([hi, lo], axis=2)


def test_residue_products_reconstruct_the_exact_integer_product():
    rnd = random.Random(5)
    m = [[rnd.randrange(-2 ** 60, 2 ** 60) for _ in range(SS.K)] for _ in range(SS.K)]
    x = [rnd.randrange(-2 ** 60, 2 ** 60) for _ in range(SS.K)]
    mods = SS.moduli_for(60 + 60 + 6)
    assert all(math.gcd(p, q) == 1 for i, p in enumerate(mods) for q in mods[i + 1:])
    assert SS.crt_product(m, x, mods) == [sum(a * b for a, b in zip(row, x)) for row in m]


def test_moduli_scheme_reaches_2_to_minus_100():
    b = bf = 110
    m = SS.dft_matrix(bf)
    mods = SS.moduli_for(b + bf + 6)
    assert len(mods) <= 32
    worst = 0.0
    for v in dd_vectors(4, 9):
        X, e = SS.to_int_vector(v, b)
        Y = SS.crt_product(m, X, mods)
        worst = max(worst, SS.dft_error(v, Y, e - b - bf))
    assert worst < 2.0 ** -100, math.log2(worst)


def test_moduli_up_to_127_cannot_cover_the_range():
    import pytest
    chosen, prod = [], 1
    for p in range(127, 2, -1):
        if all(math.gcd(p, q) == 1 for q in chosen):
            chosen.append(p)
            prod *= p
    assert prod.bit_length() < 226
    with pytest.raises(ValueError):
        SS.moduli_for(400)


def test_work_per_point():
    n = len(SS.moduli_for(226))
    plain, slices, moduli = SS.plain_ops(), SS.slice_ops(15), SS.moduli_ops(n)
    # the plain dd stage needs the least vector-ALU work; the moduli form moves the matrix work to fewer int8
    # products (n instead of S (S + 1) / 2) but pays for it in conversions
    assert plain < slices < moduli
    assert SS.int8_macs(n) < SS.int8_macs(15 * 16 // 2)
