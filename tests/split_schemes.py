"""CPU emulation of two ways to run the radix-16 DFT stage of the split-precision form on int8 matrix cores, exact in
Python integers, with a count of the vector-ALU work each needs per complex point beside the plain dd stage.

The 16-point complex DFT of a dd vector is the real 32 x 32 product y = M x (M the real form of F = e^{-2 pi i k n/16}).
Both schemes scale x by a power of two and round it to a B-bit integer vector X, and round M * 2^Bf to integers.

  slices (Ozaki scheme I, what the GPU runs): X and M are cut into S base-128 digits; the digit products
      with i + j <= S + 1 are int8 matrix products with exact int32 sums, recombined in FP64 by weight.
  moduli (Ozaki scheme II, Ozaki, Uchino and Imamura 2025): X and M are reduced modulo pairwise coprime
      p_k <= 255 (balanced residues fit int8; moduli up to 127 alone cannot exceed about 2^188, less than the
      2^226 a 2^-100 result needs); one int8 matrix product per modulus gives Y mod p_k exactly
      (|sum| <= 32 * 127 * 127 < 2^31), and the Chinese remainder theorem gives the exact integer Y = M X once
      prod p_k > 2 |Y|max. Y is then scaled back.

The emulation here runs the moduli scheme exactly (the slice scheme is the GPU's, tested against a 256-bit DFT by
test_gpu_split.py). Both give the same product when exact; the difference is the work around the matrix cores. The counts
below are vector-ALU operations (an FMA is one) per complex point of a 16-point vector, for the conversions each
scheme needs on the GPU; the int8 matrix work itself is counted separately and is cheap on MI355X. Run this file to
print the table. tests/test_split_schemes.py pins the emulation against a 256-bit DFT (mpmath).
"""
import math
from fractions import Fraction

import mpmath
import numpy as np

N = 16
K = 2 * N          # inner dimension of the real product


def dft_matrix(bf):
    """round(M * 2^bf) as Python ints: rows (Re y[0..15], Im y[0..15]), columns (Re x[0], Im x[0], Re x[1], ...)."""
    mpmath.mp.prec = 320
    m = [[0] * K for _ in range(K)]
    for r in range(K):
        k = r % N
        for c in range(K):
            n = c // 2
            ang = -2 * mpmath.pi * k * n / N
            cr, ci = mpmath.cos(ang), mpmath.sin(ang)
            row_re, col_re = r < N, c % 2 == 0
            v = cr if row_re == col_re else (-ci if row_re else ci)
            m[r][c] = int(mpmath.nint(v * mpmath.mpf(2) ** bf))
    return m


def to_int_vector(x, b):
    """x: (16, 4) doubles {re.hi, im.hi, re.lo, im.lo}. Returns (X, e): X the 32 ints round(x * 2^(b - e)) in the
    real form's order, 2^e > max |component| (dd values taken exactly)."""
    vals = []
    for n in range(N):
        vals.append(Fraction(x[n, 0]) + Fraction(x[n, 2]))
        vals.append(Fraction(x[n, 1]) + Fraction(x[n, 3]))
    big = max(abs(v) for v in vals)
    e = 0 if big == 0 else math.floor(math.log2(big)) + 1
    scale = Fraction(2) ** (b - e)
    return [round(v * scale) for v in vals], e


def moduli_for(bits):
    """Pairwise coprime moduli <= 255, largest first, until their product exceeds 2^bits."""
    chosen, prod = [], 1
    for p in range(255, 2, -1):
        if all(math.gcd(p, q) == 1 for q in chosen):
            chosen.append(p)
            prod *= p
            if prod.bit_length() > bits:
                return chosen
    raise ValueError("not enough moduli")


def balanced(v, p):
    r = v % p
    return r - p if r > p // 2 else r


def crt_product(m, X, moduli):
    """Y = M X through the residues: per modulus one int8 x int8 -> int32 product, then Garner's reconstruction."""
    P = math.prod(moduli)
    res = []
    for p in moduli:
        mp = [[balanced(v, p) for v in row] for row in m]
        xp = [balanced(v, p) for v in X]
        assert all(-128 <= v <= 127 for row in mp for v in row) and all(-128 <= v <= 127 for v in xp)
        acc = [sum(a * b for a, b in zip(row, xp)) for row in mp]
        assert all(abs(v) < 2 ** 31 for v in acc)   # exact in the int32 accumulator
        res.append([a % p for a in acc])
    Y = []
    for i in range(K):
        # Garner: mixed-radix digits d_k, Y = d_0 + d_1 p_0 + d_2 p_0 p_1 + ...
        digits, prefix = [], []
        for k, p in enumerate(moduli):
            v = res[k][i]
            for j, d in enumerate(digits):
                v = (v - d * prefix[j]) % p
            inv = pow(math.prod(moduli[:k]) % p, -1, p) if k else 1
            digits.append(v * inv % p)
            prefix.append(math.prod(moduli[:k]))
        y = sum(d * pr for d, pr in zip(digits, prefix))
        Y.append(y - P if y > P // 2 else y)
    return Y


def dft_error(x, y, scale_exp):
    """Normwise error of y * 2^scale_exp (32 ints, real form) against the 256-bit DFT of the dd vector x."""
    mpmath.mp.prec = 256
    w = [mpmath.expjpi(-mpmath.mpf(2 * k) / N) for k in range(N)]
    xs = [mpmath.mpc(mpmath.mpf(x[n, 0]) + mpmath.mpf(x[n, 2]), mpmath.mpf(x[n, 1]) + mpmath.mpf(x[n, 3]))
          for n in range(N)]
    ref = [mpmath.fsum(xs[n] * w[(k * n) % N] for n in range(N)) for k in range(N)]
    got = [mpmath.mpc(mpmath.ldexp(y[k], scale_exp), mpmath.ldexp(y[N + k], scale_exp)) for k in range(N)]
    norm = max(abs(r) for r in ref)
    return float(max(abs(g - r) for g, r in zip(got, ref)) / norm)


# ---- vector-ALU work per complex point (an FMA counts one), as a GPU kernel would need it -------------------------
def plain_ops():
    """Two radix-4 dd Stockham stages (the plain form's radix-16 work): per 4 points 8 one-TwoSum complex adds
    (16 dd adds of 11 operations) and, on 3 of 4 butterflies of the first stage, 3 dd complex products (4 products
    of 3 multiplies, one FMA and 5 adds, plus 2 dd adds)."""
    add, mul = 11, 9
    cmul = 4 * mul + 2 * add
    per_bfly = 16 * add
    stage0 = 4 * per_bfly + 3 * 3 * cmul    # 4 butterflies, 3 twiddled
    stage1 = 4 * per_bfly
    return (stage0 + stage1) / N


def slice_ops(s):
    """Per input real: s digits peeled in FP64 (t = 128 h, d = rint(t), h = t - d: 3 operations each) plus the dd
    split and scaling (about 6); per output real: s + 1 weight sums converted (1) and added into a dd accumulator by
    TwoSum (6 with the renormalisation) each; the stage's dd twiddle product on 15 of 16 outputs."""
    per_in = 6 + 3 * s
    per_out = (s + 1) * 7
    twiddle = 15 / 16 * (4 * 9 + 2 * 11)
    return 2 * per_in + 2 * per_out + twiddle


def moduli_ops(n):
    """Per input real: the 106-bit integer as three 36-bit FP64 chunks (about 8 operations), then per modulus
    c2 (2^72 mod p) + c1 (2^36 mod p) + c0 (2 FMAs), the reduction q = floor(v / p), r = v - q p (3) and the balanced
    int8 residue packed into the operand (3): 8 per modulus. Per output real: the CRT sum of n residues times their
    weights, each weight a 4-word FP64 expansion of an integer below prod p (~2^220), accumulated error-free (2
    exact products and 4 additions per word), then the reduction modulo prod p and the conversion to dd (about 40):
    about 24 per modulus + 40 (a low estimate: 51-bit word products summed over n moduli are not exact in FP64 and
    need TwoSum accumulation). Plus the twiddle product."""
    per_in = 8 + 8 * n
    per_out = 24 * n + 40
    twiddle = 15 / 16 * (4 * 9 + 2 * 11)
    return 2 * per_in + 2 * per_out + twiddle


def int8_macs(products):
    """int8 multiply-accumulates per complex point: one 32 x 32 real matrix product per slice pair or modulus."""
    return products * K * K / N


if __name__ == "__main__":
    b = bf = 110
    bound = b + bf + 6
    mods = moduli_for(bound)
    s = 15
    print(f"plain dd radix-16 stage: {plain_ops():.0f} VALU operations per point")
    print(f"slices S={s}: {slice_ops(s):.0f} VALU per point, {s * (s + 1) // 2} digit products, "
          f"{int8_macs(s * (s + 1) // 2):.0f} int8 MACs per point")
    print(f"moduli n={len(mods)} (prod > 2^{bound}): {moduli_ops(len(mods)):.0f} VALU per point, "
          f"{int8_macs(len(mods)):.0f} int8 MACs per point")
