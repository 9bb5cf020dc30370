"""Seeded rectangle_size_bound leaf cells drawn from the reference's grid.

Cells follow the reference's construction (Prover/rectangle_size_bound/
rectangle_size_bound.cu:276-310 and :171-193): lambda in [1, 2.5] split in
256, r1 in [0, 9/64] split in 128, r2 in [0, r1.ub] split in 128 with
r1.lb raised to r2.lb, r3..r6 each the k-th of 64 pieces of [0, r2.ub], and
the leaf tightened (variables.hpp:53-66).  Piece boundaries use the fused
lb + diff*i of get_subinterval (subintervals.hpp:32-43).
"""
import math

import numpy as np

W = 9.0 / 64.0


def piece(lo, hi, i, n):
    d = (hi - lo) / n
    l = math.fma(d, i, lo) if hasattr(math, "fma") else _fma(d, i, lo)
    u = hi if i + 1 == n else (math.fma(d, i + 1, lo) if hasattr(math, "fma") else _fma(d, i + 1, lo))
    return l, u


def _fma(a, b, c):
    from fractions import Fraction

    return float(Fraction(a) * Fraction(b) + Fraction(c))


def tighten(r):
    lb = 0.0
    for i in range(5, 0, -1):
        if lb < r[i][0]:
            lb = r[i][0]
        else:
            r[i] = (lb, r[i][1])
    return r


def leaf_cells(seed: int, n: int, la_range=(0, 256), r_bias=None) -> np.ndarray:
    """n leaf cells; `la_range` limits the lambda piece indices and `r_bias`
    (a fraction in (0,1]) limits radius piece indices to the top fraction,
    which concentrates samples near the large-disk, hard-to-prove corner."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, 14))
    for k in range(n):
        li = int(rng.integers(*la_range))
        la = piece(1.0, 2.5, li, 256)
        lo_i = 0 if r_bias is None else int(128 * (1 - r_bias))
        i1 = int(rng.integers(lo_i, 128))
        r1 = piece(0.0, W, i1, 128)
        i2 = int(rng.integers(lo_i, 128))
        r2 = piece(0.0, r1[1], i2, 128)
        r1 = (max(r1[0], r2[0]), r1[1])
        rad = [r1, r2]
        lo_j = 0 if r_bias is None else int(64 * (1 - r_bias))
        js = sorted((int(x) for x in rng.integers(lo_j, 64, 4)), reverse=True)  # r3 >= r4 >= r5 >= r6
        for j in js:
            rad.append(piece(0.0, r2[1], j, 64))
        rad = tighten(rad)
        out[k, 0:2] = la
        for j in range(6):
            out[k, 2 + 2 * j: 4 + 2 * j] = rad[j]
    return out
