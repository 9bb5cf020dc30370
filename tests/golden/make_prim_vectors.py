"""Writes prim_vectors.json: exact directed-rounding results of FP64
add / mul / div / sqrt, computed with rational arithmetic (fractions) and
integer square roots, independent of any floating-point rounding mode.

The cases cover signed zeros, subnormals, the smallest normal, DBL_MAX,
infinities, NaN, exact results, halfway points and seeded random operands
over the full exponent range.  Run:  python tests/golden/make_prim_vectors.py
"""
from __future__ import annotations

import json
import math
import os
import random
import struct
from fractions import Fraction

INF = float("inf")
NAN = float("nan")
DMAX = 1.7976931348623157e308
TINY = 5e-324
MIN_NORMAL = 2.2250738585072014e-308


def bits(x: float) -> str:
    return "%016x" % struct.unpack("<Q", struct.pack("<d", x))[0]


def round_dir(q: Fraction, up: bool) -> float:
    """The double nearest to q below (up=False) or above (up=True)."""
    if q == 0:
        return 0.0
    # float(Fraction) is correctly rounded to nearest (overflow raises)
    try:
        c = float(q)
    except OverflowError:
        c = INF if q > 0 else -INF
    if math.isinf(c):
        if up:
            return INF if q > 0 else -DMAX
        return DMAX if q > 0 else -INF
    if Fraction(c) == q:
        return c
    if up:
        return c if Fraction(c) > q else math.nextafter(c, INF)
    return c if Fraction(c) < q else math.nextafter(c, -INF)


def exact_zero_sign(op: str, a: float, b: float, up: bool) -> float:
    # IEEE 754: exact zero sums take sign + except in round-down (-); exact
    # zero products/quotients take the xor of the operand signs.
    if op == "add":
        if a == 0 and b == 0 and math.copysign(1, a) == math.copysign(1, b):
            return a
        return 0.0 if up else -0.0
    neg = (math.copysign(1, a) < 0) != (math.copysign(1, b) < 0)
    return -0.0 if neg else 0.0


def ref_binop(op: str, a: float, b: float, up: bool) -> float:
    if math.isnan(a) or math.isnan(b):
        return NAN
    if op == "add":
        if math.isinf(a) or math.isinf(b):
            if math.isinf(a) and math.isinf(b) and a != b:
                return NAN
            return a if math.isinf(a) else b
        q = Fraction(a) + Fraction(b)
    elif op == "mul":
        if math.isinf(a) or math.isinf(b):
            if a == 0 or b == 0:
                return NAN
            return INF if (a > 0) == (b > 0) else -INF
        q = Fraction(a) * Fraction(b)
    else:  # div
        if math.isinf(a) and math.isinf(b):
            return NAN
        if b == 0:
            if a == 0:
                return NAN
            neg = (math.copysign(1, a) < 0) != (math.copysign(1, b) < 0)
            return -INF if neg else INF
        if math.isinf(a):
            return INF if (a > 0) == (math.copysign(1, b) > 0) else -INF
        if math.isinf(b):
            return exact_zero_sign("div", a, b, up)
        q = Fraction(a) / Fraction(b)
    if q == 0:
        return exact_zero_sign(op, a, b, up)
    return round_dir(q, up)


def ref_sqrt(x: float, up: bool) -> float:
    if math.isnan(x) or x < 0:
        return NAN
    if x == 0 or math.isinf(x):
        return x
    fx = Fraction(x)
    c = math.sqrt(x)  # correctly rounded to nearest; adjust by exact squares
    if Fraction(c) ** 2 == fx:
        return c
    if up:
        return c if Fraction(c) ** 2 > fx else math.nextafter(c, INF)
    return c if Fraction(c) ** 2 < fx else math.nextafter(c, -INF)


def special_values():
    return [0.0, -0.0, TINY, -TINY, 3 * TINY, MIN_NORMAL, -MIN_NORMAL, math.nextafter(MIN_NORMAL, 0), 1.0, -1.0,
            2.0, 3.0, 0.1, 1e-300, 1e300, DMAX, -DMAX, INF, -INF, NAN, 7.0, 0.61, 0.25, 1.5, 2 ** -1000, 2 ** 1000,
            1.0 + 2 ** -52, 1.0 - 2 ** -53]


def main() -> None:
    rng = random.Random(20261203)
    pairs = []
    sv = special_values()
    for a in sv:
        for b in sv:
            pairs.append((a, b))
    for _ in range(3000):
        e1, e2 = rng.randint(-1074, 1023), rng.randint(-1074, 1023)
        a = rng.uniform(1, 2) * 2.0 ** max(e1, -1022) * rng.choice([1, -1])
        b = rng.uniform(1, 2) * 2.0 ** max(e2, -1022) * rng.choice([1, -1])
        if rng.random() < 0.1:
            a = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(52)))[0]  # subnormal
        pairs.append((a, b))
    for _ in range(2000):  # moderate range, the prover's working range
        a = rng.uniform(-4, 4)
        b = rng.uniform(-4, 4)
        pairs.append((a, b))
    for _ in range(500):  # exact and halfway cases
        m = rng.getrandbits(26) | 1
        a, b = float(m * m), float(m)
        pairs.append((a, b))
        pairs.append((a * 2.0 ** rng.randint(-60, 60), b))
    cases = []
    for a, b in pairs:
        row = {"a": bits(a), "b": bits(b)}
        for op in ("add", "mul", "div"):
            row[op + "_rd"] = bits(ref_binop(op, a, b, False))
            row[op + "_ru"] = bits(ref_binop(op, a, b, True))
        row["sqrt_rd"] = bits(ref_sqrt(a, False))
        row["sqrt_ru"] = bits(ref_sqrt(a, True))
        cases.append(row)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "prim_vectors.json")
    with open(out, "w") as f:
        json.dump({"generator": "tests/golden/make_prim_vectors.py", "cases": cases}, f, separators=(",", ":"))
    print(out, len(cases))


if __name__ == "__main__":
    main()
