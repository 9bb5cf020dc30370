"""Directed-rounding primitives: the CPU oracle against exact golden vectors
(CPU), and the gfx950 primitives against the golden vectors and the oracle
(GPU).  The golden vectors come from rational arithmetic
(tests/golden/make_prim_vectors.py), so they pin IEEE-754 round-down /
round-up semantics -- what the reference's __dadd_rd ... __dsqrt_ru compute
(Prover/algcuda/cuda-utils/cuh/include/algcuda/interval.cuh:30-355)."""
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "prim_vectors.json")
COLS = ["add_rd", "add_ru", "mul_rd", "mul_ru", "div_rd", "div_ru", "sqrt_rd", "sqrt_ru"]


def f64(h: str) -> float:
    return struct.unpack("<d", struct.pack("<Q", int(h, 16)))[0]


def load_gold():
    cases = json.load(open(GOLD))["cases"]
    a = np.array([f64(c["a"]) for c in cases])
    b = np.array([f64(c["b"]) for c in cases])
    exp = np.array([[f64(c[k]) for k in COLS] for c in cases])
    return a, b, exp


def same_bits(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Bitwise equality, any NaN equal to any NaN."""
    xb, yb = x.view(np.uint64), y.view(np.uint64)
    return (xb == yb) | (np.isnan(x) & np.isnan(y))


def oracle_table(a, b):
    out = np.empty((a.size, 8))
    for i, (x, y) in enumerate(zip(a.tolist(), b.tolist())):
        out[i] = [oracle.prim("add", "rd", x, y), oracle.prim("add", "ru", x, y),
                  oracle.prim("mul", "rd", x, y), oracle.prim("mul", "ru", x, y),
                  oracle.prim("div", "rd", x, y), oracle.prim("div", "ru", x, y),
                  oracle.prim("sqrt", "rd", x), oracle.prim("sqrt", "ru", x)]
    return out


def test_golden_vectors_are_self_consistent():
    a, b, exp = load_gold()
    # round-down never exceeds round-up; both bracket the RN result
    for k in (0, 2, 4, 6):
        lo, hi = exp[:, k], exp[:, k + 1]
        ok = np.isnan(lo) | (lo <= hi)
        assert ok.all()
    # spot identities that a transposed operand or wrong direction would break
    assert f64(json.load(open(GOLD))["cases"][0]["add_rd"]) == 0.0


def test_oracle_matches_golden():
    a, b, exp = load_gold()
    got = oracle_table(a, b)
    bad = ~same_bits(got, exp)
    assert not bad.any(), f"{bad.sum()} mismatches, first at {np.argwhere(bad)[:5].tolist()}"


def test_oracle_directed_ordering_random():
    rng = np.random.default_rng(7)
    a = rng.normal(size=2000) * 10.0 ** rng.integers(-30, 30, 2000)
    b = rng.normal(size=2000) * 10.0 ** rng.integers(-30, 30, 2000)
    t = oracle_table(a, b)
    for k in (0, 2, 4):
        lo, hi = t[:, k], t[:, k + 1]
        gap = np.abs(hi - lo)
        # at most one ulp apart: RD and RU are adjacent or equal
        assert (hi >= lo).all() and (gap <= np.spacing(np.maximum(np.abs(lo), np.abs(hi))) * 1.0000001).all()


@pytest.mark.gpu
def test_gpu_primitives_match_golden(prover):
    a, b, exp = load_gold()
    got = prover.prim_selftest(a, b)
    bad = ~same_bits(got, exp)
    assert not bad.any(), f"{bad.sum()} mismatches: " + str(
        [(a[i], b[i], COLS[k], got[i, k], exp[i, k]) for i, k in np.argwhere(bad)[:5]])


@pytest.mark.gpu
def test_gpu_primitives_match_oracle_random(prover):
    rng = np.random.default_rng(12345)
    n = 200_000
    e = rng.integers(-1074, 1024, size=(2, n))
    a = rng.uniform(1, 2, n) * np.exp2(np.maximum(e[0], -1022).astype(float)) * rng.choice([-1, 1], n)
    b = rng.uniform(1, 2, n) * np.exp2(np.maximum(e[1], -1022).astype(float)) * rng.choice([-1, 1], n)
    # half of the draws in the prover's working range
    a[: n // 2] = rng.uniform(-3, 3, n // 2)
    b[: n // 2] = rng.uniform(-3, 3, n // 2)
    got = prover.prim_selftest(a, b)
    m = 20_000
    idx = rng.choice(n, m, replace=False)
    exp = oracle_table(a[idx], b[idx])
    bad = ~same_bits(got[idx], exp)
    assert not bad.any(), f"{bad.sum()} mismatches"
    # whole batch: round-down <= round-up for every op
    for k in (0, 2, 4, 6):
        lo, hi = got[:, k], got[:, k + 1]
        assert (np.isnan(lo) | (lo <= hi)).all()


@pytest.mark.gpu
def test_gpu_div_sqrt_near_exact_results(prover):
    """Operands whose quotient or root lies within a few ulps of a double (or
    on one): the cases where a directed div/sqrt must step its Newton result
    by one ulp.  Near squares r*r +- k ulp, exact squares, near-exact
    quotients (q*b rounded, +- k ulp) and exact ones, both signs."""
    rng = np.random.default_rng(777)
    n = 6000
    r = rng.uniform(1, 2, n) * np.exp2(rng.integers(-200, 200, n).astype(float))
    sq = r * r
    k = rng.integers(-3, 4, n)
    sq = (sq.view(np.int64) + k).view(np.float64)
    q = rng.uniform(1, 2, n) * np.exp2(rng.integers(-100, 100, n).astype(float))
    b = rng.uniform(1, 2, n) * np.exp2(rng.integers(-100, 100, n).astype(float))
    b[: n // 3] = rng.integers(1, 1 << 20, n // 3).astype(float)   # exact small quotients below
    q[: n // 3] = rng.integers(1, 1 << 30, n // 3).astype(float)
    a = q * b
    a = (a.view(np.int64) + rng.integers(-2, 3, n)).view(np.float64)
    sgn_a, sgn_b = rng.choice([-1.0, 1.0], n), rng.choice([-1.0, 1.0], n)
    a, b = a * sgn_a, b * sgn_b
    # columns 0..5 use (a, b); the sqrt columns use a: run the roots as a separate batch
    got_div = prover.prim_selftest(a, b)
    got_sqrt = prover.prim_selftest(sq, b)
    exp_div = oracle_table(a, b)
    exp_sqrt = oracle_table(sq, b)
    bad = ~same_bits(got_div[:, 4:6], exp_div[:, 4:6])
    assert not bad.any(), f"{bad.sum()} division mismatches"
    bad = ~same_bits(got_sqrt[:, 6:8], exp_sqrt[:, 6:8])
    assert not bad.any(), f"{bad.sum()} sqrt mismatches"


@pytest.mark.gpu
def test_gpu_round_nearest_helpers(prover):
    """Bisection midpoints, fused get_subinterval steps and the subinterval
    width must be the round-to-nearest results (checked against the oracle)."""
    rng = np.random.default_rng(99)
    n = 5000
    a = rng.uniform(0, 3, n) * np.exp2(rng.integers(-40, 40, n).astype(float))
    b = rng.uniform(0, 3, n) * np.exp2(rng.integers(-40, 40, n).astype(float))
    c = rng.uniform(0.5, 3, n) * np.exp2(rng.integers(-5, 5, n).astype(float))
    a[:100], b[:100] = 0.3, np.nextafter(0.3, 1.0)  # adjacent midpoint cases
    got = prover.prim_rn_selftest(a, b, c)
    exp = np.array([oracle.rn_helpers(x, y, z) for x, y, z in zip(a.tolist(), b.tolist(), c.tolist())])
    bad = ~same_bits(got, exp)
    assert not bad.any(), f"{bad.sum()} mismatches: {[(a[i], b[i], c[i], k, got[i, k], exp[i, k]) for i, k in np.argwhere(bad)[:5]]}"
    # numpy computes in round-to-nearest: 0.5*(a+b) and (a-b)/c agree
    assert same_bits(got[:, 0], 0.5 * (a + b)).all()
    assert same_bits(got[:, 2], (a - b) / c).all()
