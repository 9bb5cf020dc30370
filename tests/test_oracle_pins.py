"""Pins of the CPU oracle that do not come from the oracle itself.

Each check recomputes a quantity from mathematics the reference's code only
approximates (exact rationals via fractions.Fraction, 60-digit mpmath), so a
dropped term, a wrong constant, a swapped operand or a wrong rounding
direction in the oracle fails here.  Sources of the formulas:
- interval ops: Prover/algcuda/cuda-utils/cuh/include/algcuda/interval.cuh
  (sum, difference, non-branching product, sign-case quotient, reciprocal,
  square, sqrt) and the scalar-product quirk at :186-188;
- critical_ratio: Prover/tight_rectangle/critical_ratio.cu:32-78;
- two_disks_maximize_height: Prover/rectangle_size_bound/maximize_covered.cu:106-130;
- disk_satisfies_size_bound, bound_required_height:
  Prover/rectangle_size_bound/disk_satisfies_size_bound.cu:25-37;
- circle intersection: Prover/operations.cu:30-83.
"""
import math
import random
from fractions import Fraction as F

import mpmath
import pytest

import oracle

mpmath.mp.dps = 60


def rd(x: F) -> float:
    """Largest double <= x."""
    f = float(x)
    if F(f) > x:
        f = math.nextafter(f, -math.inf)
    return f


def ru(x: F) -> float:
    f = float(x)
    if F(f) < x:
        f = math.nextafter(f, math.inf)
    return f


def rand_iv(rng, lo=-4.0, hi=4.0):
    a, b = sorted(rng.uniform(lo, hi) for _ in range(2))
    if rng.random() < 0.1:
        b = a
    return a, b


# ------------------------------------------------------------ interval ops
def test_iv_add_sub_exact_rounding():
    rng = random.Random(5)
    for _ in range(3000):
        a, b = rand_iv(rng), rand_iv(rng)
        (lo, hi), _ = oracle.iv_op(0, a, b)
        assert lo == rd(F(a[0]) + F(b[0])) and hi == ru(F(a[1]) + F(b[1]))
        (lo, hi), _ = oracle.iv_op(1, a, b)
        assert lo == rd(F(a[0]) - F(b[1])) and hi == ru(F(a[1]) - F(b[0]))


def test_iv_mul_is_hull_of_corner_products():
    rng = random.Random(6)
    for _ in range(3000):
        a, b = rand_iv(rng), rand_iv(rng)
        (lo, hi), _ = oracle.iv_op(2, a, b)
        prods = [F(x) * F(y) for x in a for y in b]
        assert lo == rd(min(prods)) and hi == ru(max(prods))


def test_iv_scalar_mul_quirk_rounds_upper_bound_down():
    """IV * n with n >= 0 rounds both bounds down (interval.cuh:186-188);
    n < 0 swaps and rounds outward."""
    rng = random.Random(7)
    seen_inward = False
    for _ in range(3000):
        a = rand_iv(rng, 0.1, 3.0)
        n = rng.choice([0.61, 1 / 3, 0.1, 2.0, 0.5, -0.61, -0.3])
        (lo, hi), _ = oracle.iv_op(4, a, (n, n))
        if n >= 0:
            assert lo == rd(F(a[0]) * F(n)) and hi == rd(F(a[1]) * F(n))
            seen_inward |= F(hi) < F(a[1]) * F(n)
        else:
            assert lo == rd(F(a[1]) * F(n)) and hi == ru(F(a[0]) * F(n))
    assert seen_inward  # the quirk is visible: the upper bound is below the exact product


def test_iv_div_sign_cases():
    rng = random.Random(8)
    for _ in range(3000):
        a = rand_iv(rng)
        b = rand_iv(rng, 0.2, 3.0)
        if rng.random() < 0.5:
            b = (-b[1], -b[0])
        (lo, hi), _ = oracle.iv_op(3, a, b)
        q = [F(x) / F(y) for x in a for y in b]
        assert lo == rd(min(q)) and hi == ru(max(q))
    (lo, hi), _ = oracle.iv_op(3, (1.0, 2.0), (-1.0, 1.0))
    assert lo == -math.inf and hi == math.inf


def test_iv_reciprocal_square_sqrt():
    rng = random.Random(9)
    for _ in range(2000):
        a = rand_iv(rng, 0.05, 9.0)
        (lo, hi), _ = oracle.iv_op(7, a)
        assert lo == rd(1 / F(a[1])) and hi == ru(1 / F(a[0]))
        # negative intervals: the reference returns [RD(1/lb), RU(1/ub)]
        # (interval.cuh:231), which is not an enclosure; every call site
        # passes a positive interval, so the quirk is kept as written
        (lo, hi), _ = oracle.iv_op(7, (-a[1], -a[0]))
        assert lo == rd(-1 / F(a[1])) and hi == ru(-1 / F(a[0]))
        s = rand_iv(rng)
        (lo, hi), _ = oracle.iv_op(5, s)
        m_lo = F(0) if s[0] <= 0 <= s[1] else min(F(s[0]) ** 2, F(s[1]) ** 2)
        m_hi = max(F(s[0]) ** 2, F(s[1]) ** 2)
        assert lo == rd(m_lo) and hi == ru(m_hi)
        (lo, hi), trap = oracle.iv_op(6, a)
        assert not trap
        assert F(lo) ** 2 <= F(a[0]) < F(math.nextafter(lo, math.inf)) ** 2
        assert F(math.nextafter(hi, -math.inf)) ** 2 < F(a[1]) <= F(hi) ** 2
    _, trap = oracle.iv_op(6, (-1e-300, 1.0))
    assert trap  # the reference traps the grid here; the oracle flags it


# ---------------------------------------------------------- critical_ratio
LAMBDA2 = mpmath.sqrt(mpmath.sqrt(7) / 2 - mpmath.mpf(1) / 4)


def cr_exact(la):
    la = mpmath.mpf(la)
    if la <= LAMBDA2:
        return mpmath.mpf(3) / 256 / la * (16 * la ** 2 + 40 + 9 / la ** 2)
    return la / 4 + 1 / (2 * la)


def test_break_even_lambda_is_continuous():
    a = mpmath.mpf(3) / 256 / LAMBDA2 * (16 * LAMBDA2 ** 2 + 40 + 9 / LAMBDA2 ** 2)
    b = LAMBDA2 / 4 + 1 / (2 * LAMBDA2)
    assert abs(a - b) < mpmath.mpf(10) ** -50
    assert abs(float(LAMBDA2) - 1.0357971111816713) < 1e-15


def test_critical_ratio_at_one_contains_195_over_256():
    lo, hi = oracle.tr_critical_ratio(1.0, 1.0)
    assert F(lo) <= F(195, 256) <= F(hi)
    assert hi - lo < 1e-14


@pytest.mark.parametrize("seed", [1, 2])
def test_critical_ratio_encloses_closed_form(seed):
    rng = random.Random(seed)
    for _ in range(400):
        la = rng.uniform(1.0, 8.0)
        lo, hi = oracle.tr_critical_ratio(la, la)
        v = cr_exact(la)
        assert mpmath.mpf(lo) <= v <= mpmath.mpf(hi), la
        assert hi - lo <= 64 * math.ulp(float(v)), (la, lo, hi)
    # straddling the break-even point: the hull of both branches
    l2 = float(LAMBDA2)
    lo, hi = oracle.tr_critical_ratio(l2 - 1e-9, l2 + 1e-9)
    for x in (l2 - 1e-9, l2, l2 + 1e-9):
        assert mpmath.mpf(lo) <= cr_exact(x) <= mpmath.mpf(hi)


def test_critical_ratio_wide_interval_encloses_range():
    lo, hi = oracle.tr_critical_ratio(1.0, 2.0)
    for k in range(0, 101):
        x = 1.0 + k / 100
        assert mpmath.mpf(lo) <= cr_exact(x) <= mpmath.mpf(hi)


# ------------------------------------------------- rectangle_size_bound helpers
def test_two_disks_maximize_height_equal_radii():
    """r1 = r2 = r (squared radii): the strip height is sqrt(4r - w^2/4)."""
    rng = random.Random(3)
    hits = 0
    for _ in range(2000):
        r = rng.uniform(0.01, 0.6)
        w = rng.uniform(0.05, 2.0)
        h = oracle.lib.or_two_disks_maximize_height(r, r, w)
        arg = 4 * mpmath.mpf(r) - mpmath.mpf(w) ** 2 / 4
        if arg <= 0:
            assert h == 0.0
            continue
        exact = mpmath.sqrt(arg)
        if h > 0:
            hits += 1
            assert mpmath.mpf(h) <= exact
            assert exact - mpmath.mpf(h) <= 1e-12 * exact + 1e-300
    assert hits > 500


def test_two_disks_maximize_height_unequal_radii():
    rng = random.Random(4)
    hits = 0
    for _ in range(2000):
        r1 = rng.uniform(0.05, 0.6)
        r2 = rng.uniform(0.02, r1)
        w = rng.uniform(0.3, 2.0)
        h = oracle.lib.or_two_disks_maximize_height(r1, r2, w)
        R1, R2, W = mpmath.mpf(r1), mpmath.mpf(r2), mpmath.mpf(w)
        arg = 2 * (R1 + R2) - W ** 2 / 4 - 4 * (R1 - R2) ** 2 / W ** 2
        if h > 0:
            hits += 1
            assert arg > 0
            exact = mpmath.sqrt(arg)
            # RU(w^2) sits in the denominator of the subtracted term
            # (maximize_covered.cu:124), so the reference's "lower bound"
            # may exceed the exact height by a few ulps: closeness only
            assert abs(exact - mpmath.mpf(h)) <= 1e-11
    assert hits > 300


def test_disk_satisfies_size_bound_and_required_height():
    rng = random.Random(10)
    both = [0, 0]
    for _ in range(5000):
        r = rng.uniform(1e-4, 0.6)
        h = rng.uniform(0.05, 2.0)
        ok = bool(oracle.lib.or_disk_satisfies_size_bound(r, h))
        exact = F(r) / (F(h) ** 2)
        if ok:
            assert exact <= F(9, 64)
        elif exact <= F(9, 64) * (1 - F(1, 10 ** 12)):
            raise AssertionError((r, h))
        both[ok] += 1
        need = oracle.lib.or_bound_required_height(r)
        ex = mpmath.sqrt(mpmath.mpf(r)) * 8 / 3
        assert mpmath.mpf(need) >= ex and mpmath.mpf(need) - ex <= 4 * math.ulp(need)
    assert min(both) > 500


# ------------------------------------------------------------- geometry
def test_intersection_points_enclose_exact_points():
    rng = random.Random(11)
    checked = 0
    for _ in range(1500):
        c1 = (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.05, 1.0))
        c2 = (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.05, 1.0))
        ok, pts = oracle.intersection(c1, c2)
        x1, y1, q1 = map(mpmath.mpf, c1)
        x2, y2, q2 = map(mpmath.mpf, c2)
        d2 = (x2 - x1) ** 2 + (y2 - y1) ** 2
        a = (q1 - q2 + d2) / (2 * mpmath.sqrt(d2))
        hh = q1 - a * a
        if not ok:
            assert hh < 1e-9 * q1  # only near-tangent or disjoint pairs are refused
            continue
        assert hh > 0
        d = mpmath.sqrt(d2)
        mx, my = x1 + a * (x2 - x1) / d, y1 + a * (y2 - y1) / d
        h = mpmath.sqrt(hh)
        ex = [(mx + h * (y2 - y1) / d, my - h * (x2 - x1) / d), (mx - h * (y2 - y1) / d, my + h * (x2 - x1) / d)]
        # order rule of the reference: the first point is the '+' one when
        # ydiff.ub < 0 or (ydiff.ub == 0 and xdiff.lb >= 0)
        dy, dx = c2[1] - c1[1], c2[0] - c1[0]
        if not (dy < 0 or (dy == 0 and dx >= 0)):
            ex.reverse()
        for (px, py), ((xl, xh), (yl, yh)) in zip(ex, pts):
            assert xl <= px <= xh and yl <= py <= yh
            assert xh - xl < 1e-9 and yh - yl < 1e-9
        checked += 1
    assert checked > 500
