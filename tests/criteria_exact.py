"""The ten criteria of the 4-D box set, the rectangle_size_bound leaf
steps C1..C13 and the tight_rectangle leaf criteria, decided in exact real
arithmetic (test support, independent of oracle/ and of the kernels).

Each function restates the mathematics of one reference criterion on a
point cell (lambda, squared radii r1 >= r2 >= r3): the same placements,
comparisons and constants (the doubles the reference uses, e.g. its 0.61,
which is below 61/100), and the reference's own quirks, but no rounding:
every quantity is an mpmath number at 60 significant digits.  A Dec object
records the smallest relative slack |a - b| / max(|a|, |b|) of every
comparison the evaluation decides, so a caller knows how far the verdict is
from flipping.  Where that slack is large, the outward-rounded FP64
criterion (a few hundred operations, each off by at most one ulp) must reach
the same verdict; a dropped term, a swapped operand or a wrong constant in a
transcription moves a quantity by far more than rounding does and flips
verdicts at clear margins.

Reference files (under Prover/): tight_rectangle/critical_ratio.cu:32-86,
can_recurse.cu:32-62, can_apply_size_bounded_covering.cu:49-52,
r1_strategies.cu:27-87, r1_r2_strategies.cu:29-134,
r1_r2_r3_strategies.cu:28-94; rectangle_size_bound/values.hpp:28-35,
disk_satisfies_size_bound.cu:25-37, maximize_covered.cu:106-130,
can_recurse.cu:26-88, 221-235, even_split_recursion.cu:42-57,
uneven_split_recursion.cu:78-125, r1_r2_opposite_corners_recursion.cu:28-71,
r1_r2_vertical_strip.cu:28-41; for the leaf steps variables.cuh:30-52,
compute_even_split_distance_bound.cu:25-76, even_split_recursion.cu:25-40,
uneven_split_recursion.cu:25-76, vertical_wall_building_recursion.cu:25-78,
r1_r2_vertical_strip.cu:43-95, multi_disk_strip.cu:31-262,
r1_in_corner_explicit_recursion.cu:35-262, maximize_covered.cu:28-104,
171-237; operations.cu:30-83; for the tight_rectangle leaf criteria
tight_rectangle.cu:217-249, two_by_two_strategies.cu:29-115,
r1_r2_r3_strategies.cu:101-260, r1_r2_large_r3_r4_gaps_strategy.cu:29-98.
"""
from __future__ import annotations

import mpmath
from mpmath import mpf

mpmath.mp.dps = 60  # 60 significant digits

INF = mpmath.inf
CR = mpf(0.61)                       # rectangle_size_bound::critical_ratio, the double
UB_DISK_WEIGHT = mpf(9) / 64         # (3/8)^2
UB_DISK_RADIUS = mpf(3) / 8
C_WORST = mpf(195) / 256             # worst-case ratio floor
TOL_LAMBDA_RSB = mpf(2.0898841580413813901)  # rectangle_size_bound/can_recurse.cu:28 (the double)
LAMBDA_2 = mpmath.sqrt(mpmath.sqrt(7) / 2 - mpf(1) / 4)  # break-even point of critical_ratio


class Dec:
    """Comparisons of one evaluation, with the smallest relative slack seen."""

    def __init__(self):
        self.slack = INF

    def cmp(self, a, op: str, b) -> bool:
        a, b = mpf(a), mpf(b)
        scale = max(abs(a), abs(b))
        self.slack = min(self.slack, abs(a - b) / scale if scale else mpf(0))
        d = a - b
        return {"<": d < 0, "<=": d <= 0, ">": d > 0, ">=": d >= 0}[op]


# ---------------------------------------------------------------- helpers

def tr_critical_ratio(la):
    """critical_ratio(IV la) on a point (critical_ratio.cu:58-78); the two
    forms agree at the break-even point, so the branch needs no slack."""
    if la <= LAMBDA_2:
        return (mpf(3) / 256) / la * (16 * la * la + 40 + 9 / (la * la))
    return la / 4 + 1 / (2 * la)


def tr_critical_ratio_wh(w, h):
    """critical_ratio(IV w, IV h) (critical_ratio.cu:32-52): the ratio of the
    longer side over the shorter, as the reference orients it."""
    if w <= h:
        return tr_critical_ratio(h / w)
    return tr_critical_ratio(w / h)


def size_bound_ok(D, r_sq, h):
    """disk_satisfies_size_bound: r^2 (1/h)^2 <= 9/64 (disk_satisfies_size_bound.cu:25-28)."""
    if h == 0:
        return False
    return D.cmp(r_sq / (h * h), "<=", UB_DISK_WEIGHT)


def tr_can_recurse(D, w, h, largest, rem):
    """tight_rectangle can_recurse(ub_w, ub_h, ub_largest, lb_rem) (can_recurse.cu:32-62):
    the worst-case ratio of the w x h rectangle, or the corollary's ratio
    for its largest disk if lower (at least 195/256), or 0.61 when the
    largest disk, scaled to the shorter side, meets the size bound."""
    cr = tr_critical_ratio_wh(w, h)
    s = min(w, h)
    norm = largest / (s * s)
    cor = mpmath.sqrt(mpmath.sqrt(norm * norm + 1) + 1) / 2
    if D.cmp(cor, "<", C_WORST):
        cor = C_WORST
    if D.cmp(cr, ">", cor):
        cr = cor
    if D.cmp(norm, "<=", UB_DISK_WEIGHT):
        cr = CR
    return D.cmp(rem, ">=", cr * w * h)


def two_disks_maximize_height(D, r1, r2, w):
    """Height of the strip of width w that two disks stacked in it cover
    (maximize_covered.cu:106-130), 0 if a disk cannot span its part."""
    d = r1 - r2
    w1 = (w / 2 + (2 / w) * d) ** 2
    w2 = (w / 2 - (2 / w) * d) ** 2
    if D.cmp(r1, "<=", w1 / 4) or D.cmp(r2, "<=", w2 / 4):
        return mpf(0)
    a, b = 2 * (r1 + r2), w * w / 4 + 4 * d * d / (w * w)
    if D.cmp(a, "<=", b):
        return mpf(0)
    return mpmath.sqrt(a - b)


def rsb_can_recurse(D, rem, largest, w, h):
    """rectangle_size_bound can_recurse (can_recurse.cu:26-88)."""
    if D.cmp(rem, "<=", 0):
        if rem < 0:
            return False
        return w <= 0 or h <= 0
    if w <= 0 or h <= 0:
        return True
    smin, smax = min(w, h), max(w, h)
    area = rem / CR
    side = area / smax
    if D.cmp(side, "<", smin):
        return False
    if D.cmp(side, ">", smax):
        side = mpmath.sqrt(area)
    if D.cmp(UB_DISK_WEIGHT * side * side, ">=", largest):
        return True
    skew = smax / smin
    if D.cmp(skew, "<=", TOL_LAMBDA_RSB):
        return D.cmp(tr_critical_ratio(skew) * w * h, "<=", rem)
    sig = (largest / (smin * smin)) ** 2
    cr = mpmath.sqrt(mpmath.sqrt(sig + 1) + 1) / 2
    if D.cmp(cr, "<", C_WORST):
        cr = C_WORST
    if D.cmp(cr * w * h, "<=", rem):
        return True
    return D.cmp(smax * smax / 4 + smin * smin / 2, "<=", rem)


def bound_worst_case_ratio(D, w, h):
    """rectangle_size_bound bound_worst_case_ratio (can_recurse.cu:221-235)."""
    if w <= 0 or h <= 0:
        return INF
    skew = max(w / h, h / w)
    r = skew / 4 + 1 / (2 * skew)
    return C_WORST if D.cmp(r, "<", C_WORST) else r


# ------------------------------------------------------- tight_rectangle

def r1_strategies(D, la, r1, ub_r2):
    """Criterion 0 (r1_strategies.cu:27-87).  Returns (verdict, strategy):
    1 full-height strip of r1 plus recursion, 2 r1's square in a corner with
    the right strip and the top rectangle, 3 the pocket placement."""
    rem = tr_critical_ratio(la) * la - r1
    if D.cmp(4 * r1, ">", 1):
        s = mpmath.sqrt(4 * r1 - 1)
        if tr_can_recurse(D, la - s, mpf(1), ub_r2, rem):
            return True, 1
    side = mpmath.sqrt(2 * r1)
    if D.cmp(side, "<", 1):
        wright = la - side
        if size_bound_ok(D, ub_r2, min(wright, mpf(1))):
            wgt = wright * CR
        else:
            wgt = tr_critical_ratio_wh(wright, mpf(1)) * wright
        top = rem - (wgt + ub_r2)
        if D.cmp(top, ">", 0) and tr_can_recurse(D, side, 1 - side, ub_r2, top):
            return True, 2
    if D.cmp(4 * r1, ">", 1):
        s1 = mpmath.sqrt(4 * r1 - 1)
        t = (la - s1 / 2) ** 2
        if D.cmp(r1, ">", t):
            h1 = 2 * mpmath.sqrt(r1 - t)
            ph, pw = (1 - h1) / 2, la - s1
            ru = (rem - ub_r2) / 2
            if D.cmp(ru, ">", 0) and tr_can_recurse(D, pw, ph, ub_r2, ru):
                return True, 3
    return False, 0


def manually_proved_two_pocket(D, la, r1, r2):
    """r1_r2_strategies.cu:29-44.  A negative pocket width la - S1 becomes
    the interval [0, la - S1] (tighten_lb raises only the lower bound), whose
    abs() and square are [0, 0] (interval.hpp:216-232): its square is
    max(la - S1, 0)^2."""
    if D.cmp(r2, "<", mpf(0.2400000000000000188)) or D.cmp(r1, "<", mpf(0.4900000000000000466)):
        return False
    eps = mpf(0.00999999999999999847)
    if D.cmp(r1, "<", (la * la + 1) / 4 - eps):
        return False
    s1 = 2 * mpmath.sqrt(r1 - mpf(1) / 4)
    return D.cmp(r2, ">=", mpf(1) / 4 + max(la - s1, mpf(0)) ** 2 - eps)


def r1_r2_simple(D, la, r1, r2, ub_r3):
    """Criterion 1 (r1_r2_strategies.cu:46-72).  Strategies: 1 the manual
    two-disk pocket, 2 side strips wider than lambda, 3 the recursion on the
    rest, whose width the reference computes as lambda + width (Q10)."""
    if manually_proved_two_pocket(D, la, r1, r2):
        return True, 1
    w = two_disks_maximize_height(D, r1, r2, mpf(1))
    if w <= 0:
        return False, 0
    if D.cmp(4 * r2, ">", 1):
        lw = mpmath.sqrt(4 * r1 - 1) + mpmath.sqrt(4 * r2 - 1)
        if D.cmp(lw, ">", w):
            w = lw
            if D.cmp(lw, ">", la):
                return True, 2
    rem = tr_critical_ratio(la) * la - r1 - r2
    return (True, 3) if tr_can_recurse(D, la + w, mpf(1), ub_r3, rem) else (False, 0)


def r1_r2_stacked(D, la, r1, r2, ub_r3):
    """Criterion 2 (r1_r2_strategies.cu:74-130): r1's square in the lower
    left corner, r2's strip above it; strategy 1 the top rectangle
    size-bounded, 2 the bottom one size-bounded, 3 both by recursion."""
    side = mpmath.sqrt(2 * r1)
    if D.cmp(side, ">", 1):
        return False, 0
    hrem = 1 - side
    if D.cmp(4 * r2, "<", hrem * hrem):
        return False, 0
    w2 = mpmath.sqrt(4 * r2 - hrem * hrem)
    w_top, w_bot = la - w2, la - side
    rem = tr_critical_ratio(la) * la - r1 - r2
    if size_bound_ok(D, ub_r3, min(w_top, hrem)):
        left = rem - (CR * w_top * hrem + ub_r3)
        ok = D.cmp(left, ">", 0) and tr_can_recurse(D, w_bot, side, ub_r3, left)
        return (True, 1) if ok else (False, 0)
    if size_bound_ok(D, ub_r3, min(w_bot, side)):
        left = rem - (CR * w_bot * side + ub_r3)
        ok = D.cmp(left, ">", 0) and tr_can_recurse(D, w_top, hrem, ub_r3, left)
        return (True, 2) if ok else (False, 0)
    left = rem - (tr_critical_ratio_wh(w_top, hrem) * w_top * hrem + ub_r3)
    ok = D.cmp(left, ">", 0) and tr_can_recurse(D, w_bot, side, ub_r3, left)
    return (True, 3) if ok else (False, 0)


def r123_horizontal(D, la, r1, r2, r3, ub_4):
    """Criterion 3 (r1_r2_r3_strategies.cu:28-46): three full-height strips;
    strategy 1 they span lambda, 2 recursion on the rest."""
    if D.cmp(4 * r3, "<=", 1):
        return False, 0
    wcov = sum(mpmath.sqrt(4 * r - 1) for r in (r1, r2, r3))
    wrem = la - wcov
    if D.cmp(la, "<=", wcov):
        return True, 1
    rem = tr_critical_ratio(la) * la - r1 - r2 - r3
    return (True, 2) if tr_can_recurse(D, wrem, mpf(1), ub_4, rem) else (False, 0)


def strip_width_ok(r1, r2, r3, w):
    """can_realize_strip_width (r1_r2_r3_strategies.cu:48-67): r1 spans the
    strip of width w at height h1, r2 and r3 side by side cover the rest of
    its height with widths summing to at least w."""
    q = 4 * r1 - w * w
    if q <= 0:
        return False
    hrem = 1 - mpmath.sqrt(q)
    a, b = 4 * r2 - hrem * hrem, 4 * r3 - hrem * hrem
    if b <= 0 or a <= 0:
        return False
    return mpmath.sqrt(a) + mpmath.sqrt(b) >= w


def cover_strip_width(D, la, r1, r2, r3):
    """The widest strip width in [0.3, lambda] the bisection of
    r1_r2_r3_strategy_cover_strip accepts, None if 0.3 is not accepted.  The
    bisection runs on exact midpoints; the decisions of its first levels,
    where the bracket is still wide, count towards the slack (later levels
    close in on the same boundary in either arithmetic)."""
    if not strip_width_ok(r1, r2, r3, mpf("0.3")):
        return None
    lo, hi = mpf("0.3"), la
    for it in range(140):  # to 2^-140 of the bracket, far below a double's resolution
        mid = (lo + hi) / 2
        ok = strip_width_ok(r1, r2, r3, mid)
        if hi - lo > mpf("1e-6"):
            _strip_slack(D, r1, r2, r3, mid)
        if ok:
            lo = mid
        else:
            hi = mid
    return lo


def r123_cover_strip(D, la, r1, r2, r3, ub_4):
    """Criterion 4 (r1_r2_r3_strategies.cu:69-94): r1, r2, r3 cover a
    full-height strip of the widest width the bisection finds, recursion
    covers the rest."""
    w = cover_strip_width(D, la, r1, r2, r3)
    if w is None:
        return False, 0
    rem = tr_critical_ratio(la) * la - r1 - r2 - r3
    return (True, 1) if tr_can_recurse(D, la - w, mpf(1), ub_4, rem) else (False, 0)


def _strip_slack(D, r1, r2, r3, w):
    """The comparisons strip_width_ok decides at w, for the slack only."""
    q = 4 * r1 - w * w
    if not D.cmp(4 * r1, ">", w * w):
        return
    hrem = 1 - mpmath.sqrt(q)
    if not (D.cmp(4 * r3, ">", hrem * hrem) and D.cmp(4 * r2, ">", hrem * hrem)):
        return
    D.cmp(mpmath.sqrt(4 * r2 - hrem * hrem) + mpmath.sqrt(4 * r3 - hrem * hrem), ">=", w)


# ---------------------------------------------------- rectangle_size_bound

def s1_even_split(D, la, r1, r2):
    """Criterion 5, shortcut_even_split_recursion (even_split_recursion.cu:42-57)."""
    big = la / 2
    if D.cmp(big, "<", 1) and not size_bound_ok(D, r1, big):
        return False, 0
    small = big - r2 / CR
    return (True, 1) if (D.cmp(small, ">=", 1) or size_bound_ok(D, r2, small)) else (False, 0)


def s23_uneven_split(D, la, r1, r2):
    """Criteria 6 and 7, shortcut_uneven_split_recursion (uneven_split_recursion.cu:
    78-125); the 4-argument form is this one with r3 in place of r2."""
    if r1 <= 0:
        return False, 0
    need = mpmath.sqrt(r1) / UB_DISK_RADIUS
    if D.cmp(need, ">=", la):
        return False, 0
    rem = la * CR - (r2 + need / CR)
    if D.cmp(la * CR, "<=", r2 + need / CR):
        return False, 0
    wrem = rem / CR
    return (True, 1) if (D.cmp(wrem, ">=", 1) or size_bound_ok(D, r2, wrem)) else (False, 0)


def s4_opposite_corners(D, la, r1, r2, r3):
    """Criterion 8, shortcut_r1_r2_opposite_corners_recursion
    (r1_r2_opposite_corners_recursion.cu:28-71): r1's and r2's squares in
    opposite corners, regions A, B, C around them."""
    w1, w2 = mpmath.sqrt(2 * r1), mpmath.sqrt(2 * r2)
    comb = w1 + w2
    if not D.cmp(comb, "<", 1):
        return False, 0
    min_ab = min(la - w2, 1 - w1, la - w1, w1)
    if not size_bound_ok(D, r3, min_ab):
        return False, 0
    wc, hc = w2, 1 - comb
    budget = (2 * CR - 1) * (r1 + r2)
    if size_bound_ok(D, r3, min(wc, hc)):
        return (False, 0) if D.cmp(2 * r3, ">", budget) else (True, 1)
    extra = (bound_worst_case_ratio(D, wc, hc) - CR) * wc * hc
    return (False, 0) if D.cmp(budget - 2 * r3, "<", extra) else (True, 2)


def s5_vertical_strip(D, la, r1, r2, r3):
    """Criterion 9, shortcut_r1_r2_vertical_strip (r1_r2_vertical_strip.cu:28-41)."""
    w = two_disks_maximize_height(D, r1, r2, mpf(1))
    if w <= 0:
        return False, 0
    rem = la * CR - (r1 + r2)
    wrem = la - w
    if D.cmp(wrem, ">=", 1) or size_bound_ok(D, r3, wrem):
        return (True, 1) if D.cmp(rem, ">=", CR * wrem) else (False, 0)
    return (True, 2) if rsb_can_recurse(D, rem, r3, wrem, mpf(1)) else (False, 0)


def c2_criteria(la, r1, r2, r3):
    """All ten criteria on the point cell (la, r1, r2, r3) as the batch
    evaluates them: [(verdict, strategy, slack)] for criteria 0..9, with
    ub_r4 = min(r3, required_weight_for(la) - r1 - r2 - r3)
    (tight_rectangle.cu:97-102)."""
    la, r1, r2, r3 = (mpf(x) for x in (la, r1, r2, r3))
    ub4 = min(r3, tr_critical_ratio(la) * la - r1 - r2 - r3)
    calls = (lambda D: r1_strategies(D, la, r1, r2), lambda D: r1_r2_simple(D, la, r1, r2, r3),
             lambda D: r1_r2_stacked(D, la, r1, r2, r3), lambda D: r123_horizontal(D, la, r1, r2, r3, ub4),
             lambda D: r123_cover_strip(D, la, r1, r2, r3, ub4), lambda D: s1_even_split(D, la, r1, r2),
             lambda D: s23_uneven_split(D, la, r1, r2), lambda D: s23_uneven_split(D, la, r1, r3),
             lambda D: s4_opposite_corners(D, la, r1, r2, r3), lambda D: s5_vertical_strip(D, la, r1, r2, r3))
    out = []
    for f in calls:
        D = Dec()
        v, s = f(D)
        out.append((bool(v), s, D.slack))
    return out


# ----------------------------------- rectangle_size_bound leaf chain (C-steps)
# Point cells (lambda, r1..r6), r1 >= ... >= r6, with weight left over:
# lambda * 0.61 - sum(r) > 0.  Intermediate_values (variables.cuh:30-52):
# R = lambda * 0.61 - sum(r) with its lower bound clamped at 0, and the
# cell-independent wall-building constants.

CR_SQ = CR * CR
_WB_DEN = mpmath.sqrt(1 + mpmath.sqrt(1 - mpf("0.25") / CR_SQ))
WEIGHT_BOUND_WALL_BUILDING = (1 - 1 / _WB_DEN) ** 2 / 2
PHASE_FACTOR_WALL_BUILDING = 4 * CR_SQ - CR * mpmath.sqrt(16 * CR_SQ - 4)


def rsb_R(la, r):
    return la * CR - sum(r)


def bound_required_height(r_sq):
    """bound_required_height (disk_satisfies_size_bound.cu:30-32): sqrt(r^2) / (3/8)."""
    return mpmath.sqrt(r_sq) / UB_DISK_RADIUS


def even_split_distance_bound(D, la, r):
    """compute_even_split_distance_bound (compute_even_split_distance_bound.cu:25-76):
    greedy two groups by weight (r1 | r2 + r3, then r4..r6 to the lighter
    group by upper bound), their distance bound, then the rest R added to
    the lighter group if the smallest disk does not already bound it."""
    g1, g2 = r[0], r[1] + r[2]
    for i in range(3, 6):
        if D.cmp(g1, "<=", g2):
            g1 += r[i]
        else:
            g2 += r[i]
    dist = abs(g1 - g2)
    R = max(rsb_R(la, r), mpf(0))
    if D.cmp(dist - r[5], "<=", R):
        return r[5]
    if D.cmp(g1, "<=", g2):
        g1 += R
    else:
        g2 += R
    dist = abs(g1 - g2)
    return dist if D.cmp(dist, "<=", r[0]) else r[0]


def rsb_c1_even_split(D, la, r):
    """C1, even_split_recursion (even_split_recursion.cu:25-40)."""
    dist = even_split_distance_bound(D, la, r)
    width = (mpf("0.5") - dist / (2 * la * CR)) * la
    if D.cmp(width, ">=", 1):
        return True, 1
    return (True, 2) if size_bound_ok(D, r[0], width) else (False, 0)


def rsb_c2_uneven_split(D, la, r):
    """C2, uneven_split_recursion (uneven_split_recursion.cu:25-76): a strip
    for r1 filled by the largest disks up to its weight, the rest
    size-bounded."""
    if r[0] <= 0:
        return False, 0
    need = bound_required_height(r[0])
    if D.cmp(need, ">=", la):
        return False, 0
    wneed = need * CR
    w, at = mpf(0), -1
    for i in range(6):
        w += r[i]
        if D.cmp(w, ">=", wneed):
            at = i
            break
    largest = r[5] if at < 0 or at == 5 else r[at + 1]
    d1 = wneed + r[5] if at < 0 else w
    if D.cmp(d1, ">=", la * CR):
        return False, 0
    wrem = la - d1 / CR
    if D.cmp(wrem, ">=", 1):
        return True, 1
    return (True, 2) if size_bound_ok(D, largest, wrem) else (False, 0)


def rsb_c4_wall_building(D, la, r):
    """C4, vertical_wall_building_recursion (vertical_wall_building_recursion.cu:
    25-78), with the reference's "remaining width" lambda + width (Q5)."""
    R = max(rsb_R(la, r), mpf(0))
    rem_w = la + mpmath.sqrt(2 * r[5])
    if r[5] <= 0 or D.cmp(R, "<=", 0):
        return False, 0
    if D.cmp(rem_w, "<", 1) and not size_bound_ok(D, r[0], rem_w):
        return False, 0
    if D.cmp(r[5], ">", WEIGHT_BOUND_WALL_BUILDING):
        return False, 0
    r6_total = r[5] + R
    below = r[5] * PHASE_FACTOR_WALL_BUILDING
    used_in_phases = CR * mpmath.sqrt(2 * r[5])
    for k in range(32):
        total = r6_total - used_in_phases
        if D.cmp(r6_total, "<=", used_in_phases):
            break
        req = CR * bound_required_height(below)
        if D.cmp(req, "<=", total):
            arw = la + (req + below) / CR
            if D.cmp(arw, ">=", 1) or size_bound_ok(D, r[0], arw):
                return True, 1
        used_in_phases += CR * mpmath.sqrt(2 * below)
        below *= PHASE_FACTOR_WALL_BUILDING
    return False, 0


def rsb_c7_vertical_strip(D, la, r):
    """C7, r1_r2_vertical_strip (r1_r2_vertical_strip.cu:43-95): r1 and r2 in
    a full-height strip, then each of r3..r6 recursion or a strip off the
    longer side of the rest."""
    w = two_disks_maximize_height(D, r[0], r[1], mpf(1))
    if w <= 0:
        return False, 0
    rem = la * CR - (r[0] + r[1])
    wrem, hrem = la - w, mpf(1)
    for md in range(2, 6):
        if rsb_can_recurse(D, rem, r[md], wrem, hrem):
            return True, 1
        ms = min(wrem, hrem)
        if D.cmp(4 * r[md], "<=", ms * ms):
            return False, 0
        cut = mpmath.sqrt(4 * r[md] - ms * ms)
        if D.cmp(wrem, "<", hrem):
            hrem -= cut
            if D.cmp(hrem, "<=", 0):
                return True, 2
        else:
            wrem -= cut
            if D.cmp(wrem, "<=", 0):
                return True, 2
        rem -= r[md]
    r7 = min(r[5], rsb_R(la, r))
    return (True, 3) if rsb_can_recurse(D, rem, r7, wrem, hrem) else (False, 0)


def circle_intersection(D, c1, c2):
    """intersection (operations.cu:30-83) of two circles (x, y, squared
    radius): (definitely intersecting, first point, second point), the first
    point being the '+' one when c2 lies below c1, or level with it and not
    to its left (the reference's order rule, Q17)."""
    (x1, y1, q1), (x2, y2, q2) = c1, c2
    d2 = (x2 - x1) ** 2 + (y2 - y1) ** 2
    if d2 == 0:
        return False, None, None
    rd = q1 - q2
    if D.cmp(2 * (q1 + q2) / d2, "<", 1 + rd * rd / (d2 * d2)):
        return False, None, None
    f1 = mpmath.sqrt(2 * (q1 + q2) / d2 - 1 - rd * rd / (d2 * d2)) / 2
    f2 = rd / (2 * d2)
    xd, yd = x2 - x1, y2 - y1
    cx, cy = (x1 + x2) / 2 + f2 * xd, (y1 + y2) / 2 + f2 * yd
    plus, minus = (cx + f1 * yd, cy - f1 * xd), (cx - f1 * yd, cy + f1 * xd)
    if D.cmp(y2, "<", y1) or (y2 == y1 and D.cmp(x2, ">=", x1)):
        return True, plus, minus
    return True, minus, plus


def dist2(p, q):
    return (p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2


def _rem_width_ok(D, la, w, largest):
    """The strip of width w leaves la - w: wide enough, or the largest
    remaining disk meets the size bound for it."""
    rem = la - w
    return D.cmp(rem, ">=", 1) or size_bound_ok(D, largest, rem)


def _c3_nd(D, la, r, k):
    """multi_disk_strip_vertical_nd (multi_disk_strip.cu:31-62): r1..rk in a
    vertical strip of the width their weight needs at ratio 0.61, stacked
    while each spans it; the rest of the rectangle for the largest remaining
    disk (r6 when all six are in the strip, as the reference has it)."""
    w = sum(r[:k]) / CR
    ht = mpf(0)
    for j in range(k):
        if D.cmp(4 * r[j], "<=", w * w):
            break
        ht += mpmath.sqrt(4 * r[j] - w * w)
    if D.cmp(ht, "<", 1):
        return False
    return _rem_width_ok(D, la, w, r[5] if k == 6 else r[k])


def _c3_stacked_r34(D, la, r):
    """r1_r2_stacked_r3_r4_horizontal (multi_disk_strip.cu:64-98)."""
    w = sum(r[:4]) / CR
    if D.cmp(4 * r[0], "<=", w * w) or D.cmp(4 * r[1], "<=", w * w):
        return False
    hrem = 1 - (mpmath.sqrt(4 * r[0] - w * w) + mpmath.sqrt(4 * r[1] - w * w))
    if D.cmp(4 * r[3], "<=", hrem * hrem) or D.cmp(4 * r[2], "<=", hrem * hrem):
        return False
    w34 = mpmath.sqrt(4 * r[2] - hrem * hrem) + mpmath.sqrt(4 * r[3] - hrem * hrem)
    if D.cmp(w34, "<", w):
        return False
    return _rem_width_ok(D, la, w, r[4])


def _c3_pair_height(D, w, ra, rb):
    """Two disks side by side in the strip of width w (multi_disk_strip.cu:
    116-126): the lower of their heights, or None if one cannot span its part."""
    diff = max(ra - rb, mpf(0))
    wa = 2 * diff / w + w / 2
    if D.cmp(4 * ra, "<=", wa * wa) or D.cmp(4 * rb, "<=", (w - wa) ** 2):
        return None
    return mpmath.sqrt(4 * ra - wa * wa)


def _c3_stacked_r34_r56(D, la, r):
    """r1_r2_stacked_r3_r4_and_r5_r6_horizontal (multi_disk_strip.cu:100-143)."""
    w = sum(r) / CR
    if D.cmp(4 * r[0], "<=", w * w) or D.cmp(4 * r[1], "<=", w * w):
        return False
    h34 = _c3_pair_height(D, w, r[2], r[3])
    if h34 is None:
        return False
    h56 = _c3_pair_height(D, w, r[4], r[5])
    if h56 is None:
        return False
    h12 = mpmath.sqrt(4 * r[0] - w * w) + mpmath.sqrt(4 * r[1] - w * w)
    if D.cmp(h12 + h34 + h56, "<", 1):
        return False
    return _rem_width_ok(D, la, w, r[5])


# The reference's height of a two-disk row in six_disk_strip_three_rows: the
# square root taken twice (Q6).
ROW_HEIGHT = lambda a: mpmath.sqrt(mpmath.sqrt(a))  # noqa: E731


def _c3_three_rows(D, la, r):
    """six_disk_strip_three_rows (multi_disk_strip.cu:145-238): r1 and r4 on
    top, r2 and r3 below, r5 and r6 covering the middle row's left and right
    ends, in the strip of the width the six disks need.  The heights of the
    two-disk rows take the square root twice, as the reference does (Q6)."""
    w = sum(r) / CR
    r7 = min(r[5], rsb_R(la, r))
    if D.cmp(la - w, "<", 1) and not size_bound_ok(D, r7, la - w):
        return False
    rows = []
    for a, b in ((0, 3), (1, 2)):
        arg = 2 * (r[a] + r[b]) - w * w / 4 - 4 * (r[a] - r[b]) ** 2 / (w * w)
        if D.cmp(arg, "<", 0):
            return False  # the reference's NaN path ends in false
        rows.append(ROW_HEIGHT(arg))
    h14, h23 = rows
    hrem = 1 - h23 - h14
    if D.cmp(hrem, "<=", 0):
        return True
    if D.cmp(r[4], "<", hrem * hrem / 4) or D.cmp(r[5], "<", hrem * hrem / 4):
        return False
    x5 = mpmath.sqrt(r[4] - hrem * hrem / 4)
    x6 = w - mpmath.sqrt(r[5] - hrem * hrem / 4)
    y5 = hrem / 2 + h23
    y1, y2 = 1 - h14 / 2, h23 / 2
    w1 = (2 / w) * (r[0] - r[3]) + w / 2
    w2 = (2 / w) * (r[1] - r[2]) + w / 2
    c1, c2 = (w - w1 / 2, y1, r[0]), (w - w2 / 2, y2, r[1])
    c3, c4 = ((w - w2) / 2, y2), ((w - w1) / 2, y1)
    c5, c6 = (x5, y5, r[4]), (x6, y5)
    ok15, p15, q15 = circle_intersection(D, c5, c1)
    ok25, p25, q25 = circle_intersection(D, c5, c2)
    if not ok15 or not ok25:
        return False
    if D.cmp(dist2(c4, p15), ">", r[3]) or D.cmp(dist2(c3, p25), ">", r[2]):
        return False
    return D.cmp(dist2(c6, q15), "<=", r[5]) and D.cmp(dist2(c6, q25), "<=", r[5])


def rsb_c3_multi_disk_strip(D, la, r):
    """C3, multi_disk_strip_vertical (multi_disk_strip.cu:240-262)."""
    for k in range(2, 7):
        if _c3_nd(D, la, r, k):
            return True, 1
    if _c3_stacked_r34(D, la, r):
        return True, 2
    if _c3_stacked_r34_r56(D, la, r):
        return True, 3
    if _c3_three_rows(D, la, r):
        return True, 4
    return False, 0


def _chords(d, x):
    """Sum of the chords the disks of squared radii d cut from a strip of
    width (or height) x: sqrt(4 d_i - x^2) where positive."""
    return sum((mpmath.sqrt(4 * q - x * x) for q in d if 4 * q > x * x), mpf(0))


def _chords_dec(D, d, x):
    """_chords with each disk's "spans the strip" decision recorded."""
    t = mpf(0)
    for q in d:
        if not D.cmp(4 * q, "<=", x * x):
            t += mpmath.sqrt(4 * q - x * x)
    return t


def _sup_decreasing(f, a, b):
    """The point where the continuous non-increasing f (f(a) >= 0 > f(b))
    crosses zero, by bisection to far below any compared slack."""
    for _ in range(120):  # 2^-120 of the interval: far below any compared slack
        m = (a + b) / 2
        if f(m) >= 0:
            a = m
        else:
            b = m
    return (a + b) / 2


def nd_can_cover(D, w, h, d):
    """nd_can_cover (maximize_covered.cu:28-71): the disks d (squared radii,
    at most 6) cover a w x h rectangle in one row or one column."""
    n = len(d)
    if n == 1:
        return D.cmp(4 * d[0], ">=", w * w + h * h)
    if n == 2:
        return (D.cmp(two_disks_maximize_height(D, d[0], d[1], w), ">=", h)
                or D.cmp(two_disks_maximize_height(D, d[0], d[1], h), ">=", w))
    if n == 3:
        return D.cmp(_chords_dec(D, d, h), ">=", w)
    return D.cmp(_chords_dec(D, d, w), ">=", h) or D.cmp(_chords_dec(D, d, h), ">=", w)


def nd_maximize_covered(D, w, d):
    """nd_maximize_covered (maximize_covered.cu:73-104, three_disks_maximize_height
    :198-237): the greatest height of a strip of width w the disks d cover
    (side by side; for four or more also stacked), below 2 sqrt(d_last), the
    bound the reference's bisection starts from; 0 if none."""
    n = len(d)
    if n == 1:
        return mpf(0) if D.cmp(4 * d[0], "<=", w * w) else mpmath.sqrt(4 * d[0] - w * w)
    if n == 2:
        return two_disks_maximize_height(D, d[0], d[1], w)
    hi = 2 * mpmath.sqrt(d[-1])
    side = mpf(0)  # side by side: the chords at height h reach w up to this height
    if not D.cmp(_chords(d, mpf(0)), "<", w):
        if D.cmp(_chords(d, hi), ">=", w):
            side = hi
        else:
            side = _sup_decreasing(lambda h: _chords(d, h) - w, mpf(0), hi)
    if n == 3:
        return side
    stacked = _chords_dec(D, d, w)  # stacked: the chords at width w
    return min(max(side, stacked), hi)  # max and min: continuous, no decision to record


def _c5_backtrack(D, d, reached, target, cross):
    """The backtracking of r1_in_corner_explicit_recursion (:60-100): the
    disks d, in order, split into consecutive groups, each covering a strip
    of extent `cross` as far as it can; true once the strips reach `target`,
    or the last group covers what is left."""
    for choice in range(1, len(d) + 1):
        group = d[:choice]
        if choice == len(d):
            return nd_can_cover(D, cross, target - reached, group)
        got = nd_maximize_covered(D, cross, group)
        if got == 0:
            continue
        if D.cmp(reached + got, ">=", target):
            return True
        if _c5_backtrack(D, d[choice:], reached + got, target, cross):
            return True
    return False


def _c5_max_remaining(la, r, nd):
    return min(r[5], rsb_R(la, r)) if nd == 5 else r[nd + 1]


def _c5_above(D, la, r, nd):
    """r1_in_corner_explicit_recursion_above_nd (:35-104): r1..r(nd+1) in a
    bottom-left strip of the width their weight needs, r1 in the corner,
    groups of the others stacked above it."""
    w1 = sum(r[: nd + 1]) / CR
    allowed = la - bound_required_height(_c5_max_remaining(la, r, nd))
    if D.cmp(w1, ">", allowed) or D.cmp(4 * r[0], "<=", w1 * w1):
        return False
    return _c5_backtrack(D, r[1: nd + 1], mpmath.sqrt(4 * r[0] - w1 * w1), mpf(1), w1)


def _c5_right(D, la, r, nd):
    """r1_in_corner_explicit_recursion_right_nd (:106-175): the same in a
    bottom strip of the height their weight needs, groups to the right of r1."""
    h1 = sum(r[: nd + 1]) / CR / la
    allowed = 1 - bound_required_height(_c5_max_remaining(la, r, nd))
    if D.cmp(h1, ">", allowed) or D.cmp(4 * r[0], "<=", h1 * h1):
        return False
    return _c5_backtrack(D, r[1: nd + 1], mpmath.sqrt(4 * r[0] - h1 * h1), la, h1)


def _c5_right_34_and_2(D, la, r):
    """r1_in_corner_right_34_and_2_recursion (:177-247): r1 in a corner
    square, r3 and r4 stacked to its right, r2 beside them, the extra weight
    of the efficient cover against the cost of the rest B."""
    w1 = mpmath.sqrt(2 * r[0])
    wrem = la - w1
    if not size_bound_ok(D, r[4], 1 - w1):
        return False
    h34 = two_disks_maximize_height(D, r[2], r[3], wrem)
    if h34 <= 0 or D.cmp(h34, ">=", w1):
        return False
    hrem = w1 - h34
    if D.cmp(4 * r[1], "<=", hrem * hrem):
        return False
    w2 = mpmath.sqrt(4 * r[1] - hrem * hrem)
    no_b = D.cmp(w2, ">=", wrem)
    if no_b or w2 > wrem:
        w2 = wrem
    a234 = w2 * w1 + h34 * (wrem - w2)
    extra = (2 * CR - 1) * r[0] + (CR * a234 - (r[1] + r[2] + r[3]))
    if D.cmp(extra, "<=", 0):
        return False
    if no_b:
        return True
    hb, wb = w1 - h34, wrem - w2
    cost = r[5]
    if not size_bound_ok(D, r[5], min(hb, wb)):
        cost += (bound_worst_case_ratio(D, wb, hb) - CR) * hb * wb
    return D.cmp(cost, "<=", extra)


def rsb_c5_r1_in_corner(D, la, r):
    """C5, r1_in_corner_explicit_recursion (r1_in_corner_explicit_recursion.cu:249-262)."""
    for nd in range(2, 6):
        if _c5_above(D, la, r, nd):
            return True, 1
        if _c5_right(D, la, r, nd):
            return True, 2
    if _c5_right_34_and_2(D, la, r):
        return True, 3
    return False, 0


WIDTH_FRACTION_INCOMPLETE = 1 / _WB_DEN
LB_W_INFTY_DENOM = 1 - mpmath.sqrt(PHASE_FACTOR_WALL_BUILDING)


def rsb_c10_r1_corner_wall_building(D, la, r):
    """C10, r1_in_corner_wall_building_recursion (r1_in_corner_wall_building_recursion.cu:
    27-149): r1 covers a square in the lower-left corner; the weight it saves
    pays for the region A above it (size-bounded recursion, or a row of wall
    building with r6 and smaller), r2 (with r3, or r3 and r4, when it does
    not fit) goes to the region B right of it, and the weight dropped to B
    by incomplete rows is bounded."""
    w1 = mpmath.sqrt(2 * r[0])
    extra_r1 = (2 * CR - 1) * r[0]
    wb, ha = la - w1, 1 - w1
    min_a = ha if D.cmp(ha, "<", w1) else w1
    max_a = w1 if min_a is ha else ha
    wb_a = UB_DISK_WEIGHT * min_a * min_a
    extra_a = r[5] if D.cmp(wb_a, ">=", r[5]) else wb_a
    if wb_a <= 0 or D.cmp(extra_a, ">", extra_r1):
        return False, 0
    if D.cmp(wb, "<", 1) and not size_bound_ok(D, r[1], wb):
        better = -INF
        h23 = two_disks_maximize_height(D, r[1], r[2], wb)
        if size_bound_ok(D, r[3], min(1 - h23, wb)):
            better = CR * h23 * wb - (r[1] + r[2])
        h234 = nd_maximize_covered(D, wb, [r[1], r[2], r[3]])
        if size_bound_ok(D, r[4], min(1 - h234, wb)):
            e = CR * h234 * wb - (r[1] + r[2] + r[3])
            if better == -INF or D.cmp(e, ">", better):
                better = e
        extra_r1 = extra_r1 + better
        if better == -INF or D.cmp(extra_r1, "<=", 0):
            return False, 0
    if D.cmp(WEIGHT_BOUND_WALL_BUILDING * max_a * max_a, "<", r[5]):
        return False, 0
    short = D.cmp(WEIGHT_BOUND_WALL_BUILDING * min_a * min_a, ">=", r[5])
    side = min_a if short else max_a
    cost_wall = CR * side * mpmath.sqrt(2 * r[5])
    if D.cmp(cost_wall, ">", extra_a):
        extra_a = cost_wall
    if D.cmp(extra_a, ">", extra_r1):
        return False, 0
    if D.cmp(r[5], "<=", wb_a):
        return True, 1
    r6_weight = la * CR - sum(r[:5])
    width_incomplete = side * WIDTH_FRACTION_INCOMPLETE
    need_a = CR * w1 * ha
    dropped_inf = CR * width_incomplete * mpmath.sqrt(2 * r[5]) / LB_W_INFTY_DENOM
    if D.cmp(r6_weight - dropped_inf, ">=", need_a):
        return True, 2
    dropped, cur = need_a, r[5]
    while D.cmp(cur, ">", wb_a):
        dropped += CR * width_incomplete * mpmath.sqrt(2 * cur)
        cur *= PHASE_FACTOR_WALL_BUILDING
    return (True, 3) if D.cmp(r6_weight - dropped, ">=", need_a) else (False, 0)


def _c9_cover_rest(D, rem, largest, w, h):
    """A w x h rectangle with weight rem: at 0.61 per area when the largest
    disk meets the size bound for it, else by can_recurse."""
    if size_bound_ok(D, largest, min(w, h)):
        return D.cmp(CR * w * h, "<=", rem)
    return rsb_can_recurse(D, rem, largest, w, h)


def _c9_one_horizontal(D, la, r, w, weight):
    """l_shaped_recursion_two_disks_vertical_one_horizontal (l_shaped_recursion.cu:
    27-122): r3 as a square in the corner of the w x 1 rest; the L around
    it split top / bottom-right or top-left / right, or r3 across the whole
    width and the rest above it."""
    R4 = weight - r[2]
    s3 = mpmath.sqrt(2 * r[2])
    htop = 1 - s3
    if size_bound_ok(D, r[3], min(w, htop)):
        rem = R4 - (CR * htop * w + r[3])
        if D.cmp(rem, ">", 0) and _c9_cover_rest(D, rem, r[4], w - s3, s3):
            return 1
    wright = la - s3
    if D.cmp(wright, ">=", 1) or size_bound_ok(D, r[3], wright):
        rem = R4 - (CR * wright + r[3])
        if D.cmp(rem, ">", 0) and _c9_cover_rest(D, rem, r[4], s3, 1 - s3):
            return 2
    if not D.cmp(4 * r[2], "<=", w * w):
        hrem = 1 - mpmath.sqrt(4 * r[2] - w * w)
        if _c9_cover_rest(D, R4, r[3], w, hrem):
            return 3
    return 0


def rsb_c9_l_shaped(D, la, r):
    """C9, l_shaped_recursion (l_shaped_recursion.cu:124-155): r1 and r2
    stacked in a full-height strip, then r3 (one horizontal) or r3 and r4
    (two) in the rest; the two-horizontal form recurses with r6 as the
    largest remaining disk, as the reference has it."""
    weight = la * CR - (r[0] + r[1])
    wrem = la - two_disks_maximize_height(D, r[0], r[1], mpf(1))
    if not D.cmp(wrem * CR, ">", weight):
        s = _c9_one_horizontal(D, la, r, wrem, weight)
        if s:
            return True, s
    w12 = two_disks_maximize_height(D, r[0], r[1], mpf(1))
    if w12 <= 0:  # exactly 0: the function decided that (its comparisons are recorded)
        return False, 0
    wrem = la - w12
    h34 = two_disks_maximize_height(D, r[2], r[3], wrem)
    if h34 <= 0:
        return False, 0
    R5 = la * CR - (r[0] + r[1] + r[2] + r[3])
    return (True, 4) if rsb_can_recurse(D, R5, r[5], wrem, 1 - h34) else (False, 0)


def _two_by_two_concrete(D, a, b, c, d):
    """compute_two_by_two_cover_concrete (two_by_two_cover.cu:28-76): a over b
    in the first column, c over d in the second, each column as wide as its
    two disks stacked can cover at full height; a column that cannot is left
    out (width 0, centres at the origin).  Returns (width, circles[0..3])."""
    def column(top, bottom):
        diff = top - bottom
        h1, h2 = mpf("0.5") + 2 * diff, mpf("0.5") - 2 * diff
        wsq = 2 * (top + bottom) - mpf("0.25") - 4 * diff * diff
        if (D.cmp(top, "<=", h1 * h1 / 4) or D.cmp(bottom, "<=", h2 * h2 / 4)
                or D.cmp(2 * (top + bottom), "<=", mpf("0.25") + 4 * diff * diff)):
            return None
        return mpmath.sqrt(wsq), h1, h2
    circles = [None] * 4
    width, w1 = mpf(0), mpf(0)
    col = column(a, b)
    if col is None:
        circles[0], circles[1] = (mpf(0), mpf(0), a), (mpf(0), mpf(0), b)
    else:
        w1, h1, h2 = col
        circles[0], circles[1] = (w1 / 2, h1 / 2, a), (w1 / 2, 1 - h2 / 2, b)
        width = w1
    col = column(c, d)
    if col is None:
        circles[3], circles[2] = (mpf(0), mpf(0), c), (mpf(0), mpf(0), d)
    else:
        w2, h1, h2 = col
        x = w1 + w2 / 2
        circles[3], circles[2] = (x, 1 - h1 / 2, c), (x, h2 / 2, d)
        width += w2
    return width, circles


def two_by_two_cover(D, r):
    """compute_two_by_two_cover (two_by_two_cover.cu:78-89): the widest of the
    three pairings of r1..r4 (ties to the later pairing, as the reference
    selects)."""
    c1 = _two_by_two_concrete(D, r[0], r[1], r[2], r[3])
    c2 = _two_by_two_concrete(D, r[0], r[2], r[1], r[3])
    c3 = _two_by_two_concrete(D, r[0], r[3], r[1], r[2])

    def wider(a, b):  # two pairings left out entirely tie at exactly 0 in every arithmetic
        return False if a[0] == 0 and b[0] == 0 else D.cmp(a[0], ">", b[0])
    if wider(c1, c2):
        return c1 if wider(c1, c3) else c3
    return c2 if wider(c2, c3) else c3


def rsb_c8_two_by_two(D, la, r):
    """C8 (rectangle_size_bound.cu:206-209, two_by_two_cover_with_strip_and_recursion.cu:
    25-105): the 2x2 cover of r1..r4 covers the rectangle, or the strip it
    leaves is covered by recursion from r5, r6 or r7 on, or r5 and r6
    placed against its right end with the disk of the 2x2 cover."""
    width, circles = two_by_two_cover(D, r)
    if width <= 0:  # both columns left out (decided inside)
        return False, 0
    if D.cmp(width, ">", la):
        return True, 1
    R = max(rsb_R(la, r), mpf(0))
    w = la - width
    R5 = la * CR - sum(r[:4])
    R6 = R5 - r[4]
    if rsb_can_recurse(D, R5, r[4], w, mpf(1)):
        return True, 2
    if D.cmp(4 * r[4], "<=", w * w):
        return False, 0
    h5 = mpmath.sqrt(4 * r[4] - w * w)
    hrem = 1 - h5
    if D.cmp(h5, ">=", 1) or rsb_can_recurse(D, R6, r[5], w, hrem):
        return True, 3
    if not D.cmp(4 * r[5], "<=", w * w):
        h6 = mpmath.sqrt(4 * r[5] - w * w)
        if D.cmp(h6, ">=", hrem) or rsb_can_recurse(D, R, r[5], w, hrem - h6):
            return True, 4
    if not D.cmp(4 * r[5], "<=", hrem * hrem):
        w6 = mpmath.sqrt(4 * r[5] - hrem * hrem)
        if D.cmp(w6, ">=", w):
            return True, 5
        c5 = (la - w / 2, h5 / 2, r[4])
        c6 = (la - w6 / 2, 1 - hrem / 2, r[5])
        ok, p0, p1 = circle_intersection(D, c5, c6)
        c4 = circles[3]
        if ok and D.cmp(p0[0], "<", p1[0]) and D.cmp(dist2(p0, c4[:2]), "<=", c4[2]):
            ok46, q0, q1 = circle_intersection(D, c6, c4)
            if ok46:
                lb_y = q0[1] if D.cmp(q0[1], ">", q1[1]) else q1[1]
                if rsb_can_recurse(D, R, r[5], w - w6, 1 - lb_y):
                    return True, 6
    return False, 0


_SQRT_4_M_INV_CRSQ = mpmath.sqrt(4 - 1 / (CR * CR))


def efficient_cover_weight_range(w, h):
    """compute_efficient_rectangle_cover_weight_range (compute_efficient_rectangle_cover_weight_range.cu:
    25-35): the squared radii of a disk that covers a w x h rectangle (h the
    shorter side) at density 0.61 or better."""
    lb = CR * CR * h * h * (2 - _SQRT_4_M_INV_CRSQ)
    ub = min(CR * h * w, CR * CR * h * h * (2 + _SQRT_4_M_INV_CRSQ))
    return lb, ub


def _c6_size_bounds(D, la, r, wa, ha, wb, hb, start):
    """r1_r2_opposite_corners_recursion_check_size_bounds (r1_r2_opposite_corners_recursion.cu:
    181-247): the disks from `start` on fill the region of the larger
    shorter side up to its weight, then the other; the extra cost of the
    split on each.  Returns (works, cost for the larger, cost for the smaller)."""
    min_a, min_b = min(wa, ha), min(wb, hb)
    a_smaller = D.cmp(min_a, "<", min_b)
    bigger_min, smaller_min = (min_b, min_a) if a_smaller else (min_a, min_b)
    bigger_area, smaller_area = (wb * hb, wa * ha) if a_smaller else (wa * ha, wb * hb)
    r7 = min(r[5], rsb_R(la, r))
    if start < 6:
        if not size_bound_ok(D, r[start], bigger_min):
            return False, 0, 0
        acc = mpf(0)
        while start < 6:
            acc += r[start]
            if D.cmp(acc, ">=", CR * bigger_area):
                break
            start += 1
    if start < 6:
        if not size_bound_ok(D, r[start], smaller_min):
            return False, 0, 0
        cost_bigger = r[start]
        acc = mpf(0)
        while start < 6:
            acc += r[start]
            if D.cmp(acc, ">=", CR * smaller_area):
                break
            start += 1
        return True, cost_bigger, (r[start] if start < 6 else r7)
    if not size_bound_ok(D, r7, smaller_min):
        return False, 0, 0
    return True, r7, r7


def rsb_c6_opposite_corners(D, la, r):
    """C6, r1_r2_opposite_corners_recursion (r1_r2_opposite_corners_recursion.cu:
    249-340): r1 and r2 as squares in opposite corners, not meeting and not
    covering the height; the regions A and B beside them by recursion, the
    gap C between them filled with disks that cover it efficiently, or by
    worst-case recursion."""
    w1, w2 = mpmath.sqrt(2 * r[0]), mpmath.sqrt(2 * r[1])
    comb = w1 + w2
    extra12 = (2 * CR - 1) * (r[0] + r[1])
    if not D.cmp(comb, "<", 1):
        return False, 0
    wa, wb, ha, hb = la - w2, la - w1, 1 - w1, w1
    wc, hc = w2, 1 - comb
    shorter, longer = (wc, hc) if D.cmp(wc, "<=", hc) else (hc, wc)
    lo, hi = efficient_cover_weight_range(longer, shorter)
    md = 2
    while md < 6:
        if D.cmp(lo, ">", r[md]) or D.cmp(r[md], ">", hi):
            break
        if D.cmp(4 * r[md], "<=", shorter * shorter):
            break
        covered = mpmath.sqrt(4 * r[md] - shorter * shorter)
        if D.cmp(longer, "<", covered):  # C is covered at density 0.61
            ok, cost_big, _ = _c6_size_bounds(D, la, r, wa, ha, wb, hb, md + 1)
            return (True, 1) if ok and D.cmp(cost_big, "<=", extra12) else (False, 0)
        longer -= covered
        if not D.cmp(shorter, "<=", longer):
            shorter, longer = longer, shorter
        lo, hi = efficient_cover_weight_range(longer, shorter)
        md += 1
    ok, cost_big, cost_small = _c6_size_bounds(D, la, r, wa, ha, wb, hb, md)
    if ok and size_bound_ok(D, cost_small, shorter):
        return (True, 2) if D.cmp(cost_small + cost_big, "<=", extra12) else (False, 0)
    cr_c = bound_worst_case_ratio(D, wc, hc)
    big = mpf(0)
    while md < 6:
        big += r[md]
        if D.cmp(big, ">=", cr_c * wc * hc):
            break
        md += 1
    ok, cost_big, cost_small = _c6_size_bounds(D, la, r, wa, ha, wb, hb, md)
    if ok:
        cost = cost_small + cost_big + (cr_c - CR) * wc * hc
        return (True, 3) if D.cmp(cost, "<=", extra12) else (False, 0)
    return False, 0


def rsb_c11_advanced_strip(D, la, r):
    """C11, advanced_multi_disk_strip_vertical (multi_disk_strip.cu:264-317):
    any proper subset of r1..r6 stacked in a full-height strip of the width
    its weight needs at 0.61, each disk spanning it; the rest for the
    largest disk not in the subset."""
    for mask in range(1, 63):
        inside = [j for j in range(6) if mask >> j & 1]
        w = sum(r[j] for j in inside) / CR
        if any(D.cmp(4 * r[j], "<=", w * w) for j in inside):
            continue
        if D.cmp(sum(mpmath.sqrt(4 * r[j] - w * w) for j in inside), "<", 1):
            continue
        largest = r[min(j for j in range(6) if not mask >> j & 1)]
        if _rem_width_ok(D, la, w, largest):
            return True, 1
    return False, 0


def max_height_strip(D, la, r1, r2):
    """two_disks_maximal_height_strip (maximize_covered.cu:132-150): r1 and r2
    side by side at the bottom of a strip of width la, as high as both
    reach; (height, circle 1, circle 2), height 0 if one cannot."""
    wd = 2 * (r1 - r2) / la
    w1, w2 = la / 2 + wd, la / 2 - wd
    if D.cmp(r1, "<=", w1 * w1 / 4) or D.cmp(r2, "<=", w2 * w2 / 4):
        return mpf(0), (mpf(0), mpf(0), r1), (mpf(0), mpf(0), r2)
    h = mpmath.sqrt(r1 - w1 * w1 / 4)
    return 2 * h, (w1 / 2, h, r1), (la - w2 / 2, h, r2)


def _trap(D):
    """A reference trap site: the cell is left out of the comparison."""
    D.slack = mpf(0)
    return False


def _c12_center(D, la, r):
    """five_disk_full_cover_3v2_recursion_center (five_disk_full_cover.cu:92-171)."""
    R6 = r[5] + max(rsb_R(la, r), mpf(0))
    height, c1, c2 = max_height_strip(D, la, r[0], r[1])
    if height <= 0:
        return 0
    hrem = 1 - height
    if D.cmp(4 * r[3], "<=", hrem * hrem) or D.cmp(4 * r[4], "<=", hrem * hrem):
        return 0
    w4, w5 = mpmath.sqrt(4 * r[3] - hrem * hrem), mpmath.sqrt(4 * r[4] - hrem * hrem)
    y45 = 1 - hrem / 2
    c4, c5 = (la - w4 / 2, y45, r[3]), (w5 / 2, y45, r[4])
    wrem = la - (w4 + w5)
    if D.cmp(4 * r[2], "<=", wrem * wrem):
        return 0
    h3 = mpmath.sqrt(4 * r[2] - wrem * wrem)
    if D.cmp(h3, ">=", hrem):
        return 1
    c3 = (w5 + wrem / 2, 1 - h3 / 2, r[2])
    ok15, p15, q15 = circle_intersection(D, c1, c5)
    ok12, p12, q12 = circle_intersection(D, c1, c2)
    ok24, p24, q24 = circle_intersection(D, c2, c4)
    if not (ok12 and ok15 and ok24):
        return _trap(D)
    if D.cmp(p15[0], ">=", q15[0]) or D.cmp(p24[0], ">=", q24[0]):
        return 0
    top12 = p12 if D.cmp(p12[1], ">", q12[1]) else q12
    in12 = D.cmp(dist2(top12, c3[:2]), "<=", r[2])
    in15 = D.cmp(dist2(q15, c3[:2]), "<=", r[2])
    in24 = D.cmp(dist2(p24, c3[:2]), "<=", r[2])
    if in12 and in15 and in24:
        return 2
    if in15 and in24:
        ok31, p31, q31 = circle_intersection(D, c1, c3)
        ok32, p32, q32 = circle_intersection(D, c2, c3)
        if ok31 and ok32 and D.cmp(p31[0], "<", q31[0]) and D.cmp(p32[0], "<", q32[0]):
            pw = p32[0] - q31[0]
            ph = max(q31[1], p32[1]) - top12[1]
            return 3 if rsb_can_recurse(D, R6, r[5], pw, ph) else 0
    return 0


def _c12_border32(D, la, r, t1, t2, t3, a, b):
    """five_disk_full_cover_border_cover32_concrete (five_disk_full_cover.cu:173-194)."""
    htop = nd_maximize_covered(D, la, [r[t1], r[t2], r[t3]])
    if htop <= 0:
        return False
    hrem = 1 - htop
    wright = two_disks_maximize_height(D, r[a], r[b], hrem)
    if wright <= 0:
        return False
    return rsb_can_recurse(D, la * CR - sum(r[:5]), r[5], la - wright, hrem)


def rsb_c12_five_disk(D, la, r):
    """C12, five_disk_full_cover (five_disk_full_cover.cu:200-206)."""
    s = _c12_center(D, la, r)
    if s:
        return True, s
    if _c12_border32(D, la, r, 0, 1, 2, 3, 4) or _c12_border32(D, la, r, 0, 1, 4, 2, 3):
        return True, 4
    return False, 0


def can_recurse_worst_case(D, rem, w, h):
    """can_recurse_worst_case (can_recurse.cu:237-261)."""
    lng, sht = (h, w) if D.cmp(w, "<", h) else (w, h)
    if D.cmp(lng / sht, "<=", TOL_LAMBDA_RSB):
        return D.cmp(w * h * C_WORST, "<=", rem)
    return D.cmp(sht * sht / 2 + lng * lng / 4, "<=", rem)


def max_height_strip_wc(D, r1, r2, R, w):
    """two_disks_maximal_height_strip_wc_recursion (maximize_covered.cu:239-278):
    the greatest height h below 2 r2 (the reference's bound) at which r1 and
    r2 side by side at the bottom of the w-wide strip leave a remainder that
    worst-case recursion covers with weight R; (height, circle 1, circle 2),
    height 0 if none.  The test passes, then fails, as h grows."""
    def test(h, Dx):
        if 4 * r1 < h * h or 4 * r2 < h * h:
            return False
        wrem = w - (mpmath.sqrt(4 * r1 - h * h) + mpmath.sqrt(4 * r2 - h * h))
        return wrem <= 0 or can_recurse_worst_case(Dx, R, wrem, h)
    lo, hi = mpf(0), 2 * r2
    for _ in range(120):
        m = (lo + hi) / 2
        if test(m, Dec()):
            lo = m
        else:
            hi = m
    if lo <= 0 or D.cmp(lo, "<=", mpf(10) ** -40):
        return mpf(0), (mpf(0), mpf(0), r1), (mpf(0), mpf(0), r2)
    w1, w2 = mpmath.sqrt(4 * r1 - lo * lo), mpmath.sqrt(4 * r2 - lo * lo)
    return lo, (w1 / 2, lo / 2, r1), (w - w2 / 2, lo / 2, r2)


def _c13_r12_r456_r3(D, la, r):
    """six_disk_full_cover_r12_r456_r3_recursion (six_disk_full_cover.cu:168-193)."""
    w12 = two_disks_maximize_height(D, r[0], r[1], mpf(1))
    if w12 <= 0:
        return False
    w456 = nd_maximize_covered(D, mpf(1), [r[3], r[4], r[5]])
    if w456 <= 0:
        return False
    wrem = la - (w12 + w456)
    if D.cmp(wrem, "<=", 0):
        return True
    if D.cmp(4 * r[2], "<=", wrem * wrem):
        return False
    hrem = 1 - mpmath.sqrt(4 * r[2] - wrem * wrem)
    return rsb_can_recurse(D, max(rsb_R(la, r), mpf(0)), r[5], wrem, hrem)


def _c13_4_plus_2(D, la, r):
    """six_disk_full_cover_4_plus_2 (six_disk_full_cover.cu:27-104): r1, r2 at
    the bottom, r3, r4 in the top corners, r5 and r6 over the two gaps."""
    height, c1, c2 = max_height_strip(D, la, r[0], r[1])
    if height <= 0:
        return False
    hrem = 1 - height
    if D.cmp(r[3], "<=", hrem * hrem / 4) or D.cmp(r[2], "<=", hrem * hrem / 4):
        return False
    w4, w3 = mpmath.sqrt(r[3] - hrem * hrem / 4), mpmath.sqrt(r[2] - hrem * hrem / 4)
    y = 1 - hrem / 2
    c4, c3 = (w4, y, r[3]), (la - w3, y, r[2])
    ok34, p34, q34 = circle_intersection(D, c3, c4)
    if not ok34:
        return False
    up_top, low_top = (q34, p34) if D.cmp(p34[1], "<", q34[1]) else (p34, q34)
    ok12, p12, q12 = circle_intersection(D, c1, c2)
    if not ok12:
        return _trap(D)
    up_bottom = q12 if D.cmp(p12[1], "<", q12[1]) else p12
    ok14, p14, q14 = circle_intersection(D, c1, c4)
    ok23, p23, q23 = circle_intersection(D, c2, c3)
    if not (ok14 and ok23):
        return False
    top_c = (up_top[0], (1 + up_top[1]) / 2)
    top = [dist2(top_c, (2 * w4, mpf(1))), dist2(top_c, (la - 2 * w3, mpf(1))), dist2(top_c, up_top)]
    bottom_c = ((low_top[0] + up_bottom[0]) / 2, (low_top[1] + up_bottom[1]) / 2)
    bottom = [dist2(bottom_c, q14), dist2(bottom_c, p23), dist2(bottom_c, low_top), dist2(bottom_c, up_bottom)]
    if all(D.cmp(d, "<=", r[5]) for d in top):
        return all(D.cmp(d, "<=", r[4]) for d in bottom)
    if all(D.cmp(d, "<=", r[4]) for d in top):
        return all(D.cmp(d, "<=", r[5]) for d in bottom)
    return False


def _c13_2v3(D, la, r):
    """six_disk_full_cover_2v3_with_recursion_r6_center (six_disk_full_cover.cu:106-166)."""
    height, c1, c2 = max_height_strip_wc(D, r[0], r[1], max(rsb_R(la, r), mpf(0)), la)
    if height <= 0:
        return False
    hrem = 1 - height
    if D.cmp(4 * r[2], "<=", hrem * hrem) or D.cmp(4 * r[3], "<=", hrem * hrem):
        return False
    w3, w4 = mpmath.sqrt(4 * r[2] - hrem * hrem), mpmath.sqrt(4 * r[3] - hrem * hrem)
    wrem = la - (w3 + w4)
    if D.cmp(wrem, "<=", 0):
        return True
    if D.cmp(r[4], "<=", wrem * wrem / 4):
        return False
    dy5 = mpmath.sqrt(r[4] - wrem * wrem / 4)
    if D.cmp(2 * dy5 + height, ">=", 1):
        return True
    c5 = (w4 + wrem / 2, 1 - dy5, r[4])
    ok15, p15_, q15 = circle_intersection(D, c1, c5)
    ok25, p25, q25_ = circle_intersection(D, c2, c5)
    if not (ok15 and ok25):
        return False
    if D.cmp(q15[1], "<=", height) and D.cmp(p25[1], "<=", height):
        return True
    wrem6 = p25[0] - q15[0]
    hrem6 = max(q15[1], p25[1]) - height
    return D.cmp(wrem6 * wrem6 + hrem6 * hrem6, "<=", 4 * r[5])


def rsb_c13_six_disk(D, la, r):
    """C13, six_disk_full_cover (six_disk_full_cover.cu:195-200)."""
    if _c13_r12_r456_r3(D, la, r):
        return True, 1
    if _c13_4_plus_2(D, la, r):
        return True, 2
    if _c13_2v3(D, la, r):
        return True, 3
    return False, 0


RSB_STEPS = {6: rsb_c1_even_split, 7: rsb_c2_uneven_split, 8: rsb_c3_multi_disk_strip, 9: rsb_c4_wall_building, 10: rsb_c5_r1_in_corner, 11: rsb_c6_opposite_corners, 14: rsb_c9_l_shaped, 15: rsb_c10_r1_corner_wall_building, 16: rsb_c11_advanced_strip, 17: rsb_c12_five_disk, 18: rsb_c13_six_disk, 12: rsb_c7_vertical_strip, 13: rsb_c8_two_by_two}


def rsb_steps(la, r):
    """The pinned leaf-chain steps on the point cell: {step: (verdict, strategy, slack)}
    with the step numbering of the oracle's per-step masks (C1 = 6, ...)."""
    la = mpf(la)
    r = [mpf(x) for x in r]
    out = {}
    for k, f in RSB_STEPS.items():
        D = Dec()
        v, s = f(D, la, r)
        out[k] = (bool(v), s, D.slack)
    return out


# ------------------------------------------ tight_rectangle leaf criteria
# Point leaves (lambda, r1 >= r2 >= r3 >= r4) as the enumeration forms them
# (tight_rectangle.cu:217-246): W = critical_ratio(lambda) * lambda,
# R4 = W - r1 - r2 - r3 and R = R4 - r4, both at least 0.

def tr_leaf_weights(la, r):
    W = tr_critical_ratio(la) * la
    R4 = W - r[0] - r[1] - r[2]
    return R4, R4 - r[3]


def _tbt_r1_large(D, la, r, R):
    """two_by_two_strategies.cu:29-51: r1 spans the height at width
    sqrt(4 r1 - 1) on the left; r2, r3, r4 stacked in the strip to its right
    cover heights sqrt(4 ri - wrem^2); recursion on the rest of that strip."""
    w1 = mpmath.sqrt(4 * r[0] - 1)
    wrem = la - w1
    if D.cmp(4 * r[3], "<=", wrem * wrem):
        return False
    if min(r[1], r[2]) < r[3]:  # unsorted radii: the reference's square root is NaN, its tests false
        return False
    htot = sum(mpmath.sqrt(4 * x - wrem * wrem) for x in (r[1], r[2], r[3]))
    if D.cmp(htot, ">=", 1):
        return True
    return tr_can_recurse(D, wrem, 1 - htot, min(r[3], R), R)


def tbt_lambda_small(D, la, r, R):
    """two_by_two_strategies.cu:53-83: r1's square in a corner (r1_large when
    its side reaches the height), r2 in the strip to its right, r4 in the strip
    above it, r3 covers the rectangle left in the far corner."""
    side = mpmath.sqrt(2 * r[0])
    if D.cmp(side, ">=", 1):
        return _tbt_r1_large(D, la, r, R)
    wrem, hrem = la - side, 1 - side
    a, b = 4 * r[1] - wrem * wrem, 4 * r[3] - hrem * hrem
    if D.cmp(a, "<=", 0) or D.cmp(b, "<=", 0):
        return False
    h2, w4 = min(mpmath.sqrt(a), side), min(mpmath.sqrt(b), side)
    h, w = 1 - h2, la - w4
    return D.cmp(4 * r[2], ">=", h * h + w * w)


def tbt_lambda_big(D, la, r, R):
    """two_by_two_strategies.cu:85-111: two stacked pairs side by side, the
    best of the three pairings, recursion on the rest of the width."""
    def th(a, b):
        return two_disks_maximize_height(D, a, b, mpf(1))
    w = max(th(r[0], r[1]) + th(r[2], r[3]), th(r[0], r[3]) + th(r[1], r[2]), th(r[0], r[2]) + th(r[1], r[3]))
    if w == 0:  # every pairing decided 0 (its comparisons are recorded)
        return False
    return tr_can_recurse(D, la - w, mpf(1), min(R, r[3]), R)


def tr_two_by_two(D, la, r, R):
    """two_by_two_strategies (:113-115).  Strategies 1 lambda_small, 2 lambda_big."""
    if tbt_lambda_small(D, la, r, R):
        return True, 1
    return (True, 2) if tbt_lambda_big(D, la, r, R) else (False, 0)


def three_disk_big_strip(D, la, r1, r2, r3, ub_r4, R4):
    """r1_r2_r3_strategies.cu:101-121: r2 and r3 cover a full-height strip
    of width S (S^2 = 2 (r2 + r3) - 1/4 - 4 (r2 - r3)^2), r1 spans the
    remaining width, recursion on the rest of the height."""
    s_sq = 2 * r2 + 2 * r3 - mpf(1) / 4 - 4 * (r2 - r3) ** 2
    if D.cmp(s_sq, "<=", 0):
        return False
    la_rem = la - mpmath.sqrt(s_sq)
    h1_sq = 4 * r1 - la_rem * la_rem
    if D.cmp(h1_sq, "<=", 0):
        return False
    if D.cmp(h1_sq, ">=", 1):
        return True
    return tr_can_recurse(D, la_rem, 1 - mpmath.sqrt(h1_sq), ub_r4, R4)


def three_disk_small_corner(D, la, r1, r2, r3, ub_r4, R4):
    """r1_r2_r3_strategies.cu:123-165: r3's square in a corner, r1 above it
    and r2 beside it (or the other way round when r2 cannot reach half the
    remaining width), recursion on the rectangle left.  The square root of
    a negative difference is NaN in the reference, and a NaN width makes its
    can_recurse false."""
    c3 = mpmath.sqrt(2 * r3)
    if D.cmp(c3, ">=", 1):
        return False
    la_rem, one_rem = la - c3, 1 - c3
    if D.cmp(la_rem / 2, "<=", mpmath.sqrt(r2)):
        a = 4 * r1 - one_rem * one_rem
        if D.cmp(a, "<", 0):
            return False
        b = 4 * r2 - la_rem * la_rem
        if D.cmp(b, "<", 0):
            return False
        return tr_can_recurse(D, la - mpmath.sqrt(a), 1 - mpmath.sqrt(b), ub_r4, R4)
    if D.cmp(la_rem / 2, ">=", mpmath.sqrt(r1)):
        return False
    if D.cmp(one_rem, ">=", 2 * mpmath.sqrt(r2)):
        return False
    c2 = mpmath.sqrt(4 * r2 - one_rem * one_rem)
    c1 = mpmath.sqrt(4 * r1 - la_rem * la_rem)
    return tr_can_recurse(D, la - c2, 1 - c1, ub_r4, R4)


def _pocket_X(D, la, r1, r2, r3):
    """The three-disk pocket placement (r1_r2_r3_strategies.cu:167-186 and
    :228-244): r1 spans the height at the left, r2 below-right of it, r3 in
    the corner above r2; returns (S1, S3, X = lambda - S1 - S3) or None when a
    disk cannot reach."""
    if D.cmp(r1, "<=", mpf(1) / 4):
        return None
    s1 = 2 * mpmath.sqrt(r1 - mpf(1) / 4)
    h2_rest = (la - s1) ** 2 / 4
    if D.cmp(r2, "<=", h2_rest):
        return None
    h2 = 2 * mpmath.sqrt(r2 - h2_rest)
    s3_rest = (1 - h2) ** 2 / 4
    if D.cmp(r3, "<=", s3_rest):
        return None
    s3 = 2 * mpmath.sqrt(r3 - s3_rest)
    return s1, s3, la - s1 - s3


def three_disk_r1_r3_pocket(D, la, r1, r2, r3, ub_r4, R4):
    """r1_r2_r3_strategies.cu:167-206: the pocket placement; the X x Y
    rectangle between r1 and r3 is left to recursion."""
    p = _pocket_X(D, la, r1, r2, r3)
    if p is None:
        return False
    s1, s3, X = p
    if D.cmp(X, "<=", 0):
        return True
    y_rest = (la - s1 / 2 - s3) ** 2
    if D.cmp(r1, "<=", y_rest):
        return False
    Y = mpf(1) / 2 - mpmath.sqrt(r1 - y_rest)
    if D.cmp(Y, "<=", 0):
        return True
    return tr_can_recurse(D, X, Y, ub_r4, R4)


def manual_three_disk_pocket(D, la, r1, r2, r3):
    """r1_r2_r3_strategies.cu:208-247: the manually proved neighbourhood of
    the three-disk worst case (lambda <= 1.05, each ri within 0.001, 0.002,
    0.003 of r*(lambda) = lambda^2/16 + 5/32 + 9/(256 lambda^2), the pocket's
    X at most 0.047); its constants are the doubles the reference uses."""
    if D.cmp(la, ">", mpf(1.0499999999999998)):
        return False
    lasq = la * la
    rstar = lasq / 16 + mpf(0.15625) + mpf(0.03515625) / lasq
    for x, eps in ((r1, mpf(0.0009999999999999998)), (r2, mpf(0.0019999999999999996)),
                   (r3, mpf(0.0029999999999999996))):
        d = rstar - x
        if D.cmp(d, "<", -eps) or D.cmp(d, ">", eps):
            return False
    p = _pocket_X(D, la, r1, r2, r3)
    if p is None:
        return False
    return not D.cmp(p[2], ">", mpf(0.046999999999999993))


def tr_three_disk(D, la, r, R4):
    """r1_r2_r3_strategies(vars, R4) (:249-260).  Strategies: 1 big strip, 2
    small corner, 3 r1-r3 pocket, 4 the manual pocket."""
    r1, r2, r3, ub_r4 = r
    fns = (lambda: three_disk_big_strip(D, la, r1, r2, r3, ub_r4, R4),
           lambda: three_disk_small_corner(D, la, r1, r2, r3, ub_r4, R4),
           lambda: three_disk_r1_r3_pocket(D, la, r1, r2, r3, ub_r4, R4),
           lambda: manual_three_disk_pocket(D, la, r1, r2, r3))
    for s, f in enumerate(fns, 1):
        if f():
            return True, s
    return False, 0


def _gap_ok(D, gap, r1, w1, r2, w2, r4):
    """can_realize_gap_width (r1_r2_large_r3_r4_gaps_strategy.cu:29-57): r1
    and r2 span the height side by side at widths w1, w2 with a gap between
    them; r4, centred on the gap at the bottom edge's height sqrt(r4 - gap^2/4),
    reaches the lower intersection point of r1 and r2."""
    y4_sq = r4 - gap * gap / 4
    if D.cmp(y4_sq, "<=", 0):
        return False
    d2 = (w2 / 2 + gap + w1 / 2) ** 2
    rd = r1 - r2
    if D.cmp(2 * (r1 + r2) / d2, "<", 1 + rd * rd / (d2 * d2)):
        return False
    # the circles' centres are level, so the first point is the lower one
    ok, p, _ = circle_intersection(Dec(), (w1 / 2, mpf(1) / 2, r1), (w2 / 2 + gap + w1, mpf(1) / 2, r2))
    return D.cmp(dist2((gap / 2 + w1, mpmath.sqrt(y4_sq)), p), "<=", r4)


def tr_gaps(D, la, r, R):
    """r1_r2_large_r3_r4_gaps_strategy (:59-98): r1 and r2 span the height
    with the widest gap between them that r4 closes (the reference's
    bisection over [0, sqrt(r1) + sqrt(r2)], on exact midpoints; the tests of
    its first levels, where the bracket is still wide, count towards the
    slack), recursion on the rest of the width."""
    r1, r2, r4 = r[0], r[1], r[3]
    if D.cmp(4 * r1, "<=", 1) or D.cmp(4 * r2, "<=", 1):
        return False, 0
    w1, w2 = mpmath.sqrt(4 * r1 - 1), mpmath.sqrt(4 * r2 - 1)
    lo, hi = mpf(0), mpmath.sqrt(r1) + mpmath.sqrt(r2)
    wide = hi * mpf("1e-6")
    for _ in range(140):
        mid = (lo + hi) / 2
        if _gap_ok(D if hi - lo > wide else Dec(), mid, r1, w1, r2, w2, r4):
            lo = mid
        else:
            hi = mid
    wrem = la - (w1 + lo + w2)
    if D.cmp(wrem, "<=", 0):
        return True, 1
    return (True, 2) if tr_can_recurse(D, wrem, mpf(1), min(R, r[3]), R) else (False, 0)


TR_LEAF = {0: "two_by_two", 1: "three_disk", 2: "gaps"}


def tr_leaf(la, r):
    """The three leaf criteria on the point leaf: {bit: (verdict, strategy, slack)}
    with the bit numbering of the oracle's leaf masks."""
    la = mpf(la)
    r = [mpf(x) for x in r]
    R4, R = tr_leaf_weights(la, r)
    out = {}
    for k, f in ((0, lambda D: tr_two_by_two(D, la, r, R)), (1, lambda D: tr_three_disk(D, la, r, R4)),
                 (2, lambda D: tr_gaps(D, la, r, R))):
        D = Dec()
        try:
            v, s = f(D)
        except ZeroDivisionError:
            v, s, D.slack = False, 0, mpf(0)
        out[k] = (bool(v), s, D.slack)
    return out


# ------------------------------- rectangle_size_bound leaf step C14 (seven disks)
# seven_disk_strategies.cu: r1..r6 and an unknown seventh disk r7 whose range
# [lower bound, min(r6, R)] is cut down by recursions and then checked in 16
# pieces.  The reference's bisections become bisections on exact midpoints;
# the tests of their first levels, where the bracket is still wide, count
# towards the slack.  r7 ranges are [lo, hi] lists; an empty range has lo > hi.

C14_ITER = 80        # 2^-80 of a bracket: far below any compared slack
C14_WIDE = mpf("1e-8")


def _bisect(D, test, lo, hi, record=True):
    """The reference's `mid = (lo + hi) / 2; test(mid) ? lo = mid : hi = mid`
    on exact midpoints; returns (lo, the value test returned at the last
    passing midpoint or None)."""
    wide = (hi - lo) * C14_WIDE
    last = None
    for _ in range(C14_ITER):
        mid = (lo + hi) / 2
        got = test(D if record and hi - lo > wide else Dec(), mid)
        if got:
            lo, last = mid, got
        else:
            hi = mid
    return lo, last


def three_disks_height(D, w, t):
    """three_disks_maximize_height(t1, t2, t3, w) (maximize_covered.cu:198-237):
    the greatest height at which the three disks side by side span w, below
    2 sqrt(t3), the bound the reference's bisection starts from."""
    hi = 2 * mpmath.sqrt(t[2])
    if D.cmp(_chords(t, mpf(0)), "<", w):
        return mpf(0)
    if D.cmp(_chords(t, hi), ">=", w):
        return hi
    lo, _ = _bisect(D, lambda D_, h: D_.cmp(_chords(t, h), ">=", w), mpf(0), hi)
    return lo


def bound_critical_ratio(D, rem, largest, w, h):
    """bound_critical_ratio (can_recurse.cu:90-160): the best covering ratio
    the remaining weight reaches on a w x h rectangle (size-bounded with a
    square or strip of the required side, the large size-bound corollary, the
    worst case), INF if none."""
    if D.cmp(rem, "<", 0):
        return INF
    if D.cmp(w, "<=", 0) or D.cmp(h, "<=", 0):
        return mpf(0)
    smin, smax = min(w, h), max(w, h)
    best = INF
    area = w * h
    req = mpmath.sqrt(largest / UB_DISK_WEIGHT)
    if D.cmp(smin, ">=", req):
        return CR if D.cmp(CR * area, "<=", rem) else INF
    if D.cmp(req, ">", smax):
        if D.cmp(mpmath.sqrt(rem / CR), ">=", req):
            best = req * req * CR / area
    else:
        need = CR * req * smax
        if D.cmp(need, "<=", rem):
            best = need / area
    sig = (largest / (smin * smin)) ** 2
    cor = mpmath.sqrt(mpmath.sqrt(sig + 1) + 1) / 2
    if D.cmp(cor, "<", C_WORST):
        cor = C_WORST
    if D.cmp(cor, "<", best) and D.cmp(cor * area, "<=", rem):
        best = cor
    cr = tr_critical_ratio(smax / smin)
    if D.cmp(cr, "<", best) and D.cmp(cr * area, "<=", rem):
        best = cr
    return best


def recursion_bound_size(D, rem, w, h):
    """recursion_bound_size (can_recurse.cu:162-219): the largest disk weight
    recursion on a w x h rectangle tolerates with the weight rem (INF: any)."""
    if D.cmp(rem, "<=", 0):
        if rem < 0:
            return mpf(0)
        return INF if (w <= 0 or h <= 0) else mpf(0)
    if D.cmp(w, "<=", 0) or D.cmp(h, "<=", 0):
        return INF
    smin, smax = min(w, h), max(w, h)
    max_cr = rem / (w * h)
    if D.cmp(tr_critical_ratio(smax / smin), "<=", max_cr):
        return INF
    sb_area = rem / CR
    side = sb_area / smax
    if D.cmp(side, "<", smin):
        return mpf(0)
    if D.cmp(side, ">", smax):
        side = mpmath.sqrt(sb_area)
    best = UB_DISK_WEIGHT * side * side
    if D.cmp(max_cr, ">=", C_WORST):
        lam = 2 * max_cr + mpmath.sqrt(4 * max_cr * max_cr - 2)
        sig = max_cr * (lam - 2 / lam) * smin * smin
        if D.cmp(sig, ">=", best):
            best = sig
    return best


def _rest(x, y):
    """The lower bound of a remaining weight x - y, where y is r7's upper
    bound.  Where that bound is R itself (the last r7 piece) the difference is
    0 in exact arithmetic, but below 0 in the reference, which subtracts R's
    upper bound from R's lower bound: a negative infinitesimal stands in."""
    d = x - y
    return d if abs(d) > mpf("1e-40") * max(abs(x), abs(y)) else mpf("-1e-40") * max(abs(x), abs(y))


def _empty(D, r7):
    return D.cmp(r7[0], ">", r7[1])


def _tighten_lb(r7, x):
    if x > r7[0]:
        r7[0] = x


def _split_largest_r7(D, r7, rem, wf, hf, ws, hs):
    """restrict_r7_t_shaped_recursion_split_recursion_largest_r7 (:116-153):
    r7 as the largest disk of both parts; the range keeps only the r7 for
    which the split does not work."""
    def works(D_, x):
        cr = bound_critical_ratio(D_, rem, x, wf, hf)
        if cr == INF:
            return False
        return rsb_can_recurse(D_, rem - (cr * wf * hf + x), x, ws, hs)
    if works(D, r7[1]):
        r7[0] = INF
        return
    lo, _ = _bisect(D, works, r7[0], r7[1])
    _tighten_lb(r7, lo)


def _split_recursion(D, r7, r, rem, largest, wf, hf, ws, hs):
    """restrict_r7_t_shaped_recursion_split_recursion (:155-215)."""
    if largest > 5:
        _split_largest_r7(D, r7, rem, wf, hf, ws, hs)
        return
    cr = bound_critical_ratio(D, rem, r[largest], wf, hf)
    if cr == INF:
        return
    need = cr * wf * hf
    win = sum(r[largest:6])
    if D.cmp(win, ">", need):
        _tighten_lb(r7, recursion_bound_size(D, rem - win, ws, hs))
        return
    _tighten_lb(r7, recursion_bound_size(D, rem - (need + r7[1]), ws, hs))
    if not _empty(D, r7):
        lo, _ = _bisect(D, lambda D_, x: not D_.cmp(recursion_bound_size(D_, rem - (need + x), ws, hs), "<", x),
                        r7[0], r7[1])
        r7[0] = lo


def _place_r6(D, q, x1, x2, lower, upper):
    """restrict_r7_t_shaped_recursion_compute_r6 (:225-258): the disk q
    centred right of the two points, spanning their height; (centre, the x
    up to which it covers the strip, 0 if the placement fails)."""
    left = min(x1[0], x2[0])
    top, bot = max(x1[1], x2[1]), min(x1[1], x2[1])
    dx = q - (top - bot) ** 2 / 4
    if q == 0 or D.cmp(dx, "<=", 0):  # q = 0: r7 at its lower bound 0
        return None, mpf(0)
    p = (left + mpmath.sqrt(dx), (bot + top) / 2)
    dy = max((p[1] - lower) ** 2, (upper - p[1]) ** 2)
    sw = q - dy
    if D.cmp(sw, "<=", 0):
        return p, mpf(0)
    return p, p[0] + mpmath.sqrt(sw)


def _meets(D, a, b):
    """The two intersection points of two circles in no particular order (the
    caller picks by coordinate), None unless they definitely intersect; only
    the intersection test is recorded (operations.cu:39-51)."""
    d2 = dist2(a, b)
    rd = a[2] - b[2]
    if d2 == 0 or D.cmp(2 * (a[2] + b[2]) / d2, "<", 1 + rd * rd / (d2 * d2)):
        return None
    _, p, q = circle_intersection(Dec(), a, b)
    return p, q


def _intersect_or_trap(D, a, b):
    ok, p, q = circle_intersection(D, a, b)
    if not ok:
        _trap(D)
    return ok, p, q


def _t_shaped_r5_large(D, r7, la, r, c1, c3, c5, R6, w12, wrem, hrem, strip_w5, lower, upper, p43, p21):
    """restrict_r7_t_shaped_recursion_r5_large (:260-314)."""
    ok15, _, x15 = _intersect_or_trap(D, c1, c5)
    ok35, _, x35 = _intersect_or_trap(D, c3, c5)
    if not (ok15 and ok35):
        return
    left = min(x15[0], x35[0])
    whr = w12 - left
    if D.cmp(whr, "<=", 0):
        if rsb_can_recurse(D, R6, r[5], wrem, mpf(1)):
            r7[0] = INF
        return
    _split_recursion(D, r7, r, R6, 5, whr, hrem, wrem, mpf(1))
    _split_recursion(D, r7, r, R6, 5, wrem, mpf(1), whr, hrem)
    if _empty(D, r7):
        return
    R7 = R6 - r[5]
    if D.cmp(R7, "<=", 0):
        return
    x1, x2 = x15, x35
    if (D.cmp(x1[1], ">=", upper) or D.cmp(x1[1], "<=", lower) or D.cmp(x2[1], ">=", upper)
            or D.cmp(x2[1], "<=", lower)):
        x1, x2 = (strip_w5, upper), (strip_w5, lower)
    _, reach = _place_r6(D, r[5], x1, x2, lower, upper)
    if reach == 0 or D.cmp(reach, "<=", 0):  # 0: the placement failed
        return
    if D.cmp(reach, ">=", w12):
        _tighten_lb(r7, recursion_bound_size(D, R7, wrem, mpf(1)))
    else:
        whr = w12 - left  # the reference recomputes this from r5's intersections
        _split_recursion(D, r7, r, R7, 6, whr, hrem, wrem, mpf(1))
        _split_recursion(D, r7, r, R7, 6, wrem, mpf(1), whr, hrem)


def _t_shaped(D, r7, la, r, R):
    """restrict_r7_t_shaped_recursion (:316-510): r2's square in the lower
    left corner with r1 beside it, r3 and r4 stacked in the strip above, r5
    (and r6, r7) in the gap between; recursion on the rest."""
    w2 = mpmath.sqrt(2 * r[1])
    h12 = w2
    w1 = mpmath.sqrt(4 * r[0] - h12 * h12)
    w12 = w1 + w2
    h34 = two_disks_maximize_height(D, r[2], r[3], w12)
    if h34 == 0:
        return
    R5 = la * CR - (r[0] + r[1] + r[2] + r[3])
    hrem = 1 - (h12 + h34)
    wrem = la - w12
    if D.cmp(hrem, "<=", 0):
        if D.cmp(wrem, "<=", 0) or rsb_can_recurse(D, R5, r[4], wrem, mpf(1)):
            r7[0] = INF
        return
    if D.cmp(wrem, "<=", 0):
        if rsb_can_recurse(D, R5, r[4], la, hrem):
            r7[0] = INF
        return
    _split_recursion(D, r7, r, R5, 4, w12, hrem, wrem, mpf(1))
    _split_recursion(D, r7, r, R5, 4, wrem, mpf(1), w12, hrem)
    if _empty(D, r7):
        return
    R6 = R5 - r[4]
    if D.cmp(R6, "<=", 0):
        return
    x5 = r[4] - hrem * hrem / 4
    if D.cmp(x5, "<=", 0):
        return
    x5 = mpmath.sqrt(x5)
    strip_w5 = 2 * x5
    p5 = (x5, h12 + hrem / 2)
    p2 = (w2 / 2, h12 / 2)
    p1 = (w2 + w1 / 2, h12 / 2)
    x4 = mpmath.sqrt(r[3] - h34 * h34 / 4)
    y4 = 1 - h34 / 2
    p4 = (x4, y4)
    w3 = w12 - 2 * x4
    p3 = (w12 - w3 / 2, y4)
    c1, c2, c3, c4, c5 = (p1 + (r[0],)), (p2 + (r[1],)), (p3 + (r[2],)), (p4 + (r[3],)), (p5 + (r[4],))
    p43, p21 = (2 * x4, 1 - h34), (w2, h12)
    lower, upper = h12, 1 - h34
    if D.cmp(dist2(p5, p43), "<=", r[4]) and D.cmp(dist2(p5, p21), "<=", r[4]):
        _t_shaped_r5_large(D, r7, la, r, c1, c3, c5, R6, w12, wrem, hrem, strip_w5, lower, upper, p43, p21)
        return
    ok25, a25, x25 = _intersect_or_trap(D, c2, c5)
    ok45, a45, x45 = _intersect_or_trap(D, c4, c5)
    if not (ok25 and ok45):
        return
    if D.cmp(a25[0], ">=", x25[0]) or D.cmp(a45[0], ">=", x45[0]):
        return
    x1, x2 = x25, x45
    if (D.cmp(x1[1], ">=", upper) or D.cmp(x1[1], "<=", lower) or D.cmp(x2[1], ">=", upper)
            or D.cmp(x2[1], "<=", lower)):
        x1, x2 = (strip_w5, upper), (strip_w5, lower)
    left = min(x1[0], x2[0])
    whr = w12 - left
    _split_recursion(D, r7, r, R6, 5, whr, hrem, wrem, mpf(1))
    _split_recursion(D, r7, r, R6, 5, wrem, mpf(1), whr, hrem)
    if _empty(D, r7):
        return
    R7 = R6 - r[5]
    if D.cmp(R7, "<=", 0):
        return
    p6, reach = _place_r6(D, r[5], x1, x2, lower, upper)
    if reach == 0 or D.cmp(reach, "<=", 0):  # 0: the placement failed
        return
    c6 = p6 + (r[5],)
    ok16, _, x16 = _intersect_or_trap(D, c1, c6)
    ok36, _, x36 = _intersect_or_trap(D, c3, c6)
    if not (ok16 and ok36):
        return
    r6_right = False
    if D.cmp(dist2(p43, p6), "<=", r[5]) and D.cmp(dist2(p21, p6), "<=", r[5]):
        left = min(x16[0], x36[0])
        d71, d72 = x36, x16
        r6_right = True
    else:
        left = reach
        # the reference builds these points with the strip borders as x and
        # the covered x as y (a quirk kept on purpose)
        d71, d72 = (upper, left), (lower, left)
    whr = w12 - left
    if D.cmp(whr, "<=", 0):
        _tighten_lb(r7, recursion_bound_size(D, R7, wrem, mpf(1)))
    else:
        _split_recursion(D, r7, r, R7, 6, whr, hrem, wrem, mpf(1))
        _split_recursion(D, r7, r, R7, 6, wrem, mpf(1), whr, hrem)
    if _empty(D, r7):
        return
    R8 = _rest(R7, r7[1])
    if D.cmp(R8, "<=", 0):
        return
    if r6_right:
        ok13, a13, x13 = circle_intersection(D, c1, c3)
        if (ok13 and D.cmp(dist2(a13, p6), "<=", r[5]) and D.cmp(x13[1], ">=", lower)
                and D.cmp(x13[1], "<=", upper)):
            d71 = d72 = x13
    _, reach7 = _place_r6(D, r7[0], d71, d72, lower, upper)
    if reach7 == 0 or D.cmp(reach7, "<=", 0):
        return
    if D.cmp(reach7, ">=", la):
        _split_largest_r7(D, r7, R8, wrem, h12, wrem, h34)
        _split_largest_r7(D, r7, R8, wrem, h34, wrem, h12)
    if D.cmp(reach7, ">=", w12):
        _tighten_lb(r7, recursion_bound_size(D, R8, wrem, mpf(1)))
    else:
        whr = w12 - reach7
        _split_largest_r7(D, r7, R8, whr, hrem, wrem, mpf(1))
        _split_largest_r7(D, r7, R8, wrem, mpf(1), whr, hrem)


def _four_disk_strip_r1_square(D, r7, la, r, R):
    """restrict_r7_four_disk_strip_r1_square (:1222-1265): r2..r5 cover a
    strip along the top, r1's square sits below it in the left corner, r6
    covers what is left beside the square."""
    h = nd_maximize_covered(D, la, [r[1], r[2], r[3], r[4]])
    if h == 0:
        return
    wh1 = mpmath.sqrt(2 * r[0])
    hrem = 1 - h
    hbl = 1 - (h + wh1)
    if D.cmp(hbl, "<=", 0):
        w1 = mpmath.sqrt(4 * r[0] - hrem * hrem)
        R6 = la * CR - sum(r[:5])
        if rsb_can_recurse(D, R6, r[5], la - w1, hrem):
            r7[0] = INF
        return
    wrem = la - wh1
    if D.cmp((wh1 * wh1 + hbl * hbl) / 4, "<=", r[5]):
        _tighten_lb(r7, recursion_bound_size(D, R, wrem, hrem))
        return
    w6 = 4 * r[5] - hbl * hbl
    if D.cmp(w6, "<=", 0):
        return
    _split_largest_r7(D, r7, R, wrem, hrem, wh1 - mpmath.sqrt(w6), hbl)


def _r7_range(D, la, r, R, Rub, wrem):
    """compute_r7_range (:1150-1190): [0, min(r6, R)], cut where the rest
    could be covered size-bounded or by the (large size-bound) worst case."""
    r7 = [mpf(0), min(r[5], Rub)]
    eff = R / wrem
    if D.cmp(eff, ">=", CR):
        msw = R / CR
        if D.cmp(msw, ">=", 1):
            return [mpf(1), mpf(0)]
        _tighten_lb(r7, UB_DISK_WEIGHT * msw * msw)
    if D.cmp(eff, ">=", C_WORST) and not _empty(D, r7):
        if D.cmp(tr_critical_ratio(1 / wrem), "<=", eff):
            return [mpf(1), mpf(0)]
        lam = 2 * eff + mpmath.sqrt(4 * eff * eff - 2)
        _tighten_lb(r7, eff * (lam - 2 / lam) * wrem * wrem)
    return r7


def _r7_in_corner(D, R, r7, w):
    """r7_in_corner_and_recurse (:38-47)."""
    h = 4 * r7[0] - w * w
    if D.cmp(h, "<", 0):
        return False
    return rsb_can_recurse(D, _rest(R, r7[1]), r7[1], w, 1 - mpmath.sqrt(h))


def _top_strip_can_cover(D, la, th, rem4, largest, R7):
    """three_disk_top_strip_can_cover (:49-114): three disks cover a strip
    of height th along the top; recursion (size-bounded, else worst case with
    the tolerated skew) covers part of the rest's width, two pairs of the
    remaining four disks the other part."""
    if th == 0:
        return False
    hrem = 1 - th
    if D.cmp(hrem, "<=", 0):
        return True
    wsb = R7 / CR / hrem
    if size_bound_ok(D, largest, min(wsb, hrem)):
        wrem = la - wsb
    else:
        wwc = R7 / C_WORST / hrem
        if wwc > 0 and D.cmp(max(wwc, hrem) / min(wwc, hrem), ">", TOL_LAMBDA_RSB):
            if D.cmp(wwc, ">", hrem):
                wwc = hrem * TOL_LAMBDA_RSB
            else:
                x = 2 * R7 - hrem * hrem / 2
                wwc = mpf(0) if D.cmp(x, "<=", 0) else mpmath.sqrt(x)
        wrem = la - wwc
    bot = two_disks_maximize_height(D, rem4[1], rem4[2], wrem)
    if bot == 0:
        return False
    if D.cmp(bot, ">=", hrem):
        return True
    h2 = two_disks_maximize_height(D, rem4[0], rem4[3], wrem)
    if h2 == 0:
        return False
    return D.cmp(h2, ">=", hrem - bot)


def _three_disk_top_strip(D, la, r, R, r7, heights):
    """three_disk_top_strip (:512-546) over every triple of the seven disks
    (r7 at its lower bound); heights caches the triples without r7."""
    disks = list(r) + [r7[0]]
    R7 = _rest(R, r7[1])
    for i in range(5):
        for j in range(i + 1, 6):
            for k in range(j + 1, 7):
                rest = [disks[g] for g in range(7) if g not in (i, j, k)]
                key = (i, j, k)
                if k < 6 and key in heights:
                    th = heights[key]
                else:
                    th = three_disks_height(D, la, [disks[i], disks[j], disks[k]])
                    if k < 6:
                        heights[key] = th
                if _top_strip_can_cover(D, la, th, rest, r7[1], R7):
                    return True
    return False


def _six_222(D, la, r, R, r7):
    """six_disks_222_with_recursion (:1192-1246): two stacked pairs side by
    side, r5 and r6 stacked in the rest of the width; the two rectangles left
    above r6 to recursion (r7 covering one of them, or the smaller group of a
    greedy split)."""
    c12 = two_disks_maximize_height(D, r[0], r[1], mpf(1))
    c34 = two_disks_maximize_height(D, r[2], r[3], mpf(1))
    if c34 == 0:
        return False
    wrem = la - (c12 + c34)
    h2 = 4 * r[5] - wrem * wrem
    if D.cmp(h2, "<=", 0):
        return False
    h1, h2 = mpmath.sqrt(4 * r[4] - wrem * wrem), mpmath.sqrt(h2)
    hrem = 1 - (h1 + h2)
    if D.cmp(hrem, "<=", 0):
        return True
    wtop = r[5] - (1 - (h1 + h2 / 2)) ** 2
    if D.cmp(wtop, "<=", 0):
        return False
    wrect = (wrem - 2 * mpmath.sqrt(wtop)) / 2
    if D.cmp((wrect * wrect + hrem * hrem) / 4, "<=", r7[0]):
        return rsb_can_recurse(D, _rest(R, r7[1]), r7[1], wrect, hrem)
    return rsb_can_recurse(D, _rest(R / 2, r7[1] / 2), r7[1], wrect, hrem)


def _r435_strip(D, la, r, R, r7):
    """r435_strip_covering_r12_gap_r67_right (:548-718): r3, r4, r5 cover a
    vertical strip, r1 above r2 the widest strip next to it whose left
    intersection r3 still covers, r6 and r7 stacked at the right."""
    c435 = three_disks_height(D, mpf(1), [r[2], r[3], r[4]])
    if c435 == 0:
        return False
    sq = c435 * c435
    h3, h4 = 4 * r[2] - sq, 4 * r[3] - sq
    if D.cmp(h3, "<=", 0) or D.cmp(h4, "<=", 0):
        return False
    h3, h4 = mpmath.sqrt(h3), mpmath.sqrt(h4)
    pos3 = (c435 / 2, h3 / 2 + h4)
    s1, s2 = mpmath.sqrt(r[0]), mpmath.sqrt(r[1])
    if D.cmp(s1 + s2, "<=", mpf(1) / 2):
        return False
    ub = 2 * s2
    wr = la - c435
    if D.cmp(wr, "<", ub):
        p1y, p2y = r[0] - wr * wr / 4, r[1] - wr * wr / 4
        if D.cmp(p1y, ">", 0) and D.cmp(p2y, ">", 0):
            x = wr / 2 + c435
            ok, a, b = circle_intersection(D, (x, 1 - mpmath.sqrt(p1y), r[0]), (x, mpmath.sqrt(p2y), r[1]))
            if ok and D.cmp(a[0], "<", b[0]) and D.cmp(dist2(pos3, a), "<=", r[2]):
                rem_h = 1 - (2 * mpmath.sqrt(p1y) + 2 * mpmath.sqrt(p2y))
                # la + x (not la - x) as the reference writes it
                if rsb_can_recurse(D, R + r[5], r[5], la + b[0], rem_h):
                    return True
        ub = wr

    def test(D_, mid):
        hh2 = r[1] - mid * mid / 4
        if D_.cmp(hh2, "<=", 0):
            return None
        hh1 = r[0] - mid * mid / 4
        if hh1 < 0:  # NaN in the reference: no intersection
            return None
        x = mid / 2 + c435
        ok, a, b = circle_intersection(D_, (x, 1 - mpmath.sqrt(hh1), r[0]), (x, mpmath.sqrt(hh2), r[1]))
        if ok and D_.cmp(a[0], "<", b[0]) and D_.cmp(dist2(pos3, a), "<=", r[2]):
            return b
        return None
    w12, best = _bisect(D, test, mpf(0), ub)
    if best is None:
        return False
    w67 = la - (c435 + w12)
    h7, h6 = 4 * r7[0] - w67 * w67, 4 * r[5] - w67 * w67
    if D.cmp(h7, ">", 0) and D.cmp(h6, ">", 0):
        h7, h6 = mpmath.sqrt(h7), mpmath.sqrt(h6)
        pos6 = (la - w67 / 2, 1 - h7 - h6 / 2)
        if D.cmp(dist2(best, pos6), "<=", r[5]):
            if rsb_can_recurse(D, _rest(R, r7[1]), r7[1], w67, 1 - (h7 + h6)):
                return True
    wide = la - best[0]
    h7, h6 = 4 * r7[0] - wide * wide, 4 * r[5] - wide * wide
    if D.cmp(h7, ">", 0) and D.cmp(h6, ">", 0):
        hr = 1 - (mpmath.sqrt(h6) + mpmath.sqrt(h7))
        wrec = w67 if D.cmp(hr, "<=", best[1]) else wide
        if rsb_can_recurse(D, _rest(R, r7[1]), r7[1], wrec, hr):
            return True
    return False


def _different_height(D, la, r, R, r7):
    """different_height_r12_and_two_by_two (:1248-1303): r2's square with r1
    above it (4 r1 - w12, unsquared, as the reference writes it: Q7), pairs
    r3/r4 and r5/r6 stacked beside; r7 across the higher gap, the rest priced
    by the large size-bound corollary."""
    w12 = mpmath.sqrt(2 * r[1])
    if D.cmp(4 * r[0], "<", w12):  # NaN in the reference, every test false
        return False
    h12 = mpmath.sqrt(4 * r[0] - w12) + w12
    wrem = la - w12
    h3456 = two_disks_maximize_height(D, r[2], r[3], wrem) + two_disks_maximize_height(D, r[4], r[5], wrem)
    a, b = 1 - h12, 1 - h3456
    if D.cmp(a, "<", b):
        hhigh, hlow, whigh, wlow = b, a, wrem, w12
    else:
        hhigh, hlow, whigh, wlow = a, b, w12, wrem
    wr7 = 4 * r7[0] - hhigh * hhigh
    if D.cmp(wr7, "<=", 0):
        return False
    whigh -= mpmath.sqrt(wr7)
    R8 = _rest(R, r7[1])
    if D.cmp(R8, "<=", 0):
        return False
    if D.cmp(whigh, "<=", 0):
        return rsb_can_recurse(D, R8, r7[1], whigh + wlow, hlow)
    s = min(hlow, wlow, whigh)
    sig = (r7[1] / (s * s)) ** 2
    cr = mpmath.sqrt(mpmath.sqrt(sig + 1) + 1) / 2
    if D.cmp(cr, "<", C_WORST):
        cr = C_WORST
    area = r7[1] / (cr * hhigh) * (hhigh - hlow) + whigh * hhigh + wlow * hlow
    return D.cmp(R8, ">=", area * cr)


def _five_disks_23(D, la, t1, t2, m1, m2, m3):
    """five_disks_maximize_height_23 (:975-1063): t1, t2 side by side at the
    bottom, m2, m1, m3 in a row above them as high as possible with m1
    covering the bottom pair's upper intersection and the middle row's lower
    intersections inside the bottom disks; 0 if none."""
    bot, b1, b2 = max_height_strip(D, la, t1, t2)
    if bot == 0:
        return mpf(0)
    pq = _meets(D, b1, b2)
    if pq is None:
        return _trap(D) or mpf(0)
    up = pq[1] if D.cmp(pq[0][1], "<=", pq[1][1]) else pq[0]

    def lower(D_, a, b):
        pq_ = _meets(D_, a, b)
        if pq_ is None:
            return None
        return pq_[0] if D_.cmp(pq_[0][1], "<=", pq_[1][1]) else pq_[1]

    def test(D_, hm):
        sq = hm * hm
        w = [4 * m - sq for m in (m1, m2, m3)]
        if any(D_.cmp(x, "<=", 0) for x in w):
            return False
        w1, w2, w3 = (mpmath.sqrt(x) for x in w)
        if D_.cmp(w1 + w2 + w3, ">=", la):
            return True
        wr = la - (w2 + w3)
        h1 = m1 - wr * wr / 4
        if D_.cmp(h1, "<=", 0):
            return False
        h1 = mpmath.sqrt(h1)
        y23 = bot + hm / 2
        c1 = ((w2 + la - w3) / 2, bot + hm - h1, m1)
        if D_.cmp(dist2(c1, up), ">", m1):
            return False
        l12 = lower(D_, c1, (w2 / 2, y23, m2))
        l13 = lower(D_, c1, (la - w3 / 2, y23, m3))
        if l12 is None or l13 is None:
            return False
        return D_.cmp(dist2(l12, b1), "<=", t1) and D_.cmp(dist2(l13, b2), "<=", t2)
    lo, _ = _bisect(D, test, mpf(0), 2 * mpmath.sqrt(m3))
    return mpf(0) if lo == 0 else bot + lo


PAIRINGS_3x2 = [((0, 1), (2, 3), (4, 5)), ((0, 1), (2, 4), (3, 5)), ((0, 1), (2, 5), (3, 4)),
                ((0, 2), (1, 3), (4, 5)), ((0, 2), (1, 4), (3, 5)), ((0, 2), (1, 5), (3, 4)),
                ((0, 3), (1, 2), (4, 5)), ((0, 3), (1, 4), (2, 5)), ((0, 3), (1, 5), (2, 4)),
                ((0, 4), (1, 2), (3, 5)), ((0, 4), (1, 3), (2, 5)), ((0, 4), (1, 5), (2, 3)),
                ((0, 5), (1, 2), (3, 4)), ((0, 5), (1, 3), (2, 4)), ((0, 5), (1, 4), (2, 3))]


def _r56_center(D, w, r):
    """six_disks_can_cover_width_r56_center (six_disks_cover_width.cu:29-90):
    pairs r1/r4 at the bottom and r2/r3 at the top of a strip of width w,
    r5 and r6 in the band between at its left and right."""
    if w == 0:  # 2 / w is infinite in the reference: the first pair's height is 0
        return False
    h14 = two_disks_maximize_height(D, r[0], r[3], w)
    h23 = two_disks_maximize_height(D, r[1], r[2], w)
    if h14 == 0 or h23 == 0:
        return False
    hrem = 1 - (h14 + h23)
    q = hrem * hrem / 4
    x5, x6 = r[4] - q, r[5] - q
    if D.cmp(x5, "<=", 0) or D.cmp(x6, "<=", 0):
        return False
    x5, x6 = mpmath.sqrt(x5), mpmath.sqrt(x6)
    if D.cmp(2 * (x5 + x6), ">=", w):
        return True
    x6 = w - x6
    y56, y14, y23 = h14 + hrem / 2, h14 / 2, 1 - h23 / 2
    a1, a2 = r[0] - h14 * h14 / 4, r[1] - h23 * h23 / 4
    if a1 < 0 or a2 < 0:  # NaN in the reference
        return False
    w1, w2 = mpmath.sqrt(a1), mpmath.sqrt(a2)
    p1, p2 = (w - w1, y14), (w - w2, y23)
    ps1, ps2 = (w - 2 * w1, h14), (w - 2 * w2, 1 - h23)
    ok, a, b = circle_intersection(D, p1 + (r[0],), p2 + (r[1],))
    if not ok or not D.cmp(a[0], "<", b[0]):
        return False
    p5, p6 = (x5, y56), (x6, y56)
    return (D.cmp(dist2(p5, ps1), "<=", r[4]) and D.cmp(dist2(p5, ps2), "<=", r[4])
            and D.cmp(dist2(p5, a), "<=", r[4]) and D.cmp(dist2(p6, b), "<=", r[5]))


def six_disks_covered_width(D, la, r):
    """six_disks_maximize_covered_width (six_disks_cover_width.cu:186-234):
    the widest full-height strip r1..r6 cover as three stacked pairs, two
    three-disk columns or the r5/r6-centre placement (by bisection up to
    lambda), and whether the placement already works at lambda."""
    w = mpf(0)
    for (a, b), (c, d), (e, f) in PAIRINGS_3x2:
        w = max(w, sum(two_disks_maximize_height(D, r[x], r[y], mpf(1)) for x, y in ((a, b), (c, d), (e, f))))
    for first in ((0, 1, 2), (0, 1, 3), (0, 1, 4), (0, 1, 5), (0, 2, 3), (0, 2, 4), (0, 2, 5), (0, 3, 4),
                  (0, 3, 5), (0, 4, 5)):
        second = tuple(i for i in range(6) if i not in first)
        w = max(w, three_disks_height(D, mpf(1), [r[i] for i in first])
                + three_disks_height(D, mpf(1), [r[i] for i in second]))
    if _r56_center(D, w, r):
        # the reference bisects up to the double above lambda: a placement
        # that works at lambda itself ends the search at lambda
        if _r56_center(D, la, r):
            return la, True
        w, _ = _bisect(D, lambda D_, x: _r56_center(D_, x, r), w, la)
    return w, False


def rsb_c14_seven_disk(D, la, r):
    """C14, seven_disk_strategies (:1305-1405).  Strategies: 1 r1..r6 cover
    the width, 2 the r7 range is empty from the start, 3 after the T-shaped
    restriction, 4 after the four-disk strip, 5 every one of the 16 r7 pieces
    is covered (by r7 in the corner, a three-disk top strip, the 2-2-2
    placement, the r4-r3-r5 strip, the different-height cover or the
    five-disk height)."""
    R = max(rsb_R(la, r), mpf(0))
    Rub = rsb_R(la, r)
    wc, reaches = six_disks_covered_width(D, la, r)
    if reaches or D.cmp(wc, ">=", la):
        return True, 1
    wrem = max(la - wc, mpf(0))
    r7 = _r7_range(D, la, r, R, Rub, wrem)
    if _empty(D, r7):
        return True, 2
    _t_shaped(D, r7, la, r, R)
    if _empty(D, r7):
        return True, 3
    _four_disk_strip_r1_square(D, r7, la, r, R)
    if _empty(D, r7):
        return True, 4
    h34 = two_disks_maximize_height(D, r[2], r[3], la)
    heights = {}
    used = set()
    for i in range(16):
        a = r7[0] + (r7[1] - r7[0]) * i / 16
        b = r7[0] + (r7[1] - r7[0]) * (i + 1) / 16
        piece = [a, b]
        checks = (lambda: _r7_in_corner(D, R, piece, wrem),
                  lambda: _three_disk_top_strip(D, la, r, R, piece, heights),
                  lambda: _six_222(D, la, r, R, piece),
                  lambda: _r435_strip(D, la, r, R, piece),
                  lambda: _different_height(D, la, r, R, piece))
        for s, f in enumerate(checks, 1):
            if f():
                used.add(s)
                break
        else:
            if h34 > 0:
                h5 = _five_disks_23(D, la, r[0], r[1], r[4], r[5], a)
                if D.cmp(h34 + h5, ">=", 1):
                    used.add(6)
                    continue
            return False, 0
    return True, 5 + 10 * sum(1 << (s - 1) for s in used)


C14_PIECE_STRATEGIES = ("r7 in the corner", "three-disk top strip", "2-2-2 placement", "r4-r3-r5 strip",
                        "different heights", "five-disk height")


def c14_stages(la, r, pieces, stages=True, strategies=range(6)):
    """C14's stages on the point cell, each with the slack of its own
    decisions: ((width r1..r6 cover, reaches lambda, slack), [(r7 range after
    compute_r7_range / the T-shaped restriction / the four-disk strip, slack)]
    (left out unless `stages`), [[(verdict, slack) of each of the six piece
    strategies] per r7 piece] (those not in `strategies` with slack 0))."""
    la = mpf(la)
    r = [mpf(x) for x in r]
    R, Rub = max(rsb_R(la, r), mpf(0)), rsb_R(la, r)
    D = Dec()
    wc, reaches = six_disks_covered_width(D, la, r)
    width = (wc, reaches, D.slack)
    wrem = max(la - wc, mpf(0))
    ranges = []
    if reaches or wc >= la or not stages:  # C14 ends here; no r7 range is formed
        # (tighten_lb(0) leaves the remaining width's upper bound la - wc)
        return width, ranges, _c14_piece_rows(la, r, R, la - wc, pieces, D, strategies)
    r7 = _r7_range(D, la, r, R, Rub, wrem)
    ranges.append((list(r7), D.slack))
    if not (r7[0] > r7[1]):
        _t_shaped(D, r7, la, r, R)
    ranges.append((list(r7), D.slack))
    if not (r7[0] > r7[1]):
        _four_disk_strip_r1_square(D, r7, la, r, R)
    ranges.append((list(r7), D.slack))
    return width, ranges, _c14_piece_rows(la, r, R, wrem, pieces, D, strategies)


def _c14_piece_rows(la, r, R, wrem, pieces, D, strategies):
    out = []
    for a, b in pieces:
        piece = [mpf(a), mpf(b)]
        row = [(False, mpf(0))] * 6  # strategies not evaluated: no slack
        for k in strategies:
            Dk = Dec()
            if k == 0:
                v = _r7_in_corner(Dk, R, piece, wrem)
            elif k == 1:
                v = _three_disk_top_strip(Dk, la, r, R, piece, {})
            elif k == 2:
                v = _six_222(Dk, la, r, R, piece)
            elif k == 3:
                v = _r435_strip(Dk, la, r, R, piece)
            elif k == 4:
                v = _different_height(Dk, la, r, R, piece)
            else:
                h34 = two_disks_maximize_height(Dk, r[2], r[3], la)
                v = h34 > 0 and Dk.cmp(h34 + _five_disks_23(Dk, la, r[0], r[1], r[4], r[5], piece[0]), ">=", 1)
            row[k] = (bool(v), min(Dk.slack, D.slack if k == 0 else INF))
        out.append(row)
    return out
