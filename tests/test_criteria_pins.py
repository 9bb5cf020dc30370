"""Independent pins of the ten criteria of the 4-D box set.

tests/criteria_exact.py states each criterion of the reference in exact
real arithmetic (mpmath, 60 digits), written from the reference's formulas
and not from oracle/.  On point cells the oracle's outward-rounded FP64
verdict must equal the exact verdict wherever every comparison the exact
evaluation decides has a relative slack of at least 1e-11: a few hundred
rounded operations move a quantity by about 1e-14 at most, while a
transcription slip (a dropped term, a swapped operand, a sign or a wrong
constant) moves it by far more.  This pins both directions: the oracle
proves no cell the mathematics rejects, and misses none it accepts.

The point sets are aimed at each criterion's regime so that every
criterion, and every strategy inside it, both succeeds and fails; points
bisected onto each criterion's exact decision boundary add cells whose
slack is as small as 1e-8.  A mutation check plants plausible slips in the
exact statement and requires each to be caught.  (The GPU kernels are
compared with the oracle bit for bit by tests/test_c2_batch.py.)
"""
import json
import multiprocessing
import os

import mpmath
import numpy as np
import pytest

import criteria_exact as ce
import oracle

CLEAR = 1e-11
STRATEGIES = {0: {1, 2, 3}, 1: {1, 2, 3}, 2: {1, 2, 3}, 3: {1, 2}, 4: {1}, 5: {1}, 6: {1}, 7: {1}, 8: {1, 2},
              9: {1, 2}}


def _points(rng, n, la_rng, r1_rng, r_rng=None):
    la = rng.uniform(*la_rng, n)
    if r_rng is None:
        r = np.sort(rng.uniform(*r1_rng, (n, 3)), axis=1)[:, ::-1]
    else:
        r1 = rng.uniform(*r1_rng, n)
        r23 = np.sort(rng.uniform(*r_rng, (n, 2)), axis=1)[:, ::-1]
        r = np.column_stack([r1, np.minimum(r23, r1[:, None])])
    return np.column_stack([la, r])


def point_sets():
    rng = np.random.default_rng(0xC1C1EC)
    return np.concatenate([
        _points(rng, 250, (1.0, 2.5), (0.2, 0.8)),            # tight_rectangle radii
        _points(rng, 250, (1.0, 2.5), (0.0, 9.0 / 64.0)),     # size-bounded radii
        _points(rng, 250, (1.0, 8.0), (0.0, 0.3)),            # the wide grid of the branch-and-bound
        _points(rng, 150, (1.0, 3.0), (0.1, 0.5)),
        _points(rng, 150, (1.0, 1.4), (0.6, 0.8), (0.0, 0.15)),  # r1's pocket placement
    ])


def exact(points, fns=None):
    """(n, 10) verdicts, strategies and slacks of the exact statement."""
    v = np.zeros((len(points), 10), bool)
    st = np.zeros((len(points), 10), int)
    sl = np.zeros((len(points), 10))
    for i, p in enumerate(points):
        for k, (a, b, c) in enumerate(ce.c2_criteria(*p) if fns is None else _with(fns, p)):
            v[i, k], st[i, k], sl[i, k] = a, b, float(c)
    return v, st, sl


def _with(fns, p):
    la, r1, r2, r3 = (mpmath.mpf(x) for x in p)
    ub4 = min(r3, ce.tr_critical_ratio(la) * la - r1 - r2 - r3)
    args = ((la, r1, r2), (la, r1, r2, r3), (la, r1, r2, r3), (la, r1, r2, r3, ub4), (la, r1, r2, r3, ub4),
            (la, r1, r2), (la, r1, r2), (la, r1, r3), (la, r1, r2, r3), (la, r1, r2, r3))
    out = []
    for f, a in zip(fns, args):
        D = ce.Dec()
        v, s = f(D, *a)
        out.append((bool(v), s, D.slack))
    return out


def oracle_bits(points):
    p = np.asarray(points, dtype=np.float64)
    soa = np.ascontiguousarray(np.stack([p[:, 0], p[:, 0], p[:, 1], p[:, 1], p[:, 2], p[:, 2], p[:, 3], p[:, 3]]))
    mask, _ = oracle.c2_eval(soa)
    return (mask[:, None] >> np.arange(10)[None, :]) & 1 == 1


def boundary_points(points, v, sl, per_criterion=4):
    """Cells on both sides of each criterion's exact decision boundary:
    bisect the segment between a cell the criterion clearly proves and one it
    clearly rejects (convex combinations keep the radii sorted), then step
    off the boundary by 1e-4, 1e-6 and 1e-8 of the segment."""
    rng = np.random.default_rng(7)
    out = []
    for k in range(10):
        yes = np.flatnonzero(v[:, k] & (sl[:, k] >= CLEAR))
        no = np.flatnonzero(~v[:, k] & (sl[:, k] >= CLEAR))
        if len(yes) == 0 or len(no) == 0:
            continue
        for _ in range(per_criterion):
            a = [mpmath.mpf(x) for x in points[rng.choice(yes)]]
            b = [mpmath.mpf(x) for x in points[rng.choice(no)]]
            lo, hi = mpmath.mpf(0), mpmath.mpf(1)  # t = 0 proven, t = 1 rejected
            for _ in range(64):
                t = (lo + hi) / 2
                q = [x + t * (y - x) for x, y in zip(a, b)]
                if ce.c2_criteria(*q)[k][0]:
                    lo = t
                else:
                    hi = t
            for d in (1e-4, 1e-6, 1e-8):
                for t in (lo - d, hi + d):
                    out.append([float(x + t * (y - x)) for x, y in zip(a, b)])
    return np.array(out)


@pytest.fixture(scope="module")
def pinned():
    pts = point_sets()
    v, st, sl = exact(pts)
    bnd = boundary_points(pts, v, sl)
    bv, bst, bsl = exact(bnd)
    return (np.concatenate([pts, bnd]), np.concatenate([v, bv]), np.concatenate([st, bst]),
            np.concatenate([sl, bsl]), len(pts))


def test_oracle_verdicts_equal_exact_verdicts_at_clear_slack(pinned):
    pts, v, st, sl, _ = pinned
    o = oracle_bits(pts)
    clear = sl >= CLEAR
    bad = np.argwhere(clear & (o != v))
    assert len(bad) == 0, [(int(i), int(k), bool(o[i, k]), bool(v[i, k]), float(sl[i, k]), list(pts[i]))
                           for i, k in bad[:10]]
    # both directions are exercised, for every criterion and every strategy
    for k in range(10):
        assert (clear[:, k] & v[:, k]).sum() >= 15, f"criterion {k} proves too few cells"
        assert (clear[:, k] & ~v[:, k]).sum() >= 15, f"criterion {k} rejects too few cells"
        fired = set(st[clear[:, k] & v[:, k], k].tolist())
        assert STRATEGIES[k] <= fired, f"criterion {k}: strategies {STRATEGIES[k] - fired} never decide"


def test_boundary_cells_are_close_and_agree(pinned):
    pts, v, st, sl, n_plain = pinned
    o = oracle_bits(pts[n_plain:])
    bsl, bv = sl[n_plain:], v[n_plain:]
    for k in range(10):
        near = bsl[:, k] < 1e-5
        assert near.sum() >= 4, f"criterion {k}: no cells near its boundary"
        agree = near & (bsl[:, k] >= CLEAR)
        assert (o[agree, k] == bv[agree, k]).all()


def _mutants():
    """Plausible transcription slips, each as a replacement of one exact criterion."""
    m = {}

    def r1_no_split_cost(D, la, r1, ub_r2):  # strategy 2 without the splitting cost ub_r2
        rem = ce.tr_critical_ratio(la) * la - r1
        side = mpmath.sqrt(2 * r1)
        if D.cmp(4 * r1, ">", 1) and ce.tr_can_recurse(D, la - mpmath.sqrt(4 * r1 - 1), mpmath.mpf(1), ub_r2, rem):
            return True, 1
        if D.cmp(side, "<", 1):
            wright = la - side
            wgt = wright * ce.CR if ce.size_bound_ok(D, ub_r2, min(wright, mpmath.mpf(1))) else \
                ce.tr_critical_ratio_wh(wright, mpmath.mpf(1)) * wright
            top = rem - wgt
            if D.cmp(top, ">", 0) and ce.tr_can_recurse(D, side, 1 - side, ub_r2, top):
                return True, 2
        return False, 0
    m["dropped splitting cost (criterion 0)"] = (0, r1_no_split_cost)

    def simple_fixed_q10(D, la, r1, r2, ub_r3):  # lambda - width instead of the reference's lambda + width
        w = ce.two_disks_maximize_height(D, r1, r2, mpmath.mpf(1))
        if w <= 0:
            return False, 0
        rem = ce.tr_critical_ratio(la) * la - r1 - r2
        return (True, 3) if ce.tr_can_recurse(D, la - w, mpmath.mpf(1), ub_r3, rem) else (False, 0)
    m["sign of the remaining width (criterion 1, quirk Q10)"] = (1, simple_fixed_q10)

    def strip_sign(D, la, r1, r2, r3, ub_4):  # lambda + w for the width left of the strip
        w = ce.cover_strip_width(D, la, r1, r2, r3)
        if w is None:
            return False, 0
        rem = ce.tr_critical_ratio(la) * la - r1 - r2 - r3
        return (True, 1) if ce.tr_can_recurse(D, la + w, mpmath.mpf(1), ub_4, rem) else (False, 0)
    m["sign of the remaining width (criterion 4)"] = (4, strip_sign)

    def even_mul(D, la, r1, r2):  # r2 * 0.61 instead of r2 / 0.61
        big = la / 2
        if D.cmp(big, "<", 1) and not ce.size_bound_ok(D, r1, big):
            return False, 0
        small = big - r2 * ce.CR
        return (True, 1) if (D.cmp(small, ">=", 1) or ce.size_bound_ok(D, r2, small)) else (False, 0)
    m["operation slip r2 / 0.61 -> r2 * 0.61 (criterion 5)"] = (5, even_mul)

    def uneven_swapped(D, la, r1, r2):  # the size bound checked on r1 instead of r2
        if r1 <= 0:
            return False, 0
        need = mpmath.sqrt(r1) / ce.UB_DISK_RADIUS
        if D.cmp(need, ">=", la) or D.cmp(la * ce.CR, "<=", r2 + need / ce.CR):
            return False, 0
        wrem = (la * ce.CR - (r2 + need / ce.CR)) / ce.CR
        return (True, 1) if (D.cmp(wrem, ">=", 1) or ce.size_bound_ok(D, r1, wrem)) else (False, 0)
    m["swapped operand r2 -> r1 (criterion 6)"] = (6, uneven_swapped)

    def vstrip_area(D, la, r1, r2, r3):  # the weight requirement without the 0.61
        w = ce.two_disks_maximize_height(D, r1, r2, mpmath.mpf(1))
        if w <= 0:
            return False, 0
        rem = la * ce.CR - (r1 + r2)
        wrem = la - w
        if D.cmp(wrem, ">=", 1) or ce.size_bound_ok(D, r3, wrem):
            return (True, 1) if D.cmp(rem, ">=", wrem) else (False, 0)
        return (True, 2) if ce.rsb_can_recurse(D, rem, r3, wrem, mpmath.mpf(1)) else (False, 0)
    m["dropped factor 0.61 (criterion 9)"] = (9, vstrip_area)
    return m


def test_planted_slips_are_caught(pinned):
    """Each plausible slip, planted in one exact criterion, disagrees with
    the oracle on some cell at clear slack: the comparison pins the
    criteria's formulas, not just their rough behaviour."""
    pts, v, st, sl, n_plain = pinned
    o = oracle_bits(pts)
    base = [ce.r1_strategies, ce.r1_r2_simple, ce.r1_r2_stacked, ce.r123_horizontal, ce.r123_cover_strip,
            ce.s1_even_split, ce.s23_uneven_split, ce.s23_uneven_split, ce.s4_opposite_corners,
            ce.s5_vertical_strip]
    sample = np.arange(0, n_plain, 3)
    for name, (k, f) in _mutants().items():
        fns = list(base)
        fns[k] = f
        if k == 6:
            fns[7] = f
        mv, _, msl = exact(pts[sample], fns)
        caught = (msl[:, k] >= CLEAR) & (mv[:, k] != o[sample, k])
        assert caught.any(), f"slip not caught: {name}"


# --------------------------------------------- rectangle_size_bound leaf steps

RSB_STRATEGIES = {6: {1, 2}, 7: {1, 2}, 8: {1, 2, 3, 4}, 9: {1}, 10: {1, 2}, 11: {1, 2, 3}, 12: {1, 3},
                  13: {1, 2, 3, 4}, 14: {1, 3, 4}, 15: {1, 2}, 16: {1}, 17: {1, 4}, 18: {1, 3}}

# Cells whose C7 (r1_r2_vertical_strip) is decided by its last recursion, with
# all six disks placed (found by a seeded random search; radii beyond 9/64,
# which the criterion accepts as it accepts any input).
C7_LAST = [[float.fromhex(x) for x in c] for c in (
    ['0x1.847f8e46becebp+0', '0x1.8e6e42ad8d75bp-3', '0x1.74bb84dbb90a1p-3', '0x1.7098ab7bf229dp-3',
     '0x1.4f9999d5000f7p-3', '0x1.d35efbf9248c4p-4', '0x1.cf8da5ce5719fp-5'],
    ['0x1.36ccbb221e819p+0', '0x1.291dc48deb66bp-3', '0x1.1c8222dd928f3p-3', '0x1.06239e07b7794p-3',
     '0x1.e27b30807f9cdp-4', '0x1.28b31f5fea714p-4', '0x1.a07a3747ec74ep-5'],
    ['0x1.ba526feec5591p+0', '0x1.1829b90094e46p-2', '0x1.e98a431737f4ep-3', '0x1.bdb4a6da5c134p-3',
     '0x1.ee3f08a9da42fp-4', '0x1.97a77f7637698p-4', '0x1.541bffa5d600fp-4'],
    ['0x1.5a1c24f8e3243p+0', '0x1.597f2759440cap-3', '0x1.52b3d0f58d97dp-3', '0x1.294f0e596b163p-3',
     '0x1.15f81c194d31ap-3', '0x1.a850591c23438p-4', '0x1.3ec58a11fc791p-5'])]


def rsb_point_cells():
    rng = np.random.default_rng(0xC1C1EC + 1)

    def pts(n, la_rng, rmax):
        la = rng.uniform(*la_rng, n)
        r = np.sort(rng.uniform(0, rmax, (n, 6)), axis=1)[:, ::-1]
        keep = la * 0.61 - r.sum(axis=1) > 1e-3  # weight left over (a non-vacuous cell)
        return np.column_stack([la, r])[keep]

    return np.concatenate([pts(500, (1.0, 2.5), 9.0 / 64.0), pts(500, (1.0, 2.5), 0.06), pts(500, (1.0, 8.0), 0.3),
                           pts(300, (1.0, 2.5), 0.03), np.array(C7_LAST)])


def test_rsb_leaf_steps_equal_exact_at_clear_slack():
    """The leaf-chain steps C1..C13 (C14 excepted), each evaluated alone on
    point cells by the oracle's per-step masks, against their exact
    statement: C1 even split, C2 uneven split, C3 multi-disk strips, C4
    vertical wall building, C5 r1 in a corner with explicit recursion, C6 r1
    and r2 in opposite corners, C7 r1, r2 strip and greedy strips, C8 2x2
    cover with strip and recursion, C9 L-shaped recursion, C10 r1 in a
    corner with wall building, C11 advanced multi-disk strip, C12 five-disk
    and C13 six-disk full covers; a planted slip in each of C2..C13 but C7
    is caught."""
    P = rsb_point_cells()
    cells = np.zeros((len(P), 14))
    cells[:, 0] = cells[:, 1] = P[:, 0]
    for k in range(6):
        cells[:, 2 + 2 * k] = cells[:, 3 + 2 * k] = P[:, 1 + k]
    mask = oracle.rsb_masks(cells)
    fired = {k: set() for k in ce.RSB_STEPS}
    counts = {k: [0, 0] for k in ce.RSB_STEPS}
    for i, p in enumerate(P):
        for k, (v, s, sl) in ce.rsb_steps(p[0], p[1:]).items():
            if sl < CLEAR:
                continue
            o = bool(mask[i] >> k & 1)
            assert o == v, (k, list(p), o, v, float(sl))
            counts[k][int(v)] += 1
            if v:
                fired[k].add(s)
    for k in ce.RSB_STEPS:
        assert min(counts[k]) >= 15, f"step {k}: proves {counts[k][1]}, rejects {counts[k][0]}"
        assert RSB_STRATEGIES[k] <= fired[k], f"step {k}: strategies {RSB_STRATEGIES[k] - fired[k]} never decide"

    # planted slips: C2 with the largest remaining disk one index early,
    # C4 with its remaining width lambda - w instead of the reference's lambda + w (Q5)
    def c2_index(D, la, r):
        if r[0] <= 0:
            return False, 0
        need = ce.bound_required_height(r[0])
        if D.cmp(need, ">=", la):
            return False, 0
        wneed, w, at = need * ce.CR, mpmath.mpf(0), -1
        for i in range(6):
            w += r[i]
            if D.cmp(w, ">=", wneed):
                at = i
                break
        largest = r[5] if at < 0 or at == 5 else r[at]
        d1 = wneed + r[5] if at < 0 else w
        if D.cmp(d1, ">=", la * ce.CR):
            return False, 0
        wrem = la - d1 / ce.CR
        return (True, 1) if (D.cmp(wrem, ">=", 1) or ce.size_bound_ok(D, largest, wrem)) else (False, 0)

    def _c4_sub(D, la, r):
        R = max(ce.rsb_R(la, r), mpmath.mpf(0))
        rem_w = la - mpmath.sqrt(2 * r[5])
        if r[5] <= 0 or D.cmp(R, "<=", 0):
            return False, 0
        if D.cmp(rem_w, "<", 1) and not ce.size_bound_ok(D, r[0], rem_w):
            return False, 0
        if D.cmp(r[5], ">", ce.WEIGHT_BOUND_WALL_BUILDING):
            return False, 0
        below, used = r[5] * ce.PHASE_FACTOR_WALL_BUILDING, ce.CR * mpmath.sqrt(2 * r[5])
        for _ in range(32):
            total = r[5] + R - used
            if D.cmp(r[5] + R, "<=", used):
                break
            req = ce.CR * ce.bound_required_height(below)
            if D.cmp(req, "<=", total):
                arw = la - (req + below) / ce.CR
                if D.cmp(arw, ">=", 1) or ce.size_bound_ok(D, r[0], arw):
                    return True, 1
            used += ce.CR * mpmath.sqrt(2 * below)
            below *= ce.PHASE_FACTOR_WALL_BUILDING
        return False, 0

    def c3_single_root(D, la, r):  # C3's two-disk rows with one square root (Q6 "fixed")
        saved = ce.ROW_HEIGHT
        ce.ROW_HEIGHT = mpmath.sqrt
        try:
            return ce.rsb_c3_multi_disk_strip(D, la, r)
        finally:
            ce.ROW_HEIGHT = saved

    def c5_index(D, la, r):  # C5's largest remaining disk one index early
        saved = ce._c5_max_remaining
        ce._c5_max_remaining = lambda la_, r_, nd: min(r_[5], ce.rsb_R(la_, r_)) if nd == 5 else r_[nd]
        try:
            return ce.rsb_c5_r1_in_corner(D, la, r)
        finally:
            ce._c5_max_remaining = saved

    def c8_index(D, la, r):  # C8's recursion from r5 on with r6 as its largest disk
        saved = ce.rsb_can_recurse
        ce.rsb_can_recurse = lambda D_, rem, largest, w, h: saved(D_, rem, r[5] if largest is r[4] else largest, w, h)
        try:
            return ce.rsb_c8_two_by_two(D, la, r)
        finally:
            ce.rsb_can_recurse = saved

    def c9_r5(D, la, r):  # C9's two-horizontal recursion with r5 as the largest disk (the quirk "fixed")
        saved = ce.rsb_can_recurse
        ce.rsb_can_recurse = lambda D_, rem, largest, w, h: saved(D_, rem, r[4] if largest is r[5] else largest, w, h)
        try:
            return ce.rsb_c9_l_shaped(D, la, r)
        finally:
            ce.rsb_can_recurse = saved

    def c10_factor(D, la, r):  # C10 with the wall-building weight bound's factor 1/2 as 1/4
        saved = ce.WEIGHT_BOUND_WALL_BUILDING
        ce.WEIGHT_BOUND_WALL_BUILDING = saved / 2
        try:
            return ce.rsb_c10_r1_corner_wall_building(D, la, r)
        finally:
            ce.WEIGHT_BOUND_WALL_BUILDING = saved

    def c6_index(D, la, r):  # C6's regions A and B filled from one disk too far on
        saved = ce._c6_size_bounds
        ce._c6_size_bounds = lambda D_, la_, r_, wa, ha, wb, hb, start: saved(D_, la_, r_, wa, ha, wb, hb,
                                                                              min(start + 1, 6))
        try:
            return ce.rsb_c6_opposite_corners(D, la, r)
        finally:
            ce._c6_size_bounds = saved

    def c11_all_six(D, la, r):  # C11 with the set of all six disks allowed (r6 taken as the rest's largest)
        v, st = ce.rsb_c11_advanced_strip(D, la, r)
        if v:
            return v, st
        w = sum(r) / ce.CR
        if any(D.cmp(4 * x, "<=", w * w) for x in r):
            return False, 0
        if D.cmp(sum(mpmath.sqrt(4 * x - w * w) for x in r), "<", 1):
            return False, 0
        return (True, 1) if ce._rem_width_ok(D, la, w, r[5]) else (False, 0)

    def c12_weight(D, la, r):  # C12's border covers with r5 left in the remaining weight
        saved = ce.rsb_can_recurse

        def slipped(D_, rem, largest, w, h):
            return saved(D_, rem + r[4] if largest is r[5] else rem, largest, w, h)
        ce.rsb_can_recurse = slipped
        try:
            return ce.rsb_c12_five_disk(D, la, r)
        finally:
            ce.rsb_can_recurse = saved

    def c13_index(D, la, r):  # C13's recursions with r5 as the largest disk
        saved = ce.rsb_can_recurse
        ce.rsb_can_recurse = lambda D_, rem, largest, w, h: saved(D_, rem, r[4] if largest is r[5] else largest, w, h)
        try:
            return ce.rsb_c13_six_disk(D, la, r)
        finally:
            ce.rsb_can_recurse = saved

    for k, f in ((7, c2_index), (9, _c4_sub), (8, c3_single_root), (10, c5_index), (11, c6_index), (13, c8_index),
                 (14, c9_r5), (15, c10_factor), (16, c11_all_six), (17, c12_weight), (18, c13_index)):
        caught = False
        for i, p in enumerate(P):
            D = ce.Dec()
            v, _ = f(D, mpmath.mpf(p[0]), [mpmath.mpf(x) for x in p[1:]])
            if D.slack >= CLEAR and v != bool(mask[i] >> k & 1):
                caught = True
                break
        assert caught, f"planted slip in step {k} not caught"


# --------------------------------------------- tight_rectangle leaf criteria

def tr_leaf_points():
    """Point leaves (lambda, r1 >= r2 >= r3 >= r4) with weight left over:
    the radii share a fraction of critical_ratio(lambda) * lambda, in
    proportions from seeded Dirichlet draws aimed at the two-by-two, three-
    disk and gap placements, plus leaves around the three-disk worst case
    r*(lambda) that the manually proved pocket covers."""
    rng = np.random.default_rng(0xC1C1EC + 2)

    def weight(la):
        return np.array([float(ce.tr_critical_ratio(mpmath.mpf(x))) for x in la]) * la

    def leaves(n, la_rng, frac, conc):
        la = rng.uniform(*la_rng, n)
        r = np.sort(rng.dirichlet(conc, n), axis=1)[:, ::-1] * (weight(la) * rng.uniform(*frac, n))[:, None]
        return np.column_stack([la, r])

    la = rng.uniform(1.0, 1.06, 150)
    rs = la * la / 16 + 0.15625 + 0.03515625 / (la * la)
    d = rng.uniform(0, 1, (150, 3)) * np.array([0.0012, 0.0024, 0.0036])
    r4 = rng.uniform(0, 1, 150) * (weight(la) - 3 * rs + d.sum(axis=1))
    pocket = np.column_stack([la, np.sort(np.column_stack([rs[:, None] - d, r4]), axis=1)[:, ::-1]])
    P = np.concatenate([leaves(150, (1.0, 2.1), (0.6, 0.999), (1, 1, 1, 1)),
                        leaves(150, (1.0, 2.1), (0.8, 0.999), (3, 2, 1, 1)),
                        leaves(150, (1.0, 1.5), (0.85, 0.999), (4, 4, 4, 1)),
                        leaves(150, (1.0, 1.6), (0.85, 0.999), (6, 6, 2, 2)),
                        leaves(100, (1.0, 1.3), (0.9, 0.999), (8, 8, 8, 0.5)), pocket])
    return P[weight(P[:, 0]) - P[:, 1:].sum(axis=1) > 1e-6]


def _tr_leaf_fns(la, r, R4, R, over=None):
    f = {8: lambda D: ce.tbt_lambda_small(D, la, r, R),
         9: lambda D: ce.tbt_lambda_big(D, la, r, R),
         10: lambda D: ce.three_disk_big_strip(D, la, r[0], r[1], r[2], r[3], R4),
         11: lambda D: ce.three_disk_small_corner(D, la, r[0], r[1], r[2], r[3], R4),
         12: lambda D: ce.three_disk_r1_r3_pocket(D, la, r[0], r[1], r[2], r[3], R4),
         13: lambda D: ce.manual_three_disk_pocket(D, la, r[0], r[1], r[2]),
         2: lambda D: ce.tr_gaps(D, la, r, R)[0]}
    if over:
        for k, g in over.items():
            f[k] = (lambda g_: lambda D: g_(D, la, r, R4, R))(g)
    return f


def tr_leaf_exact(P, over=None, only=None):
    """{bit: (verdicts, slacks)} of the exact leaf strategies on the points."""
    out = {}
    for i, p in enumerate(P):
        la, r = mpmath.mpf(p[0]), [mpmath.mpf(x) for x in p[1:]]
        R4, R = ce.tr_leaf_weights(la, r)
        for k, f in _tr_leaf_fns(la, r, R4, R, over).items():
            if only is not None and k not in only:
                continue
            D = ce.Dec()
            v = bool(f(D))
            vs, ss = out.setdefault(k, (np.zeros(len(P), bool), np.zeros(len(P))))
            vs[i], ss[i] = v, float(D.slack)
    return out


def tr_leaf_oracle(P):
    cells = np.zeros((len(P), 10))
    cells[:, 0] = cells[:, 1] = P[:, 0]
    for k in range(4):
        cells[:, 2 + 2 * k] = cells[:, 3 + 2 * k] = P[:, 1 + k]
    return oracle.tr_leaf_masks(cells)


def test_tr_leaf_criteria_equal_exact_at_clear_slack():
    """The tight_rectangle leaf criteria (tight_rectangle.cu:247-249), each
    strategy on its own: two-by-two with lambda small (r1's square, r2 and
    r4 beside it, r3 in the far corner; r1 large: r2..r4 stacked beside r1's
    full-height strip) and lambda big (two stacked pairs), the three-disk big
    strip, small corner, r1-r3 pocket and the manually proved pocket, and the
    gaps strategy (with its bisection); the oracle's bits against the exact
    statement, both ways, where the slack is clear.  The leaves' criteria
    and strategies each prove and reject at least 15 leaves, and a planted
    slip in each is caught."""
    P = tr_leaf_points()
    m = tr_leaf_oracle(P)
    assert not (m >> 31).any()
    ex = tr_leaf_exact(P)
    for k, (v, sl) in ex.items():
        clear = sl >= CLEAR
        o = ((m >> k) & 1).astype(bool)
        bad = np.flatnonzero(clear & (o != v))
        assert bad.size == 0, (k, [list(P[i]) for i in bad[:3]])
        assert min((clear & v).sum(), (clear & ~v).sum()) >= 15, f"bit {k}: {(clear & v).sum()} proven, " \
                                                                  f"{(clear & ~v).sum()} rejected"
    # the criteria are the strategies' disjunctions
    assert (((m >> 8) | (m >> 9)) & 1 == m & 1).all()
    assert (((m >> 10) | (m >> 11) | (m >> 12) | (m >> 13)) & 1 == (m >> 1) & 1).all()

    def small_r3_r4_swapped(D, la, r, R4, R):  # r3 in the strip above r1's square, r4 in the corner
        return ce.tbt_lambda_small(D, la, [r[0], r[1], r[3], r[2]], R)

    def big_one_pairing(D, la, r, R4, R):  # only the pairing (r1, r2) + (r3, r4)
        def th(a, b):
            return ce.two_disks_maximize_height(D, a, b, mpmath.mpf(1))
        w = th(r[0], r[1]) + th(r[2], r[3])
        return w != 0 and ce.tr_can_recurse(D, la - w, mpmath.mpf(1), min(R, r[3]), R)

    def strip_sign(D, la, r, R4, R):  # + 4 (r2 - r3)^2 in the strip width
        s_sq = 2 * r[1] + 2 * r[2] - mpmath.mpf(1) / 4 + 4 * (r[1] - r[2]) ** 2
        if D.cmp(s_sq, "<=", 0):
            return False
        la_rem = la - mpmath.sqrt(s_sq)
        h1 = 4 * r[0] - la_rem * la_rem
        if D.cmp(h1, "<=", 0):
            return False
        return D.cmp(h1, ">=", 1) or ce.tr_can_recurse(D, la_rem, 1 - mpmath.sqrt(h1), r[3], R4)

    def corner_weight(D, la, r, R4, R):  # the recursion given R (r4 already spent) instead of R4
        return ce.three_disk_small_corner(D, la, r[0], r[1], r[2], r[3], R)

    def pocket_no_half(D, la, r, R4, R):  # la - S1 - S3 for the pocket's height term (S1 not halved)
        p = ce._pocket_X(D, la, r[0], r[1], r[2])
        if p is None:
            return False
        s1, s3, X = p
        if D.cmp(X, "<=", 0):
            return True
        y_rest = X ** 2
        if D.cmp(r[0], "<=", y_rest):
            return False
        Y = mpmath.mpf(1) / 2 - mpmath.sqrt(r[0] - y_rest)
        return D.cmp(Y, "<=", 0) or ce.tr_can_recurse(D, X, Y, r[3], R4)

    def manual_eps(D, la, r, R4, R):  # r3's tolerance 0.002 instead of 0.003
        lasq = la * la
        rstar = lasq / 16 + mpmath.mpf(0.15625) + mpmath.mpf(0.03515625) / lasq
        d = rstar - r[2]
        if D.cmp(d, "<", -0.002) or D.cmp(d, ">", 0.002):
            return False
        return ce.manual_three_disk_pocket(D, la, r[0], r[1], r[2])

    def gaps_centre(D, la, r, R4, R):  # r4's centre at the gap's far end instead of its middle
        saved = ce._gap_ok

        def slipped(D_, gap, r1, w1, r2, w2, r4):
            y4 = r4 - gap * gap / 4
            if D_.cmp(y4, "<=", 0):
                return False
            ok, p, _ = ce.circle_intersection(ce.Dec(), (w1 / 2, mpmath.mpf(1) / 2, r1),
                                              (w2 / 2 + gap + w1, mpmath.mpf(1) / 2, r2))
            return ok and D_.cmp(ce.dist2((gap + w1, mpmath.sqrt(y4)), p), "<=", r4)
        ce._gap_ok = slipped
        try:
            return ce.tr_gaps(D, la, r, R)[0]
        finally:
            ce._gap_ok = saved

    slips = {8: small_r3_r4_swapped, 9: big_one_pairing, 10: strip_sign, 11: corner_weight, 12: pocket_no_half,
             13: manual_eps, 2: gaps_centre}
    sample = np.arange(0, len(P), 2)
    for k, f in slips.items():
        v, sl = tr_leaf_exact(P[sample], {k: f}, only={k})[k]
        o = ((m[sample] >> k) & 1).astype(bool)
        assert ((sl >= CLEAR) & (v != o)).any(), f"slip in bit {k} not caught"


# ------------------------------------------- rectangle_size_bound leaf step C14

C14_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c14_cells.json")
C14_RANGE_TOL = 1e-7  # relative; the range bounds are bisection results of the reference's recursions


def c14_fixture():
    with open(C14_FIXTURE) as f:
        d = json.load(f)
    return tuple(np.array([[float.fromhex(x) for x in row] for row in d[k]])
                 for k in ("bnb", "different_heights", "two_two_two"))


def _cells14(P):
    cells = np.zeros((len(P), 14))
    cells[:, 0] = cells[:, 1] = P[:, 0]
    for k in range(6):
        cells[:, 2 + 2 * k] = cells[:, 3 + 2 * k] = P[:, 1 + k]
    return cells


def _c14_verdict(p):
    D = ce.Dec()
    v, s = ce.rsb_c14_seven_disk(D, mpmath.mpf(p[0]), [mpmath.mpf(x) for x in p[1:]])
    return bool(v), s, float(D.slack)


def _c14_stages(args):
    """Exact stages of one cell; None where a planted slip leaves the reals
    (a square root of a negative number): that cell is then left out."""
    p, pieces, stages, strategies = args if len(args) == 4 else (*args, True, range(6))
    try:
        (w, reaches, wsl), ranges, rows = ce.c14_stages(p[0], p[1:], pieces, stages, strategies)
    except (TypeError, ValueError, ZeroDivisionError):
        return None
    return ((float(w), reaches, float(wsl)), [([float(x) for x in rg], float(sl)) for rg, sl in ranges],
            [[(v, float(sl)) for v, sl in row] for row in rows])


def _pool_map(fn, items):
    """fn over items on forked workers (which see the exact statement as the
    parent has it, planted slips included)."""
    with multiprocessing.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
        return pool.map(fn, items, chunksize=2)


def c14_pieces(P, n, seed):
    """n seeded r7 pieces [a, b] in [0, min(r6, R)] per cell."""
    rng = np.random.default_rng(seed)
    top = np.minimum(P[:, 6], P[:, 0] * 0.61 - P[:, 1:].sum(axis=1))
    a = rng.uniform(0, 1, (len(P), n)) * top[:, None]
    b = a + rng.uniform(0, 1, (len(P), n)) * (top[:, None] - a)
    return np.stack([a, b], axis=2)


def c14_stage_inputs():
    bnb, dh, t222 = c14_fixture()
    P = np.concatenate([bnb[:96], dh[:, :7], t222[:, :7]])
    pieces = np.concatenate([c14_pieces(bnb[:96], 4, 0xC14), np.repeat(dh[:, None, 7:9], 4, axis=1),
                             np.repeat(t222[:, None, 7:9], 4, axis=1)])
    return P, pieces


def c14_stage_mismatches(P, pieces, res, out, bits):
    """(width/range mismatches, piece mismatches, per-strategy [rejected,
    covered] counts at clear slack) of exact stages against the oracle's."""
    stage_bad, piece_bad = [], []
    cnt = np.zeros((6, 2), int)
    for i, r in enumerate(res):
        if r is None:
            continue
        (w, reaches, wsl), ranges, rows = r
        if wsl >= CLEAR and not reaches and abs(w - out[i, 0]) > 1e-9 * w:
            stage_bad.append((i, "width", w, out[i, 0]))
        for s, (rg, sl) in enumerate(ranges):
            if sl < CLEAR:
                continue
            lo, hi = out[i, 1 + 2 * s], out[i, 2 + 2 * s]
            if (rg[0] > rg[1]) != (lo > hi):
                stage_bad.append((i, s, rg, lo, hi))
            elif rg[0] <= rg[1] and (abs(rg[0] - lo) > C14_RANGE_TOL * rg[1] or abs(rg[1] - hi) > 1e-12 * rg[1]):
                stage_bad.append((i, s, rg, lo, hi))
        for j, row in enumerate(rows):
            for k, (v, sl) in enumerate(row):
                if sl < CLEAR:
                    continue
                if bool(bits[i, j] >> k & 1) != v:
                    piece_bad.append((i, j, k))
                cnt[k, int(v)] += 1
    return stage_bad, piece_bad, cnt


def test_c14_verdicts_equal_exact_at_clear_slack():
    """C14 (seven_disk_strategies.cu:1305-1405) on the centres of cells that
    reach it in the branch-and-bound: the oracle's per-step bit 19 against
    the exact statement, both ways, at clear slack; it proves and rejects at
    least 15 cells, and each way it proves (r1..r6 spanning lambda, an r7
    range empty from the start, after the T-shaped restriction, after the
    four-disk strip, every r7 piece covered) decides some cell."""
    bnb = c14_fixture()[0]
    mask = oracle.rsb_masks(_cells14(bnb))
    res = _pool_map(_c14_verdict, list(bnb))
    seen, cnt = set(), [0, 0]
    for i, (v, s, sl) in enumerate(res):
        if sl < CLEAR:
            continue
        assert bool(mask[i] >> 19 & 1) == v, (list(bnb[i]), v, s, sl)
        cnt[int(v)] += 1
        if v:
            seen.add(min(s, 5))
    assert min(cnt) >= 15, cnt
    assert seen == {1, 2, 3, 4, 5}, seen


def test_c14_stages_equal_exact():
    """C14's stages, each on its own: the width r1..r6 cover, the r7 range
    after compute_r7_range, the T-shaped restriction and the four-disk strip
    (emptiness, and bounds within 1e-7 of the range), and each of the six
    piece strategies on seeded r7 pieces (r7 in the corner, three-disk top
    strip, 2-2-2, r4-r3-r5 strip, different heights, five-disk height), the
    oracle's against the exact statement at clear slack; every piece strategy
    covers some pieces and fails on others."""
    P, pieces = c14_stage_inputs()
    out, bits = oracle.rsb_c14_probe(_cells14(P), pieces)
    res = _pool_map(_c14_stages, list(zip(P, pieces)))
    assert all(r is not None for r in res)
    stage_bad, piece_bad, cnt = c14_stage_mismatches(P, pieces, res, out, bits)
    assert not stage_bad, stage_bad[:3]
    assert not piece_bad, piece_bad[:5]
    assert (cnt[:, 1] >= 3).all() and (cnt[:, 0] >= 15).all(), cnt.tolist()


def _swap(r, i, j):
    r = list(r)
    r[i], r[j] = r[j], r[i]
    return r


C14_SLIPS = {
    # name: (module attribute, replacement factory, "stages" or the piece strategy compared)
    "r7 in the corner: remaining height 1 + sqrt(h)": (
        "_r7_in_corner", lambda f: (lambda D, R, r7, w: (lambda h: False if D.cmp(h, "<", 0) else ce.rsb_can_recurse(
            D, ce._rest(R, r7[1]), r7[1], w, 1 + mpmath.sqrt(h)))(4 * r7[0] - w * w)), 0),
    "top strip: r7's lower bound taken for its upper bound": (
        "_three_disk_top_strip", lambda f: (lambda D, la, r, R, r7, heights: f(D, la, r, R, [r7[0], r7[0]],
                                                                               heights)), 1),
    "2-2-2: pairs r1/r3 and r2/r4 for r1/r2 and r3/r4": ("_six_222", lambda f: (lambda D, la, r, R, r7: f(
        D, la, _swap(r, 1, 2), R, r7)), 2),
    "r4-r3-r5 strip: r3 and r4 swapped": ("_r435_strip", lambda f: (lambda D, la, r, R, r7: f(D, la, _swap(r, 2, 3),
                                                                                             R, r7)), 3),
    "different heights: r1 and r2 swapped": ("_different_height", lambda f: (lambda D, la, r, R, r7: f(
        D, la, _swap(r, 0, 1), R, r7)), 4),
    "five-disk height: m1 and m2 swapped": ("_five_disks_23", lambda f: (lambda D, la, t1, t2, m1, m2, m3: f(
        D, la, t1, t2, m2, m1, m3)), 5),
    "r7 range: efficiency over lambda instead of the remaining width": (
        "_r7_range", lambda f: (lambda D, la, r, R, Rub, wrem: f(D, la, r, R, Rub, la)), "stages"),
    "T-shaped restriction: r1 and r2 swapped": ("_t_shaped", lambda f: (lambda D, r7, la, r, R: f(
        D, r7, la, _swap(r, 0, 1), R)), "stages"),
    "four-disk strip: r1 and r2 swapped": ("_four_disk_strip_r1_square", lambda f: (lambda D, r7, la, r, R: f(
        D, r7, la, _swap(r, 0, 1), R)), "stages"),
    "six-disk width: r5 and r6 swapped": ("six_disks_covered_width", lambda f: (lambda D, la, r: f(
        D, la, _swap(r, 4, 5))), "stages"),
}


def test_c14_planted_slips_are_caught():
    """A plausible slip planted in each of C14's parts disagrees with the
    oracle at clear slack (widths and ranges, or piece strategies)."""
    P, pieces = c14_stage_inputs()
    out, bits = oracle.rsb_c14_probe(_cells14(P), pieces)
    for name, (attr, make, what) in C14_SLIPS.items():
        saved = getattr(ce, attr)
        setattr(ce, attr, make(saved))
        try:
            stages = what == "stages"
            res = _pool_map(_c14_stages, [(p, pc, stages, () if stages else (what,)) for p, pc in zip(P, pieces)])
        finally:
            setattr(ce, attr, saved)
        stage_bad, piece_bad, _ = c14_stage_mismatches(P, pieces, res, out, bits)
        assert (stage_bad if stages else piece_bad), f"slip not caught: {name}"
