"""The tight_rectangle report of a final critical set
(Prover/tight_rectangle/tight_rectangle.cu:498-536) through
cc_tr_report_criticals.  Its Max_distances lines come from a one-workgroup
reduction on the device; the reference (:343-456) and the CPU oracle walk
the cells in order on one thread, each running extreme keeping the first
cell that attains it, and S1/h2/S3 restarting from a cell's own values
while the pocket width's running maximum is negative.  The cases below put
equal extremes in cells far apart (other lanes of the workgroup), restarts
at several places, and NaN terms (two-disk cells with r1 < 1/4, whose
scalar sqrt the reference takes without a trap), and compare the text byte
for byte with the oracle's.  The oracle's loop itself is pinned on the CPU
against the same quantities in exact arithmetic, with planted slips, and
against its restart rule."""
import numpy as np
import pytest

from circlecover_amd.native import CC_E_TRAP, CCError
from report_format import max_distances, tr_report

W = 1e-7  # interval width of the synthetic cells


def _cell(la, r):
    c = np.empty(10)
    c[0], c[1] = la, la + W
    for k in range(4):
        c[2 + 2 * k], c[3 + 2 * k] = r[k], r[k] + W
    return c


def two_disk_cells(rng, n):
    """Near the two-disk worst case: r1 the circumcircle of the lambda x 1
    rectangle, r2^2 = 1/4, r3 and r4 below 0.1."""
    out = []
    for _ in range(n):
        la = rng.uniform(1.0, 1.3)
        r1 = 0.25 * (la * la + 1.0) + rng.uniform(-2e-3, 2e-3)
        r2 = 0.25 + rng.uniform(-2e-3, 2e-3)
        out.append(_cell(la, (r1, r2, rng.uniform(0.0, 0.09), rng.uniform(0.0, 0.05))))
    return np.array(out).reshape(-1, 10)


def three_disk_cells(rng, n, pw=None):
    """Three-disk cells built from S1, h2 and a target pocket width
    la - S1 - S3 (drawn from [-0.3, 0.1] unless given)."""
    out = []
    for k in range(n):
        la = rng.uniform(1.0, 1.2)
        s1 = rng.uniform(0.2, 0.5)
        h2 = rng.uniform(0.2, 0.5)
        p = rng.uniform(-0.3, 0.1) if pw is None else pw[k]
        s3 = la - s1 - p
        r1 = (s1 * s1 + 1.0) / 4.0
        r2 = ((la - s1) ** 2 + h2 * h2) / 4.0
        r3 = (s3 * s3 + (1.0 - h2) ** 2) / 4.0
        out.append(_cell(la, (r1, r2, r3, rng.uniform(0.01, 0.05))))
    return np.array(out).reshape(-1, 10)


def _check(prover, cells):
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 10)
    assert not max_distances(cells).trap
    got = prover.tr_report_criticals(cells)
    assert got == tr_report(cells)
    return got


def test_synthetic_cells_are_the_intended_kind():
    """The builders give two-disk cells (r3 < 0.1), three-disk cells without
    a trap in the oracle, and pocket widths on both sides of 0."""
    rng = np.random.default_rng(1)
    two = two_disk_cells(rng, 50)
    three = three_disk_cells(rng, 400)
    assert (two[:, 7] < 0.1).all() and (three[:, 7] >= 0.1).all()
    r = max_distances(np.concatenate([two, three]))
    assert r.have_2d and r.have_3d and not r.trap
    assert r.pocket_3d > 0.0
    r = max_distances(three_disk_cells(rng, 50, pw=np.full(50, -0.2)))
    assert -0.21 < r.pocket_3d < 0.0


@pytest.mark.gpu
def test_gpu_tr_report_empty_set(prover):
    assert prover.tr_report_criticals(np.empty((0, 10))) == tr_report(np.empty((0, 10)))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2, 63, 255, 256, 257, 1000, 5000, 100000])
def test_gpu_tr_report_mixed_sets_match_oracle(prover, n):
    rng = np.random.default_rng(n)
    cells = np.concatenate([two_disk_cells(rng, n), three_disk_cells(rng, n)])
    cells = cells[rng.permutation(cells.shape[0])]
    text = _check(prover, cells)
    assert "Two disks" in text and "Three disks" in text


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [3, 4])
def test_gpu_tr_report_one_kind_only(prover, seed):
    rng = np.random.default_rng(seed)
    assert "Three disks" not in _check(prover, two_disk_cells(rng, 700))
    assert "Two disks" not in _check(prover, three_disk_cells(rng, 700))


@pytest.mark.gpu
def test_gpu_tr_report_ties_keep_the_first_cell(prover):
    """Copies of every cell that differ only in r4, which enters none of
    the distances (and r3 for two-disk cells, while below 0.1): each extreme
    is attained by three cells at positions on different lanes, and the
    dominating cells printed (with their r4) must be the first ones."""
    rng = np.random.default_rng(7)
    for make in (two_disk_cells, three_disk_cells):
        base = make(rng, 1000)
        copies = []
        for r4, order in ((0.011, np.arange(1000)), (0.022, rng.permutation(1000)), (0.033, np.arange(1000)[::-1])):
            c = base[order].copy()
            c[:, 8], c[:, 9] = r4, r4 + W
            copies.append(c)
        _check(prover, np.concatenate(copies))
        _check(prover, np.concatenate(copies[::-1]))


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["all_negative_max_twice", "late_nonnegative", "below_minus_one",
                                  "nonnegative_first", "negative_rises_then_zero"])
def test_gpu_tr_report_pocket_restarts(prover, case):
    """The S1/h2/S3 hulls restart at the last cell whose pocket width beats
    a negative running maximum."""
    rng = np.random.default_rng(11)
    n = 1200
    if case == "all_negative_max_twice":
        pw = rng.uniform(-0.3, -0.05, n)
        pw[400] = pw[900] = -0.01
    elif case == "late_nonnegative":
        pw = rng.uniform(-0.3, -0.02, n)
        pw[1100] = 0.05
        pw[1150] = 0.08
    elif case == "below_minus_one":
        pw = np.full(n, -1.2) + rng.uniform(-0.1, 0.0, n)
    elif case == "nonnegative_first":
        pw = rng.uniform(-0.3, 0.1, n)
        pw[0] = 0.02
    else:
        pw = np.linspace(-0.3, -0.001, n)
        pw[-1] = 0.0
    _check(prover, three_disk_cells(rng, n, pw=pw))


@pytest.mark.gpu
def test_gpu_tr_report_nan_terms_are_ignored(prover):
    """Two-disk cells with r1 < 1/4: the scalar sqrt of a negative number is
    NaN, and no NaN enters a running extreme."""
    rng = np.random.default_rng(5)
    cells = two_disk_cells(rng, 600)
    cells[::3, 2] = 0.2
    cells[::3, 3] = 0.2 + W
    _check(prover, cells)


@pytest.mark.gpu
def test_gpu_tr_report_trap_is_an_error(prover):
    """A three-disk cell with 4 r1 - 1 possibly negative reaches the
    reference's interval sqrt trap (ACU/interval.cuh:339-350)."""
    rng = np.random.default_rng(6)
    cells = three_disk_cells(rng, 300)
    cells[123, 2] = 0.2499
    assert max_distances(cells).trap
    with pytest.raises(CCError) as e:
        prover.tr_report_criticals(cells)
    assert e.value.code == CC_E_TRAP


# ------------------------------------------------------------------ oracle pins
# The oracle's Max_distances loop (or_tr_max_distances) against the same
# quantities in exact arithmetic (mpmath, 60 digits), from the formulas of
# tight_rectangle.cu:372-414 and not from oracle/: every extreme must bound
# the exact one from the outside and lie within rounding distance of it
# (R3 and R4, whose interval forms overestimate, only from the outside), and
# each dominating cell must attain its exact extreme.  Each planted slip in
# the exact statement must be caught.


def exact_max_distances(cells, slip=None):
    import mpmath
    from mpmath import mpf

    from criteria_exact import tr_critical_ratio

    def W(la):  # required_weight_for(la) = critical_ratio(la) * la, increasing for la >= 1
        return tr_critical_ratio(la) * la

    def th(n, d):
        # (16 la^4 + 40 la^2 + 9) / (256 la^2) in interval arithmetic: the
        # numerator's la and the denominator's vary independently
        return (16 * n ** 4 + (4 if slip == "th_40" else 40) * n ** 2 + 9) / (256 * d ** 2)

    e = {"two": [], "three": []}
    for i, c in enumerate(np.asarray(cells)):
        la0, la1, r1l, r1h, r2l, r2h, r3l, r3h = (mpf(float(x)) for x in c[:8])
        if c[7] < 0.1:
            r1t = lambda la: (la * la + (-1 if slip == "r1t_minus" else 1)) / 4
            r2 = (r1l, r1h) if slip == "d2_r1" else (r2l, r2h)
            t = {"pos": i + 1, "la": (la0, la1), "d1": (r1l - r1t(la1), r1h - r1t(la0)),
                 "d2": (r2[0] - mpf(1) / 4, r2[1] - mpf(1) / 4),
                 "R3": W(la1) - r1l - (0 if slip == "R3_no_r2" else r2l)}
            a = r1l - (mpf(1) / 2 if slip == "S1_half" else mpf(1) / 4)
            if a > 1e-12:
                S1 = 2 * mpmath.sqrt(a)
                t["q"] = la1 / S1
                half = la1 - S1 / 2
                b = r1l - half * half
                if b > 1e-12:
                    t["h1"] = 2 * mpmath.sqrt(b)
            e["two"].append(t)
        else:
            lam = (la0 + la1) / 2
            q1, q2, q3 = (r1l + r1h) / 2, (r2l + r2h) / 2, (r3l + r3h) / 2
            S1 = mpmath.sqrt(4 * q1 - 1)
            h2 = mpmath.sqrt(4 * q2 - (lam - S1) ** 2)
            S3 = mpmath.sqrt(4 * q3 - (1 - h2) ** 2)
            thl, thh = th(la0, la1), th(la1, la0)
            e["three"].append({"pos": i + 1, "la": (la0, la1), "d1": (r1l - thh, r1h - thl),
                               "d2": (r2l - thh, r2h - thl), "d3": (r3l - thh, r3h - thl),
                               "R4": W(la1) - r1l - r2l - r3l, "pw": lam - S1 - S3})
    return e


def max_distances_mismatches(r, e, cells):
    """The fields of the oracle's result r that disagree with the exact
    statement e (slack: 1e-13 absolute for rounded quantities of size
    about 1; 1e-6 for the interval forms of R3 and R4)."""
    from mpmath import mpf

    eps, bad = mpf("1e-13"), []
    cells = np.asarray(cells)

    def crit_ok(crit, pos, side, key, group, start):
        # the printed cell attains the exact extreme (the zero cell if no
        # cell goes past the start value)
        vals = [t[key][side] for t in group]
        ext = min(vals + [start]) if side == 0 else max(vals + [start])
        if ext == start:
            return not any(crit.v)
        return any(np.array_equal(np.array(crit.v), cells[t["pos"] - 1]) and abs(t[key][side] - ext) <= eps
                   for t in group)

    def hull_ok(got, group, key, start, name):
        lo = min([t[key][0] for t in group] + [start])
        hi = max([t[key][1] for t in group] + [start])
        if not (lo - eps <= got[0] <= lo and hi <= got[1] <= hi + eps):
            bad.append(name)

    two, three = e["two"], e["three"]
    if bool(r.have_2d) != bool(two) or bool(r.have_3d) != bool(three):
        bad.append("have")
    if two:
        la = (min([t["la"][0] for t in two] + [2]), max([t["la"][1] for t in two] + [2]))
        if (r.la_2d[0], r.la_2d[1]) != (float(la[0]), float(la[1])):
            bad.append("la_2d")
        hull_ok(r.d_r1_2d, two, "d1", 0, "d_r1_2d")
        hull_ok(r.d_r2_2d, two, "d2", 0, "d_r2_2d")
        for k, key, name in ((0, "d1", "d_r1_2d_crit"), (1, "d2", "d_r2_2d_crit")):
            crit = r.d_r1_2d_crit if key == "d1" else r.d_r2_2d_crit
            if not (crit_ok(crit[0], 0, 0, key, two, 0) and crit_ok(crit[1], 0, 1, key, two, 0)):
                bad.append(name)
        R3 = max([t["R3"] for t in two] + [0])
        if not (R3 <= r.ub_R3_2d <= R3 + mpf("1e-6")):
            bad.append("ub_R3_2d")
        q = max([t["q"] for t in two if "q" in t] + [1])
        if not (q <= r.ub_la_by_S1 <= q + eps):
            bad.append("ub_la_by_S1")
        h1 = min([t["h1"] for t in two if "h1" in t] + [1])
        if not (h1 - eps <= r.lb_h1_2d <= h1):
            bad.append("lb_h1_2d")
    if three:
        la = (min([t["la"][0] for t in three] + [1]), max([t["la"][1] for t in three] + [1]))
        if (r.la_3d[0], r.la_3d[1]) != (float(la[0]), float(la[1])):
            bad.append("la_3d")
        for key, hull, crit, name in (("d1", r.d_r1_3d, r.d_r1_3d_crit, "d_r1_3d"),
                                      ("d2", r.d_r2_3d, r.d_r2_3d_crit, "d_r2_3d"),
                                      ("d3", r.d_r3_3d, r.d_r3_3d_crit, "d_r3_3d")):
            hull_ok(hull, three, key, 0, name)
            if not (crit_ok(crit[0], 0, 0, key, three, 0) and crit_ok(crit[1], 0, 1, key, three, 0)):
                bad.append(name + "_crit")
        R4 = max([t["R4"] for t in three] + [0])
        if not (R4 <= r.ub_R4_3d <= R4 + mpf("1e-6")):
            bad.append("ub_R4_3d")
        if not (r.pocket_3d >= max([t["pw"] for t in three] + [-1])):
            bad.append("pocket_3d")
    return bad


def _pin_cells():
    rng = np.random.default_rng(21)
    two = two_disk_cells(rng, 120)
    two[::10, 2] = 0.2  # r1 < 1/4: the scalar terms are NaN and ignored
    three = three_disk_cells(rng, 120)
    cells = np.concatenate([two, three])
    return cells[rng.permutation(cells.shape[0])]


def test_oracle_max_distances_equal_exact():
    cells = _pin_cells()
    r = max_distances(cells)
    assert not r.trap
    assert max_distances_mismatches(r, exact_max_distances(cells), cells) == []


@pytest.mark.parametrize("slip", ["th_40", "r1t_minus", "d2_r1", "R3_no_r2", "S1_half"])
def test_oracle_max_distances_planted_slips_are_caught(slip):
    cells = _pin_cells()
    assert max_distances_mismatches(max_distances(cells), exact_max_distances(cells, slip), cells) != []


def test_oracle_max_distances_restart_rule():
    """S1, h2 and S3 restart from a cell's own values when its pocket width
    beats a running maximum that is still negative (tight_rectangle.cu:
    405-413): a cell with a small S1 before the first width >= 0 drops out
    of the hulls, and without any width above -1 nothing restarts and the
    hulls keep their start value 0."""
    rng = np.random.default_rng(8)
    pw = np.full(300, -0.2)
    pw[150] = 0.05
    cells = three_disk_cells(rng, 300, pw=pw)
    la, s1, h2, p = 1.1, 0.05, 0.3, -0.2  # cell 0: S1 = 0.05
    s3 = la - s1 - p
    cells[0] = _cell(la, ((s1 * s1 + 1) / 4, ((la - s1) ** 2 + h2 * h2) / 4, (s3 * s3 + (1 - h2) ** 2) / 4, 0.02))
    S1 = np.sqrt(4 * cells[:, 2] - 1)
    r = max_distances(cells)
    assert not r.trap and r.pocket_3d > 0.04
    assert S1[0] < 0.06 and S1[1:].min() > 0.19
    assert r.S1[0] > 0.19 and r.S1[0] <= S1[150:].min() and r.S1[1] >= np.sqrt(4 * cells[150:, 3] - 1).max()
    # every width below -1: no restart, the hulls start at 0
    r = max_distances(three_disk_cells(rng, 50, pw=np.full(50, -1.3)))
    assert r.pocket_3d == -1.0 and r.S1[0] == 0.0 and r.h2[0] == 0.0 and r.S3[0] == 0.0
