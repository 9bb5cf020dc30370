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
for byte with the oracle's."""
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
@pytest.mark.parametrize("n", [1, 2, 63, 255, 256, 257, 1000, 5000])
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
