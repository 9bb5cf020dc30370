"""rectangle_size_bound drivers: the reference's initial grid
(rectangle_size_bound.cu:276-310), its per-cell kernel enumeration
(:161-236) and the widest-dimension branch-and-bound over 7-D cells.  GPU
against the CPU oracle, bit-exact: unproven sets after canonical ordering,
cells per depth and verdict histograms."""
import numpy as np
import pytest

import oracle
from circlecover_amd import rsb_grid

W = 9.0 / 64.0
C3_GRID = (256, 64, 64)
C3_DEPTH = 24


def c3_cells(n=None, seed=0):
    """The 256x64x64 grid over lambda in [1, 8], or a seeded sample of n cells
    (sorted, so the sample keeps the grid's order)."""
    g = rsb_grid(1.0, 8.0, *C3_GRID)
    if n is None:
        return g
    return g[np.sort(np.random.default_rng(seed).choice(g.shape[0], n, replace=False))]


def ref_grid():
    return rsb_grid(1.0, 2.5, 256, 128, 128)


def hard_ref_cells(n, seed=0):
    """Reference initial cells that the cell shortcuts S1/S2 do not prove."""
    g = ref_grid()[::41]
    m = oracle.rsb_masks(g)
    hard = g[(m & np.uint64(3)) == 0]
    rng = np.random.default_rng(seed)
    return hard[rng.choice(hard.shape[0], n, replace=False)]


# ----------------------------------------------------------------- CPU side
def test_host_grid_matches_oracle():
    for la_hi, g in ((8.0, C3_GRID), (2.5, (32, 16, 16))):
        a = rsb_grid(1.0, la_hi, *g)
        b = oracle.rsb_grid(np.array([1.0, la_hi] + [0.0, W] * 6), *g)
        assert a.shape == b.shape and (a == b).all()


def test_grid_cells_are_tightened_and_nested():
    g = c3_cells(10000)
    assert g.shape[0] > 0
    for k in range(1, 5):
        assert (g[:, 2 + 2 * k] >= g[:, 4 + 2 * k]).all()  # radii[k].lb >= radii[k+1].lb
    assert (g[:, 3] >= g[:, 5]).all()  # r2 within r1's range
    for k in range(2, 6):
        assert (g[:, 3 + 2 * k] <= g[:, 5]).all()  # r3..r6 bounded by r2.ub


def test_oracle_bnb_bookkeeping():
    cells = c3_cells(3000, seed=1)
    un, per, hist, trap = oracle.rsb_bnb(cells, C3_DEPTH, threads=4)
    un1, per1, hist1, trap1 = oracle.rsb_bnb(cells, C3_DEPTH, threads=1)
    assert (per == per1).all() and (hist == hist1).all() and trap == trap1
    assert (hist.sum(axis=1) == per).all()
    # every unproven cell below the limit is split in two
    assert (per[1:] == 2 * hist[:-1, 0]).all()
    assert un.shape[0] == hist[C3_DEPTH, 0]


def test_oracle_bnb_small_depth_reports_unproven():
    cells = c3_cells(3000, seed=2)
    un, per, hist, _ = oracle.rsb_bnb(cells, 2)
    assert un.shape[0] == hist[2, 0] > 0


def test_oracle_enumeration_leaves_inside_cell():
    cells = hard_ref_cells(6)
    for c in cells:
        leaves, _ = oracle.rsb_find_critical(c, (4, 4, 4, 4))
        for k in range(7):
            assert (leaves[:, 2 * k] >= c[2 * k] - 0.0).all() or k > 0
            assert (leaves[:, 2 * k + 1] <= c[2 * k + 1]).all()


# ----------------------------------------------------------------- GPU side
def _same(gpu, cpu):
    g, c = oracle.canonical(gpu), oracle.canonical(cpu)
    assert g.shape == c.shape, (g.shape, c.shape)
    assert (g == c).all()


@pytest.mark.gpu
@pytest.mark.parametrize("n,depth", [(16384, C3_DEPTH), (4096, 3), (1024, 0)])
def test_gpu_rsb_bnb_matches_oracle(prover, n, depth):
    cells = c3_cells(n, seed=n)
    g, st = prover.rsb_bnb(cells, depth)
    c, per, hist, trap = oracle.rsb_bnb(cells, depth)
    d = st.depth_reached
    assert list(st.cells_per_depth)[: d + 1] == [int(x) for x in per[: d + 1]]
    assert int(per[d + 1:].sum()) == 0
    assert list(st.verdict_hist) == [int(x) for x in hist.sum(axis=0)]
    assert st.trap_cells == trap
    _same(g, c)


@pytest.mark.gpu
def test_gpu_rsb_enumerate_coarse_pieces(prover):
    cells = hard_ref_cells(40, seed=2)
    counts = (4, 4, 4, 4)
    g, gtrap = prover.rsb_enumerate(cells, counts, with_trap=True)
    res = [oracle.rsb_find_critical(x, counts) for x in cells]
    c = np.concatenate([r[0] for r in res])
    assert c.shape[0] > 0
    assert gtrap == any(r[1] for r in res)
    _same(g, c)


@pytest.mark.gpu
@pytest.mark.parametrize("is_init", [True, False])
def test_gpu_rsb_reference_launch(prover, is_init):
    counts = (64, 64, 64, 64) if is_init else (16, 16, 16, 32)
    for cell in hard_ref_cells(3, seed=3):
        g = prover.rsb_find_critical_in(cell, is_init)
        c, trap = oracle.rsb_find_critical(cell, counts)
        assert not trap
        _same(g, c)
