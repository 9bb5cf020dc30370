"""Batched branch-and-bound (persistent stage kernels, device work queues,
hipGraph over depths, exact bisection shortcuts): the same unproven cells,
cells per depth, verdict histogram and trap count as the CPU oracle and as
the level-synchronous GPU driver."""
import numpy as np
import pytest

import oracle
from test_rsb_bnb import C3_DEPTH, c3_cells


def _same_stats(a, b):
    assert a.depth_reached == b.depth_reached
    d = a.depth_reached
    assert list(a.cells_per_depth)[: d + 1] == list(b.cells_per_depth)[: d + 1]
    assert list(a.verdict_hist) == list(b.verdict_hist)
    assert a.trap_cells == b.trap_cells


@pytest.mark.gpu
@pytest.mark.parametrize("n,depth,params", [
    (16384, C3_DEPTH, None),
    (16384, C3_DEPTH, (64, 64, 1, 0, 1)),
    (16384, C3_DEPTH, (256, 128, 8, 1, 4)),
    (4096, 3, (1024, 128, 4, 1, 2)),
    (1024, 0, (128, 64, 2, 1, 1)),
])
def test_gpu_batched_matches_oracle(prover, n, depth, params):
    cells = c3_cells(n, seed=n)
    g, st = prover.rsb_bnb_batched(cells, depth, params)
    c, per, hist, trap = oracle.rsb_bnb(cells, depth)
    d = st.depth_reached
    assert list(st.cells_per_depth)[: d + 1] == [int(x) for x in per[: d + 1]]
    assert int(per[d + 1:].sum()) == 0
    assert list(st.verdict_hist) == [int(x) for x in hist.sum(axis=0)]
    assert st.trap_cells == trap
    assert (oracle.canonical(g) == oracle.canonical(c)).all()


@pytest.mark.gpu
def test_gpu_batched_matches_plain_full_grid(prover):
    cells = c3_cells()
    a, sa = prover.rsb_bnb(cells, C3_DEPTH)
    b, sb = prover.rsb_bnb_batched(cells, C3_DEPTH)
    _same_stats(sa, sb)
    assert (oracle.canonical(a) == oracle.canonical(b)).all()
    # the graph is reused on a second call with the same shape
    c, sc = prover.rsb_bnb_batched(cells, C3_DEPTH)
    _same_stats(sa, sc)
    assert (oracle.canonical(a) == oracle.canonical(c)).all()


@pytest.mark.gpu
def test_gpu_autotune_keeps_results(prover):
    cells = c3_cells(8192, seed=5)
    best, ms = prover.rsb_autotune(cells, 12)
    assert best[0] % 64 == 0 and ms > 0
    a, sa = prover.rsb_bnb(cells, 12)
    b, sb = prover.rsb_bnb_batched(cells, 12)
    _same_stats(sa, sb)
    assert (oracle.canonical(a) == oracle.canonical(b)).all()


@pytest.mark.gpu
def test_gpu_batched_empty(prover):
    g, st = prover.rsb_bnb_batched(np.empty((0, 14)), 5)
    assert g.shape[0] == 0 and st.depth_reached == -1
