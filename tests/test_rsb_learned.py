"""Learned step order in the batched branch-and-bound: the prover's own
records of (cell, first succeeding step, depth), a ranker trained on them
on the GPU, and the batched run that tries each cell's predicted step first.
The interval criteria stay the only judge: the unproven cells and the cells
per depth equal the CPU oracle's whatever the ranker predicts."""
import ctypes

import numpy as np
import pytest

import oracle
from circlecover_amd.native import RsbRanker, RANKER_PARAMS
from test_rsb_bnb import C3_DEPTH, c3_cells


def _same_tree(st, per, unproven, oracle_unproven):
    d = st.depth_reached
    assert list(st.cells_per_depth)[: d + 1] == [int(x) for x in per[: d + 1]]
    assert int(per[d + 1:].sum()) == 0
    assert np.array_equal(oracle.canonical(unproven), oracle.canonical(oracle_unproven))


@pytest.mark.gpu
def test_gpu_bnb_log_records_every_cell(prover):
    """The log holds one record per evaluated cell: its depth histogram is
    cells_per_depth, its steps are the verdict histogram, and the oracle's
    verdict of each logged cell is its logged first step."""
    cells = c3_cells(2048, seed=21)
    _, st, log = prover.rsb_bnb_log(cells, C3_DEPTH)
    d = st.depth_reached
    assert len(log) == sum(st.cells_per_depth[: d + 1])
    assert list(np.bincount(log["depth"], minlength=d + 1)) == list(st.cells_per_depth[: d + 1])
    assert list(np.bincount(log["first_step"] + 1, minlength=21)) == list(st.verdict_hist)
    sample = log[np.random.default_rng(3).choice(len(log), 3000, replace=False)]
    v, _ = oracle.rsb_verdicts(sample["cell"])
    assert np.array_equal(v, sample["first_step"])


@pytest.mark.gpu
def test_gpu_learned_order_keeps_the_tree(prover):
    """A ranker trained on the records of one sample of the grid, used on
    another: same cells per depth and unproven cells as the oracle, every
    cell proven by some step; the ranker predicts most first steps."""
    _, _, log = prover.rsb_bnb_log(c3_cells(8192, seed=31), C3_DEPTH)
    model, losses = prover.rsb_ranker_train(log, epochs=20, batch=4096, seed=1)
    assert np.isfinite(losses).all() and losses[-1] < losses[0]
    steps = list(model.class_step)
    assert steps[10] == -1 and all(-1 <= s <= 18 for s in steps)
    cells = c3_cells(8192, seed=32)
    got, st = prover.rsb_bnb_learned(cells, C3_DEPTH, model)
    want, per, hist, _ = oracle.rsb_bnb(cells, C3_DEPTH)
    _same_tree(st, per, got, want)
    plain = hist.sum(axis=0)
    assert st.verdict_hist[0] == plain[0]  # unproven at their depth: the same cells
    assert sum(st.verdict_hist) == sum(plain)


@pytest.mark.gpu
def test_gpu_learned_order_with_an_arbitrary_ranker(prover):
    """Random weights and a reversed step table (C13 first, ...): the
    predictions are arbitrary, the tree is still the oracle's."""
    rng = np.random.default_rng(5)
    model = RsbRanker()
    p = rng.uniform(-1, 1, RANKER_PARAMS).astype(np.float32)
    ctypes.memmove(model.net.params, p.ctypes.data, p.nbytes)
    norm = np.array([1.7, -9, 0.07, -9, 0.05, -9, 0.04, -9] + [1.0] * 8, dtype=np.float32)
    ctypes.memmove(model.net.norm, norm.ctypes.data, norm.nbytes)
    for c, s in enumerate([18, 17, 16, 15, 14, 13, 12, 8, 5, 0, -1]):
        model.class_step[c] = s
    cells = c3_cells(4096, seed=41)
    got, st = prover.rsb_bnb_learned(cells, C3_DEPTH, model)
    want, per, _, _ = oracle.rsb_bnb(cells, C3_DEPTH)
    _same_tree(st, per, got, want)
