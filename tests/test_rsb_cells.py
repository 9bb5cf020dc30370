"""rectangle_size_bound leaf-cell verdicts (7-D cells: lambda, r1..r6): the GPU
chain against the CPU oracle, bit-exact, on cells drawn from the reference's
grid (Prover/rectangle_size_bound/rectangle_size_bound.cu:171-310).

Two views are compared: the first proving step (the reference's
short-circuit order, rectangle_size_bound.cu:66-168) and the mask of every
step evaluated on its own, which pins each criterion separately even where
an earlier one would have proved the cell."""
import numpy as np
import pytest

import oracle
from rsb_cells import leaf_cells

N_STEPS = 20
DISTRIBUTIONS = {
    "uniform": dict(seed=11, n=6000),
    "large_r_small_la": dict(seed=12, n=6000, la_range=(0, 24), r_bias=0.35),
    "larger_r_la40": dict(seed=13, n=6000, la_range=(0, 40), r_bias=0.6),
    "top_corner": dict(seed=14, n=4000, la_range=(0, 8), r_bias=0.15),
}

_cache = {}


def cells_of(name):
    if name not in _cache:
        d = dict(DISTRIBUTIONS[name])
        _cache[name] = leaf_cells(d.pop("seed"), d.pop("n"), **d)
    return _cache[name]


def all_cells():
    return np.concatenate([cells_of(k) for k in DISTRIBUTIONS])


def edge_cells():
    """Point cells, empty radius intervals, zero radii, lambda at both ends of
    its range and at the tolerated maximum."""
    rows = []
    for la in ((1.0, 1.0), (1.0, 1.0 + 2 ** -40), (2.0898841580413813, 2.0898841580413818), (2.5, 2.5), (1.5, 1.6)):
        for r in ((0.14, 0.13, 0.05, 0.04, 0.03, 0.02), (0.0,) * 6, (9 / 64, 9 / 64, 9 / 64, 9 / 64, 9 / 64, 9 / 64),
                  (0.1, 0.1, 0.1, 0.1, 0.1, 0.1), (0.12, 0.02, 0.0, 0.0, 0.0, 0.0)):
            for w in (0.0, 2.0 ** -20, 2.0 ** -7):
                row = list(la)
                for x in r:
                    row += [x, x + w]
                rows.append(row)
    # empty intervals: r3.lb > r3.ub, and lambda empty
    rows.append([1.2, 1.3, 0.1, 0.11, 0.05, 0.06, 0.05, 0.04, 0.01, 0.02, 0.0, 0.01, 0.0, 0.01])
    rows.append([1.3, 1.2, 0.1, 0.11, 0.05, 0.06, 0.03, 0.04, 0.01, 0.02, 0.0, 0.01, 0.0, 0.01])
    return np.array(rows, dtype=np.float64)


def lowest_bit(mask):
    m = mask & np.uint64((1 << N_STEPS) - 1)
    out = np.full(m.shape, -1, dtype=np.int64)
    for k in reversed(range(N_STEPS)):
        out[(m >> np.uint64(k)) & np.uint64(1) == 1] = k
    return out


def test_oracle_verdict_is_first_step_of_mask():
    cells = np.concatenate([all_cells(), edge_cells()])
    v, trap = oracle.rsb_verdicts(cells)
    m = oracle.rsb_masks(cells)
    assert (v == lowest_bit(m)).all()


def test_oracle_every_step_fires_and_fails():
    m = oracle.rsb_masks(all_cells())
    bits = ((m[:, None] >> np.arange(N_STEPS).astype(np.uint64)) & np.uint64(1)).astype(bool)
    assert bits.any(axis=0).all(), np.flatnonzero(~bits.any(axis=0))
    assert (~bits).any(axis=0).all(), np.flatnonzero(~(~bits).any(axis=0))


def test_oracle_threads_agree():
    cells = cells_of("large_r_small_la")
    a = oracle.rsb_verdicts(cells, threads=1)
    b = oracle.rsb_verdicts(cells, threads=5)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
    assert (oracle.rsb_masks(cells, threads=1) == oracle.rsb_masks(cells, threads=3)).all()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(DISTRIBUTIONS))
def test_gpu_rsb_verdicts_match_oracle(prover, name):
    cells = cells_of(name)
    gv, gt = prover.rsb_eval_cells(cells)
    ov, ot = oracle.rsb_verdicts(cells)
    bad = np.flatnonzero((gv != ov) | (gt != ot))
    assert bad.size == 0, f"{bad.size} cells differ; first {bad[:5]}: gpu {gv[bad[:5]]} oracle {ov[bad[:5]]}"


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(DISTRIBUTIONS))
def test_gpu_rsb_masks_match_oracle(prover, name):
    cells = cells_of(name)
    gm = prover.rsb_eval_masks(cells)
    om = oracle.rsb_masks(cells)
    diff = gm ^ om
    bad = np.flatnonzero(diff)
    steps = sorted({int(k) for k in range(64) if ((diff >> np.uint64(k)) & np.uint64(1)).any()})
    assert bad.size == 0, f"{bad.size} cells differ in steps {steps}; first {bad[:3]}: {cells[bad[:1]].tolist()}"


@pytest.mark.gpu
def test_gpu_rsb_edge_cells(prover):
    cells = edge_cells()
    gv, gt = prover.rsb_eval_cells(cells)
    ov, ot = oracle.rsb_verdicts(cells)
    assert (gv == ov).all() and (gt == ot).all()
    assert (prover.rsb_eval_masks(cells) == oracle.rsb_masks(cells)).all()


@pytest.mark.gpu
def test_gpu_rsb_empty_batch(prover):
    gv, gt = prover.rsb_eval_cells(np.empty((0, 14)))
    assert gv.size == 0 and gt.size == 0
