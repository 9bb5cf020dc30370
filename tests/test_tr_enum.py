"""tight_rectangle enumeration (kernel_find_critical_intervals_in,
Prover/tight_rectangle/tight_rectangle.cu:169-262), its host merge
(merge_adjacent_r1, :121-159) and the small branch-and-bound on top of it:
GPU against the CPU oracle, bit-exact after canonical ordering (the
reference's own output order is that of an atomicAdd)."""
import numpy as np
import pytest

import oracle
from circlecover_amd import TR_LAMBDA_MAX, tr_initial_cells

C1_COUNTS = (64, 32, 32, 32)
C1_LEVELS = 3


def c1_cells(n=64):
    return tr_initial_cells(1.0, 1.5, 64)[:n]


def ref_init_cell(i):
    """The i-th of the reference's 256 initial lambda cells (tight_rectangle.cu:460-468)."""
    return tr_initial_cells(1.0, TR_LAMBDA_MAX, 256)[i]


def oracle_bnb(cells, counts, levels):
    leaves, trap = oracle.tr_find_critical(cells, counts)
    per = [leaves.shape[0]]
    for _ in range(levels):
        if leaves.shape[0] == 0:
            break
        leaves, t = oracle.tr_find_critical(oracle.tr_bisect(leaves), (1, 1, 1, 1))
        trap |= t
        per.append(leaves.shape[0])
    return leaves, per, trap


def check_leaf_invariants(leaves, cells):
    """Every leaf lies in one input cell, its radii are tightened
    (r_i.lb >= r_{i+1}.lb) and nonempty."""
    la = leaves[:, 0:2]
    for k in range(4):
        lo, hi = leaves[:, 2 + 2 * k], leaves[:, 3 + 2 * k]
        assert (lo <= hi).all()
        if k < 3:
            assert (lo >= leaves[:, 4 + 2 * k]).all()
    inside = np.zeros(leaves.shape[0], dtype=bool)
    for c in cells:
        ok = (la[:, 0] >= c[0]) & (la[:, 1] <= c[1])
        for k in range(4):
            ok &= (leaves[:, 3 + 2 * k] <= c[3 + 2 * k])
        inside |= ok
    assert inside.all()


# ----------------------------------------------------------------- CPU side
def test_oracle_enumeration_threads_agree_and_invariants():
    cells = c1_cells(6)
    a, ta = oracle.tr_find_critical(cells, (16, 8, 8, 8), threads=1)
    b, tb = oracle.tr_find_critical(cells, (16, 8, 8, 8), threads=4)
    assert (oracle.canonical(a) == oracle.canonical(b)).all() and ta == tb
    assert a.shape[0] > 0
    check_leaf_invariants(a, cells)


def test_oracle_finer_pieces_refine():
    """Refining every piece count can only shrink the unproven region: each
    fine leaf's lambda and r1 lie inside the union of coarse leaves."""
    cells = c1_cells(3)
    coarse, _ = oracle.tr_find_critical(cells, (8, 4, 4, 4))
    fine, _ = oracle.tr_find_critical(cells, (16, 8, 8, 8))
    for f in fine:
        hit = (coarse[:, 0] <= f[0]) & (coarse[:, 1] >= f[1])
        assert hit.any()


def test_bisect_halves_partition_the_parent():
    rng = np.random.default_rng(3)
    cells = c1_cells(4)
    leaves, _ = oracle.tr_find_critical(cells, (8, 4, 4, 4))
    leaves = leaves[rng.permutation(leaves.shape[0])[:200]]
    ch = oracle.tr_bisect(leaves).reshape(-1, 2, 10)
    widths = leaves[:, 1::2] - leaves[:, 0::2]
    dim = np.argmax(widths, axis=1)  # first maximum = lowest index on ties
    for p, (a, b), d in zip(leaves, ch, dim):
        lo, hi = 2 * d, 2 * d + 1
        assert a[lo] == p[lo] and b[hi] == p[hi] and a[hi] == b[lo]
        assert p[lo] < a[hi] < p[hi] or p[lo] == p[hi]
        keep = [j for j in range(10) if j not in (lo, hi)]
        assert (a[keep] == p[keep]).all() and (b[keep] == p[keep]).all()


def test_merge_adjacent_r1_properties():
    cells = c1_cells(2)
    leaves, _ = oracle.tr_find_critical(cells[:1], (16, 8, 8, 8))
    assert leaves.shape[0] > 10
    merged = oracle.tr_merge_adjacent_r1(leaves[np.random.default_rng(1).permutation(leaves.shape[0])])
    # r1 intervals of different merged records neither overlap nor touch
    r1 = merged[np.argsort(merged[:, 2]), 2:4]
    assert (r1[1:, 0] > r1[:-1, 1]).all()
    # every leaf is inside exactly one merged record (all four radii)
    for v in leaves:
        ok = np.ones(merged.shape[0], dtype=bool)
        for k in range(4):
            ok &= (merged[:, 2 + 2 * k] <= v[2 + 2 * k]) & (merged[:, 3 + 2 * k] >= v[3 + 2 * k])
        assert ok.sum() == 1


def test_oracle_c1_runs():
    leaves, per, trap = oracle_bnb(c1_cells(8), C1_COUNTS, C1_LEVELS)
    assert not trap and per[0] > 0
    check_leaf_invariants(leaves, c1_cells(8))


# ----------------------------------------------------------------- GPU side
def _same(gpu, cpu):
    g, c = oracle.canonical(gpu), oracle.canonical(cpu)
    assert g.shape == c.shape, (g.shape, c.shape)
    bad = np.flatnonzero((g != c).any(axis=1))
    assert bad.size == 0, f"{bad.size} leaves differ; first gpu {g[bad[0]].tolist()} cpu {c[bad[0]].tolist()}"


@pytest.mark.gpu
def test_gpu_tr_enumerate_c1_cells(prover):
    cells = c1_cells(16)
    g = prover.tr_enumerate(cells, C1_COUNTS)
    c, trap = oracle.tr_find_critical(cells, C1_COUNTS)
    assert not trap
    _same(g, c)


@pytest.mark.gpu
def test_gpu_tr_c1_branch_and_bound(prover):
    cells = c1_cells()
    g, gper = prover.tr_bnb(cells, C1_COUNTS, C1_LEVELS)
    c, cper, trap = oracle_bnb(cells, C1_COUNTS, C1_LEVELS)
    assert not trap
    assert list(gper[: len(cper)]) == cper
    _same(g, c)


@pytest.mark.gpu
@pytest.mark.parametrize("i", [0, 10])
def test_gpu_tr_reference_init_launch(prover, i):
    """One launch of the reference's initialisation, 128^4 pieces."""
    cell = ref_init_cell(i)
    g = prover.tr_find_critical_in(cell, is_init=True)
    c, trap = oracle.tr_find_critical(cell[None], (128, 128, 128, 128))
    assert not trap and c.shape[0] > 0
    _same(g, c)


@pytest.mark.gpu
def test_gpu_tr_reference_subdivision_launch(prover):
    """A generation-1 launch (counts 128,128,128,64) on a merged critical
    record of the initialisation, as tight_rectangle.cu:476-496 does."""
    init, _ = oracle.tr_find_critical(ref_init_cell(0)[None], (128, 128, 128, 128))
    merged = oracle.tr_merge_adjacent_r1(init)
    cell = merged[0]
    g = prover.tr_find_critical_in(cell, is_init=False)
    c, trap = oracle.tr_find_critical(cell[None], (128, 128, 128, 64))
    assert not trap
    _same(g, c)


@pytest.mark.gpu
def test_gpu_tr_overflow_protocol(prover):
    import ctypes

    from circlecover_amd.native import CC_E_OVERFLOW, lib

    cell = np.ascontiguousarray(c1_cells(1)[0])
    cnt = np.array(C1_COUNTS, dtype=np.int32)
    out = np.empty((1, 10))
    n = ctypes.c_size_t()
    rc = lib.cc_tr_enumerate(prover.ctx, cell.ctypes.data, 1, cnt.ctypes.data, out.ctypes.data, 1, ctypes.byref(n))
    full = prover.tr_enumerate(cell[None], C1_COUNTS)
    assert full.shape[0] > 1
    assert rc == CC_E_OVERFLOW and n.value == full.shape[0]


@pytest.mark.gpu
def test_gpu_tr_empty_batch(prover):
    assert prover.tr_enumerate(np.empty((0, 10)), C1_COUNTS).shape[0] == 0


def merge_by_r1_piece(leaves):
    """One launch's merge as the device path forms it: the hull of the
    leaves sharing an r1 upper bound (one r1 piece), then the sweep of those
    hulls in order of their least r1 lower bound, joined where it is at most
    the running r1 upper bound."""
    ub = leaves[:, 3]
    keys, inv = np.unique(ub, return_inverse=True)
    hulls = np.empty((keys.size, 10))
    for k in range(keys.size):
        g = leaves[inv == k]
        hulls[k, 0:2] = g[0, 0:2]
        hulls[k, 2::2] = g[:, 2::2].min(axis=0)
        hulls[k, 3::2] = g[:, 3::2].max(axis=0)
    hulls = hulls[np.argsort(hulls[:, 2], kind="stable")]
    out = []
    for h in hulls:
        if out and h[2] <= out[-1][3]:
            out[-1][2::2] = np.minimum(out[-1][2::2], h[2::2])
            out[-1][3::2] = np.maximum(out[-1][3::2], h[3::2])
        else:
            out.append(h.copy())
    return np.array(out).reshape(-1, 10)


@pytest.mark.parametrize("which", ["initial", "initial_fine_r1", "generation"])
def test_merge_by_r1_piece_is_the_reference_merge(which):
    """The premise of the device-side merge: every leaf of a launch keeps its
    r1 piece's upper bound and r1.lb <= r1.ub, so the hulls per piece swept
    in order give merge_adjacent_r1's records, bit for bit (oracle leaves of
    several launches; also with every other r1 piece's leaves left out,
    which opens gaps, so that the merge ends in several records)."""
    if which == "initial":
        launches = [ref_init_cell(i) for i in (0, 1, 40, 100)]
        counts = (32, 32, 32, 32)
    elif which == "initial_fine_r1":
        launches = [ref_init_cell(i) for i in (0, 100)]
        counts = (128, 16, 16, 16)
    else:
        first, _ = oracle.tr_find_critical(ref_init_cell(0)[None], (32, 32, 32, 32))
        launches = list(oracle.tr_merge_adjacent_r1(first))[:4]
        counts = (16, 16, 16, 8)
    most = 0
    for cell in launches:
        leaves, trap = oracle.tr_find_critical(np.asarray(cell)[None], counts)
        assert not trap
        if leaves.shape[0] == 0:
            continue
        assert (leaves[:, 2] <= leaves[:, 3]).all()
        _, piece = np.unique(leaves[:, 3], return_inverse=True)
        for part in (leaves, leaves[piece % 2 == 0]):
            ref = oracle.tr_merge_adjacent_r1(part)
            got = merge_by_r1_piece(part)
            assert got.shape == ref.shape
            assert (got.view(np.uint64) == ref.view(np.uint64)).all()
            most = max(most, ref.shape[0])
    assert most >= 2
