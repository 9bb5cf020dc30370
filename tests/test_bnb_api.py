"""cc_run_bnb, the branch-and-bound in one call: every form returns the CPU
oracle's unproven cells in lexicographic order with the oracle's statistics,
from caller cells or from a grid, and cc_interrupt stops a running call
without affecting the next one (the reference clears its SIGINT flag when an
interruptible computation starts, Prover/interrupt.cpp:38-41)."""
import threading

import numpy as np
import pytest

import oracle
from circlecover_amd import rsb_grid
from circlecover_amd.native import CC_BNB_BATCHED, CC_BNB_PLAIN, CC_BNB_PLAIN_EXACT, CC_E_INTERRUPTED, CCError
from test_rsb_bnb import C3_DEPTH, c3_cells


def _lexsorted(cells):
    keys = [cells[:, k] for k in range(13, -1, -1)]  # (lb, ub) of lambda, r1..r6, most significant first
    return np.array_equal(cells, cells[np.lexsort(keys)])


@pytest.mark.gpu
@pytest.mark.parametrize("form", [CC_BNB_PLAIN, CC_BNB_BATCHED, CC_BNB_PLAIN_EXACT])
def test_gpu_run_bnb_matches_oracle(prover, form):
    print(f"run_bnb form {form}: 4096 sampled cells, depth {C3_DEPTH}", flush=True)
    cells = c3_cells(4096, seed=11)
    got, st = prover.run_bnb(C3_DEPTH, form, cells=cells)
    want, per, hist, trap = oracle.rsb_bnb(cells, C3_DEPTH)
    assert np.array_equal(got, oracle.canonical(want)) and _lexsorted(got)
    d = st.depth_reached
    assert list(st.cells_per_depth)[: d + 1] == [int(x) for x in per[: d + 1]]
    assert list(st.verdict_hist) == [int(x) for x in hist.sum(axis=0)] and st.trap_cells == trap
    # the same run from the configuration's grid
    grid = (1.0, 8.0, 32, 8, 8)
    got_g, st_g = prover.run_bnb(12, form, grid=grid)
    want_g, per_g, _, _ = oracle.rsb_bnb(rsb_grid(1.0, 8.0, 32, 8, 8), 12)
    assert np.array_equal(got_g, oracle.canonical(want_g))
    assert list(st_g.cells_per_depth)[: st_g.depth_reached + 1] == [int(x) for x in per_g[: st_g.depth_reached + 1]]


@pytest.mark.gpu
def test_gpu_interrupt_stops_only_the_running_call(prover):
    """A thread keeps calling cc_interrupt while the level-synchronous
    branch-and-bound of the full 256x64x64 grid runs: the call stops at a
    level boundary with CC_E_INTERRUPTED.  The next call on the same context
    runs to the end (the flag is cleared when a call starts)."""
    cells = c3_cells(2048, seed=5)
    want, _ = prover.run_bnb(C3_DEPTH, CC_BNB_PLAIN, cells=cells)
    prover.interrupt()  # set before a call starts: has no effect on it
    again, _ = prover.run_bnb(C3_DEPTH, CC_BNB_PLAIN, cells=cells)
    assert np.array_equal(again, want)
    done = threading.Event()

    def poke():
        while not done.is_set():
            prover.interrupt()
            done.wait(0.005)

    t = threading.Thread(target=poke)
    t.start()
    try:
        with pytest.raises(CCError) as e:
            prover.run_bnb(C3_DEPTH, CC_BNB_PLAIN, grid=(1.0, 8.0, 256, 64, 64))
    finally:
        done.set()
        t.join()
    assert e.value.code == CC_E_INTERRUPTED
    after, _ = prover.run_bnb(C3_DEPTH, CC_BNB_PLAIN, cells=cells)
    assert np.array_equal(after, want)
