"""Writes tests/golden/c14_cells.json: point cells in the regime of the
rectangle_size_bound leaf step C14 (seven_disk_strategies), as inputs for
the exact-arithmetic pins of tests/test_criteria_pins.py.

- `bnb`: the centres of cells that reach C14 in the branch-and-bound over the
  256x64x64 grid of lambda in [1, 8] (depths 8..15 of a seeded 20,000-cell
  sample; a cell reaches C14 when the CPU oracle's chain proves it by C14 or
  leaves it open), radii sorted, with weight left over.
- `different_heights`, `two_two_two`: cells and r7 pieces on which the
  different-height strategy (seven_disk_strategies.cu:1248-1303) or the
  2-2-2 placement (:1192-1246) covers the piece, found by a seeded random
  search in exact arithmetic (both rarely decide a piece of a
  branch-and-bound cell).

Only inputs are stored; the test derives every expected verdict from the
exact statement.  Run from the repository root:
python tests/golden/make_c14_cells.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import mpmath  # noqa: E402
import numpy as np  # noqa: E402

import criteria_exact as ce  # noqa: E402
import oracle  # noqa: E402

N_BNB = 160
N_DH = 24


def bnb_cells():
    grid = oracle.rsb_grid([1.0, 8.0] + [0.0, 9.0 / 64.0] * 6, 256, 64, 64)
    rng = np.random.default_rng(5)
    cells = grid[rng.choice(len(grid), 20000, replace=False)]
    pool = []
    for depth in range(16):
        v, _ = oracle.rsb_verdicts(cells)
        if depth >= 8:
            pool.append(cells[(v == 19) | (v == -1)])
        live = cells[v == -1]
        if len(live) > 20000:
            live = live[rng.choice(len(live), 20000, replace=False)]
        cells = oracle.rsb_bisect(live)
    pool = np.concatenate(pool)
    pool = pool[rng.choice(len(pool), 2 * N_BNB, replace=False)]
    P = np.column_stack([(pool[:, 2 * k] + pool[:, 2 * k + 1]) / 2 for k in range(7)])
    P[:, 1:] = np.sort(P[:, 1:], axis=1)[:, ::-1]
    return P[P[:, 0] * 0.61 - P[:, 1:].sum(axis=1) > 1e-6][:N_BNB]


def rare_strategy_cells(strategy, seed):
    """N_DH cells with an r7 piece that `strategy` covers at clear slack."""
    rng = np.random.default_rng(seed)
    hits = []
    while len(hits) < N_DH:
        la = rng.uniform(1.0, 3.0)
        r = np.sort(rng.dirichlet(rng.uniform(0.3, 5, 6)) * la * 0.61 * rng.uniform(0.3, 0.95))[::-1]
        R = la * 0.61 - r.sum()
        top = min(r[5], R)
        a = rng.uniform(0.5, 1) * top
        b = a + rng.uniform(0, 1) * (top - a)
        D = ce.Dec()
        if strategy(D, mpmath.mpf(la), [mpmath.mpf(x) for x in r], mpmath.mpf(R),
                    [mpmath.mpf(a), mpmath.mpf(b)]) and D.slack > 1e-9:
            hits.append([la, *r, a, b])
    return np.array(hits)


def main():
    out = {"bnb": [[x.hex() for x in row] for row in bnb_cells()],
           "different_heights": [[x.hex() for x in row] for row in rare_strategy_cells(ce._different_height, 9)],
           "two_two_two": [[x.hex() for x in row] for row in rare_strategy_cells(ce._six_222, 10)]}
    with open(os.path.join(ROOT, "tests", "golden", "c14_cells.json"), "w") as f:
        json.dump(out, f, indent=0)
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
