"""Reference directed-operation counts of profiled workloads — TEST
INFRASTRUCTURE (the oracle's counting build, run on the CPU).

Counts the directed FP64 operations the reference's sequence performs on a
seeded sample of a workload's units, with the oracle built with op counting
(one thread, so the calling thread's counter sees all of it).  The
profiles under profiles/ divide these counts by the kernel times rocprofv3
measures for the same units:

  python3 oracle/count_ops.py c3 N SEED           subtrees (depth 24) of N seeded initial
                                                 cells of the 256x64x64 grid over [1, 8]
  python3 oracle/count_ops.py rsb-enum F C N SEED  N seeded cells of lambda pieces [F, F+C)
                                                 of the rectangle_size_bound proof grid,
                                                 enumerated with 64^4 pieces
  python3 oracle/count_ops.py tr-enum F C        tight_rectangle initial launches [F, F+C)
                                                 of 256, 128^4 pieces
"""
import ctypes
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np  # noqa: E402

import oracle  # noqa: E402

LAMBDA_MAX_TR = 2.0898841580413818342  # tight_rectangle.cu:462


def _counting_call(fn):
    """Runs fn() on the counting build and returns (result, operations)."""
    L = oracle.count_lib()
    L.or_op_count.restype = ctypes.c_uint64
    L.or_op_count_reset.restype = None
    saved = oracle.lib
    oracle.lib = L  # route the wrappers to the counting build
    try:
        L.or_op_count_reset()
        res = fn()
        return res, int(L.or_op_count())
    finally:
        oracle.lib = saved


def rsb_grid(outer, nla, nr1, nr2):
    return oracle.rsb_grid(np.asarray(outer, dtype=np.float64), nla, nr1, nr2)


def c3(n, seed):
    g = rsb_grid([1.0, 8.0] + [0.0, 9.0 / 64.0] * 6, 256, 64, 64)
    cells = g[np.sort(np.random.default_rng(seed).choice(len(g), n, replace=False))]
    (unproven, per, _, _), ops = _counting_call(lambda: oracle.rsb_bnb(cells, 24, threads=1))
    evaluated = int(per.sum())
    return {"workload": "c3", "initial_cells": n, "grid_cells": len(g), "seed": seed, "cells_evaluated": evaluated,
            "ops": ops, "ops_per_initial_cell": ops / n, "ops_per_evaluated_cell": ops / evaluated}


def rsb_enum(first, count, n, seed):
    pieces = oracle.subintervals(1.0, 2.5, 256)
    cells = []
    for a in range(first, first + count):
        outer = np.array([pieces[a][0], pieces[a][1]] + [0.0, 9.0 / 64.0] * 6)
        cells.append(rsb_grid(outer, 1, 128, 128))
    cells = np.concatenate(cells)
    sample = cells[np.sort(np.random.default_rng(seed).choice(len(cells), n, replace=False))]
    total, leaves = 0, 0
    for c in sample:
        (lv, _), ops = _counting_call(lambda: oracle.rsb_find_critical(c, (64, 64, 64, 64), threads=1))
        total += ops
        leaves += len(lv)
    return {"workload": "rsb-enum", "lambda_pieces": [first, count], "slice_cells": len(cells), "sample": n,
            "seed": seed, "ops_sample": total, "ops_per_cell": total / n, "unproven_leaves_sample": leaves}


def tr_enum(first, count):
    pieces = oracle.subintervals(1.0, LAMBDA_MAX_TR, 256)
    cells = np.array([[pieces[a][0], pieces[a][1]] + [0.0, np.finfo(np.float64).max] * 4
                      for a in range(first, first + count)])
    (leaves, _), ops = _counting_call(lambda: oracle.tr_find_critical(cells, (128, 128, 128, 128), threads=1))
    return {"workload": "tr-enum", "launches": [first, count], "ops": ops, "ops_per_launch": ops / count,
            "unproven_leaves": int(len(leaves))}


def main(argv):
    t = time.time()
    kind, args = argv[1], [int(x) for x in argv[2:]]
    out = {"c3": c3, "rsb-enum": rsb_enum, "tr-enum": tr_enum}[kind](*args)
    out["cpu_s"] = time.time() - t
    print(json.dumps(out))


if __name__ == "__main__":
    main(sys.argv)
