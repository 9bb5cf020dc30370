"""Writes tests/golden/tr_generations.json: the critical records of the
tight_rectangle proof restricted to the initial lambda launches [FIRST,
FIRST+COUNT) of 256, after the initial launches and after each refinement
generation, as the CPU oracle computes them with the reference's driver
(Prover/tight_rectangle/tight_rectangle.cu:458-537: initial counts 128^4,
generation schedule {1,2,2,...} at :50-53, subdivision counts 128,128,128,64
at :45-62, merge_adjacent_r1 per launch at :129-159).

The oracle needs minutes per subdivision launch at the reference's counts,
so the GPU test reads these records instead of recomputing them.  Run from
the repository root: python tests/golden/make_tr_generations.py [GENERATIONS]
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402

FIRST, COUNT = 0, 2
SCHEDULE = (1, 2, 2, 2, 2, 1, 2, 1, 2, 1, 2)
LAMBDA_MAX = 2.0898841580413818342  # tight_rectangle.cu:462


def enc(rows):
    return [[float(x).hex() for x in r] for r in rows]


def main():
    gens = int(sys.argv[1]) if len(sys.argv) > 1 else 3
    t0 = time.time()
    la = oracle.subintervals(1.0, LAMBDA_MAX, 256)
    cur = []
    for a in range(FIRST, FIRST + COUNT):
        c = np.array([la[a][0], la[a][1]] + [0.0, np.finfo(np.float64).max] * 4)
        leaves, trap = oracle.tr_find_critical(c[None], (128, 128, 128, 128))
        assert not trap
        cur.extend(oracle.tr_merge_adjacent_r1(leaves))
    out = {"lambda_launches": [FIRST, COUNT], "generations": [enc(cur)]}
    print(f"generation 0: {len(cur)} records ({time.time() - t0:.0f} s)", flush=True)
    save(out)
    for g in range(1, gens + 1):
        nxt = []
        for s in cur:
            pieces = oracle.subintervals(s[0], s[1], SCHEDULE[g - 1])
            for p in pieces:
                c = np.array(s, dtype=np.float64)
                c[0:2] = p
                leaves, trap = oracle.tr_find_critical(c[None], (128, 128, 128, 64))
                assert not trap
                nxt.extend(oracle.tr_merge_adjacent_r1(leaves))
        cur = nxt
        out["generations"].append(enc(cur))
        print(f"generation {g}: {len(cur)} records ({time.time() - t0:.0f} s)", flush=True)
        save(out)
        if not cur:
            break


def save(out):
    """Writes the generations so far (the file is valid after each one)."""
    path = os.path.join(ROOT, "tests", "golden", "tr_generations.json")
    with open(path + ".tmp", "w") as f:
        json.dump(out, f, indent=0)
    os.replace(path + ".tmp", path)


if __name__ == "__main__":
    main()
