"""The RSB enumeration lists a row's open (non-empty) leaves by a binary
search over its r6 pieces, which is exact only if the open leaves of a row
form a prefix of the pieces whenever the pieces' bounds strictly increase.
Checked here on the CPU oracle's own verdicts (5 = empty after tightening)
for rows of the full proof's grid, all 64 r6 pieces each."""
import numpy as np

import oracle
from circlecover_amd import rsb_grid, subintervals


def test_open_leaves_of_a_row_are_a_prefix():
    rng = np.random.default_rng(7)
    la = subintervals(1.0, 2.5, 256)
    mixed = 0
    rows = 0
    for a in rng.choice(48, 16, replace=False):
        cells = rsb_grid(la[a][0], la[a][1], 1, 128, 128)
        for ci in rng.choice(cells.shape[0], 4, replace=False):
            c = cells[ci]
            r3, r4, r5, r6 = (subintervals(c[2 * k], c[2 * k + 1], 64) for k in (3, 4, 5, 6))
            for _ in range(4):
                i3, i4, i5 = rng.integers(0, 64, 3)
                leaves = np.repeat(c[None], 64, axis=0)
                leaves[:, 6:8], leaves[:, 8:10], leaves[:, 10:12] = r3[i3], r4[i4], r5[i5]
                leaves[:, 12:14] = r6
                _, v = oracle.rsb_count_ops(leaves)
                if (v <= 4).any():
                    continue  # a level-A shortcut proves the row: never enumerated
                # the kernel's condition: strictly increasing piece bounds
                assert (r6[1:, 0] > r6[:-1, 0]).all() and (r6[:, 1] > r6[:, 0]).all()
                is_open = v != 5
                k = int(is_open.sum())
                assert is_open[:k].all() and not is_open[k:].any(), (a, ci, i3, i4, i5)
                rows += 1
                mixed += 0 < k < 64
    assert rows > 50 and mixed > 5
