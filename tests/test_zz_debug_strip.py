"""Diagnostics: where a cover-strip bisection on the GPU departs from the CPU
oracle (reports the first differing iteration of each differing box)."""
import numpy as np
import pytest

import oracle
from circlecover_amd import gen_c2_boxes


@pytest.mark.gpu
def test_cover_strip_traces_agree(prover):
    soa = gen_c2_boxes(0xC1C1EC0000000002, 1 << 14)
    gm, _ = prover.eval_batch(soa)
    om, _ = oracle.c2_eval(soa)
    bad = np.flatnonzero(((gm ^ om) >> 4) & 1)
    report = []
    for i in bad[:6]:
        box = soa[:, i].copy()
        gt, gv = prover.debug_trace_strip(box, 256)
        ot, ov = oracle.trace_strip(box, 256)
        diff = np.flatnonzero((gt.view(np.uint64) != ot.view(np.uint64)).any(axis=1))
        j = int(diff[0]) if diff.size else -1
        report.append(f"box {i} {[x.hex() for x in box]} gpu={gv} cpu={ov} first diff iter {j}: "
                      f"gpu {[x.hex() for x in gt[j]] if j >= 0 else None} cpu {[x.hex() for x in ot[j]] if j >= 0 else None}")
    assert bad.size == 0, "\n".join(report)
