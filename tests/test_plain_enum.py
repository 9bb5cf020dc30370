"""The plain form of both proofs' enumerations (cc_set_enumeration_form):
the reference's kernel shape (Prover/tight_rectangle/tight_rectangle.cu:
169-277, Prover/rectangle_size_bound/rectangle_size_bound.cu:161-251), with
the reference's operation sequence (CC_ENUM_PLAIN) or the exact shortcuts
(CC_ENUM_PLAIN_EXACT).  Its leaves must be those of the CPU oracle, as a
set, and the drop-in executable must write the batched form's report byte
for byte."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from circlecover_amd import TR_LAMBDA_MAX, subintervals, tr_initial_cells
from circlecover_amd.native import CC_ENUM_BATCHED, CC_ENUM_PLAIN, CC_ENUM_PLAIN_EXACT

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "circlecover_amd",
                   "circlecover_prover")
FORMS = (CC_ENUM_PLAIN, CC_ENUM_PLAIN_EXACT)


def _with_form(prover, form, fn):
    prover.set_enumeration_form(form)
    try:
        return fn()
    finally:
        prover.set_enumeration_form(CC_ENUM_BATCHED)


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS)
def test_gpu_tr_plain_enumeration_matches_oracle(prover, form):
    cells = tr_initial_cells(1.0, 1.5, 64)[::7]
    counts = (32, 16, 16, 16)
    got = _with_form(prover, form, lambda: prover.tr_enumerate(cells, counts))
    want, trap = oracle.tr_find_critical(cells, counts)
    assert not trap and want.shape[0] > 0
    assert np.array_equal(oracle.canonical(got), oracle.canonical(want))


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS)
def test_gpu_rsb_plain_enumeration_matches_oracle(prover, form):
    outer = np.array([1.0, 2.5] + [0.0, 9.0 / 64.0] * 6)
    sub = outer.copy()
    sub[0:2] = subintervals(1.0, 2.5, 256)[40]
    cells = oracle.rsb_grid(sub, 1, 8, 8)[::5]
    counts = (16, 16, 16, 16)
    got, gtrap = _with_form(prover, form, lambda: prover.rsb_enumerate(cells, counts, with_trap=True))
    want = []
    wtrap = 0
    for c in cells:
        lv, t = oracle.rsb_find_critical(c, counts)
        want.extend(lv)
        wtrap |= t
    want = np.array(want).reshape(-1, 14)
    assert bool(gtrap) == bool(wtrap) and want.shape[0] > 0
    assert np.array_equal(oracle.canonical(got), oracle.canonical(want))


def _report(args):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


@pytest.mark.gpu
def test_gpu_plain_proofs_write_the_batched_reports():
    """tight_rectangle's first two lambda launches with one refinement
    generation, and a coarse rectangle_size_bound slice with generation 0:
    the same report in every enumeration form."""
    tr = ["--only", "tr", "--tr-la-range", "0,2", "--tr-generations", "1"]
    rsb = ["--only", "rsb", "--rsb-grid", "256,8,8", "--rsb-la-range", "40,1", "--rsb-generations", "0"]
    for args in (tr, rsb):
        want = _report(args)
        assert want.count("\n") > 2
        for form in ("plain", "plain-exact"):
            assert _report(args + ["--enum-form", form]) == want, (args, form)


def test_enumeration_form_argument_is_checked():
    r = subprocess.run([EXE, "--only", "tr", "--enum-form", "fast"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--enum-form" in r.stderr
