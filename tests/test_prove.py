"""The drop-in entry points end to end: the circlecover_prover executable
(the reference's main, Prover/prover.cpp:33-37) on shortened schedules of
the same procedure, its stdout compared byte for byte with the oracle's
enumeration, the reference's merges and report format restated in Python."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from circlecover_amd import TR_LAMBDA_MAX, tr_initial_cells
from circlecover_amd.report import compare_reports, read_report
from report_format import begin_case, end_case, iv, merge_critical, rsb_report, tr_report

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "circlecover_amd",
                   "circlecover_prover")


def test_report_format_helpers():
    b = begin_case("Critical intervals for rectangles (tight)")
    assert len(b) == 121 and b.startswith("-" * 38 + " Critical") and b.endswith("-\n")
    assert end_case() == "-" * 120 + "\n"
    # the same text as C++ iostream with precision 19, width 26, right-aligned
    assert iv(1.0, 2.5) == "[" + " " * 25 + "1, " + " " * 23 + "2.5]"
    assert iv(0.1, 1 / 3) == "[     0.1000000000000000056,      0.3333333333333333148]"


def test_merge_critical_joins_equal_prefix():
    a = np.array([[1, 2] + [0, 1] * 6, [1, 2] + [0, 1] * 5 + [0, 3], [1, 3] + [0, 1] * 6], dtype=float)
    m0 = merge_critical(a, 0)
    assert m0.shape[0] == 2 and m0[0, 13] == 3
    m6 = merge_critical(a, 6)
    assert m6.shape[0] == 3


def run_prover(args):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


@pytest.mark.gpu
def test_gpu_tr_prove_report_matches_oracle():
    cells = tr_initial_cells(1.0, TR_LAMBDA_MAX, 256)[0:2]
    crit = []
    for c in cells:
        leaves, trap = oracle.tr_find_critical(c[None], (128, 128, 128, 128))
        assert not trap
        crit.extend(oracle.tr_merge_adjacent_r1(leaves))
    assert len(crit) > 0
    expected = tr_report(np.array(crit))
    got = run_prover(["--only", "tr", "--tr-la-range", "0,2", "--tr-generations", "0"])
    assert got == expected
    assert compare_reports(got, expected) == [] and read_report(got)[0].lines


@pytest.mark.gpu
def test_gpu_rsb_prove_report_matches_oracle():
    from circlecover_amd import subintervals

    outer = np.array([1.0, 2.5] + [0.0, 9.0 / 64.0] * 6)
    sub = outer.copy()
    sub[0:2] = subintervals(1.0, 2.5, 256)[40]
    leaves = []
    for c in oracle.rsb_grid(sub, 1, 8, 8):
        lv, trap = oracle.rsb_find_critical(c, (64, 64, 64, 64))
        assert not trap
        leaves.extend(lv)
    assert len(leaves) > 0
    expected = rsb_report(merge_critical(np.array(leaves), 0))
    got = run_prover(["--only", "rsb", "--rsb-grid", "256,8,8", "--rsb-la-range", "40,1", "--rsb-generations", "0"])
    assert got == expected
    assert compare_reports(got, expected) == [] and read_report(got)[0].lines


@pytest.mark.gpu
def test_gpu_rsb_prove_trap_fails_the_run():
    """On this coarse grid some leaf of lambda piece 0 reaches a trap site
    (the oracle agrees); the reference aborts there, the prover exits with an
    error instead of printing a report."""
    r = subprocess.run([EXE, "--only", "rsb", "--rsb-grid", "256,8,8", "--rsb-la-range", "0,1",
                        "--rsb-generations", "0"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 1 and "error 3" in r.stderr


def _run_stop(args):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout


@pytest.mark.gpu
def test_gpu_tr_prove_resumes_from_checkpoint(tmp_path):
    """A proof stopped after a batch, resumed, stopped again inside a
    generation and resumed to the end writes the uninterrupted run's report;
    the checkpoint file is gone once the proof completes."""
    args = ["--only", "tr", "--tr-la-range", "0,40", "--tr-generations", "2"]
    full = run_prover(args)
    ck = str(tmp_path / "ck")
    code, out = _run_stop([*args, "--checkpoint", ck, "--max-batches", "1"])
    assert code == 3 and out == "" and os.path.exists(ck + ".tr")
    code, out = _run_stop([*args, "--checkpoint", ck, "--max-batches", "3"])
    assert code == 3 and out == "" and os.path.exists(ck + ".tr")
    assert run_prover([*args, "--checkpoint", ck]) == full
    assert not os.path.exists(ck + ".tr")


@pytest.mark.gpu
def test_gpu_rsb_prove_resumes_from_checkpoint(tmp_path):
    args = ["--only", "rsb", "--rsb-grid", "256,8,8", "--rsb-la-range", "40,70", "--rsb-generations", "0"]
    full = run_prover(args)
    ck = str(tmp_path / "ck")
    code, out = _run_stop([*args, "--checkpoint", ck, "--max-batches", "1"])
    assert code == 3 and out == "" and os.path.exists(ck + ".rsb")
    assert run_prover([*args, "--checkpoint", ck]) == full
    assert not os.path.exists(ck + ".rsb")
    # a checkpoint of other parameters is refused
    _run_stop([*args, "--checkpoint", ck, "--max-batches", "1"])
    other = ["--only", "rsb", "--rsb-grid", "256,8,8", "--rsb-la-range", "40,71", "--rsb-generations", "0"]
    r = subprocess.run([EXE, *other, "--checkpoint", ck], capture_output=True, text=True, timeout=600)
    assert r.returncode == 1 and "checkpoint" in r.stderr
