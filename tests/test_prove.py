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
from report_format import begin_case, end_case, iv, max_distances, merge_critical, rsb_report, tr_report

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


def check_tr_run(args, criticals):
    """A tight_rectangle run whose final critical set is `criticals` writes
    the oracle's report; where a three-disk cell of the set takes the square
    root of a possibly negative interval in the Max_distances step, the
    reference's kernel traps there (ACU/interval.cuh:339-350), and the run
    fails with the trap error and no report instead.  Returns whether it
    trapped."""
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=600)
    if len(criticals) and max_distances(criticals).trap:
        assert r.returncode == 1 and r.stdout == "" and "error 3" in r.stderr, r.stderr[-2000:]
        return True
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout == tr_report(criticals)
    return False


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


@pytest.mark.gpu
def test_gpu_tr_prove_reference_entry_point_completes():
    """The executable without flags runs the reference's main through the C++
    entry point circlecover::tight_rectangle::find_critical_intervals() with
    the reference's constants (256 initial launches of 128^4 pieces, eleven
    generations at most); the proof completes with an empty critical set, so
    the report is the banner and the closing rule only."""
    r = subprocess.run([EXE, "--only", "tr"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout == begin_case("Critical intervals for rectangles (tight)") + end_case()
    assert "Running generation 1" in r.stderr


@pytest.mark.gpu
def test_gpu_tr_prove_sigint_reports_last_complete_generation():
    """SIGINT during the refinement ends the run with the report of the last
    complete generation and exit status 0, as the reference's
    was_interrupted check does (Prover/interrupt.cpp:30-46,
    TR/tight_rectangle.cu:481-489).  The signal is sent once stderr shows
    generation 2 running; the generation it interrupts is the last one the
    run announced, so the expected report is that of one generation fewer."""
    import re
    import signal

    args = ["--only", "tr", "--tr-la-range", "0,40", "--tr-generations", "3"]
    p = subprocess.Popen([EXE, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    err = []
    try:
        for line in p.stderr:
            err.append(line)
            if "Running generation 2" in line:
                p.send_signal(signal.SIGINT)
                break
        rest_out, rest_err = p.communicate(timeout=600)
    finally:
        if p.poll() is None:
            p.kill()
            p.wait()
    err.append(rest_err)
    assert p.returncode == 0, "".join(err)[-2000:]
    m = re.search(r"Interrupted during generation (\d+): reporting generation (\d+)", "".join(err))
    # (a run that ended before the signal arrived reports all its generations)
    reported = int(m.group(2)) if m else 3
    if m:
        assert int(m.group(1)) >= 2 and reported == int(m.group(1)) - 1
    expected = run_prover(["--only", "tr", "--tr-la-range", "0,40", "--tr-generations", str(reported)])
    assert rest_out == expected


SCHEDULE = (1, 2, 2, 2, 2, 1, 2, 1, 2, 1, 2)  # tight_rectangle.cu:50-53


def _tr_oracle_generations(first, count, init_counts, gen_counts, generations):
    """The critical records after the initial launches [first, first+count)
    of 256 and after each generation, by the oracle with the reference's
    driver steps: launches of get_subinterval pieces of lambda by the
    schedule, each launch's leaves merged by merge_adjacent_r1
    (tight_rectangle.cu:129-159, 476-496)."""
    cur = []
    for c in tr_initial_cells(1.0, TR_LAMBDA_MAX, 256)[first:first + count]:
        leaves, trap = oracle.tr_find_critical(c[None], init_counts)
        assert not trap
        cur.extend(oracle.tr_merge_adjacent_r1(leaves))
    sets = [np.array(cur)]
    for g in range(1, generations + 1):
        nxt = []
        for s in cur:
            for piece in oracle.subintervals(s[0], s[1], SCHEDULE[g - 1]):
                c = np.array(s, dtype=np.float64)
                c[0:2] = piece
                leaves, trap = oracle.tr_find_critical(c[None], gen_counts)
                assert not trap
                nxt.extend(oracle.tr_merge_adjacent_r1(leaves))
        cur = nxt
        sets.append(np.array(cur))
        if not cur:
            break
    return sets


def _rsb_oracle_generations(piece, nr1, nr2, init_counts, gen_grid, gen_counts, generations):
    """The critical records of the rectangle_size_bound proof on lambda piece
    `piece` of 256 with an nr1 x nr2 initial grid, after merge<0> and after
    each generation (host grid gen_grid per critical record, merge<2>,
    merge<1>, merge<3> by g mod 3; rectangle_size_bound.cu:103-130, 336-371)."""
    from circlecover_amd import subintervals

    outer = np.array([1.0, 2.5] + [0.0, 9.0 / 64.0] * 6)
    sub = outer.copy()
    sub[0:2] = subintervals(1.0, 2.5, 256)[piece]
    leaves = []
    for c in oracle.rsb_grid(sub, 1, nr1, nr2):
        lv, trap = oracle.rsb_find_critical(c, init_counts)
        assert not trap
        leaves.extend(lv)
    cur = merge_critical(np.array(leaves), 0)
    sets = [cur]
    for g in range(1, generations + 1):
        nxt = []
        for c in cur:
            for cell in oracle.rsb_grid(c, *gen_grid):
                lv, trap = oracle.rsb_find_critical(cell, gen_counts)
                assert not trap
                nxt.extend(lv)
        cur = merge_critical(np.array(nxt), 3 if g % 3 == 0 else 2 if g % 3 == 1 else 1)
        sets.append(cur)
        if not len(cur):
            break
    return sets


@pytest.mark.gpu
def test_gpu_tr_prove_generations_match_oracle():
    """Generations 0..3 of the tight_rectangle proof on 8 initial lambda
    launches, with r1..r4 piece counts 32^4 for the initial launches and
    16,16,16,8 for the generation launches (the critical sets have 12, 8 and
    32 records, then none): each report equals the oracle's, byte for byte.
    The initial launches' 12 wide cells include a three-disk cell with r1
    possibly below 1/4, so that run ends at the reference's trap."""
    sets = _tr_oracle_generations(0, 8, (32, 32, 32, 32), (16, 16, 16, 8), 3)
    assert [len(s) for s in sets] == [12, 8, 32, 0]
    base = ["--only", "tr", "--tr-la-range", "0,8", "--tr-counts", "32,32,32,32", "--tr-gen-counts", "16,16,16,8"]
    trapped = []
    for g, s in enumerate(sets):
        print(f"tr generations {g}", flush=True)
        trapped.append(check_tr_run([*base, "--tr-generations", str(g)], s))
    assert trapped == [True, False, False, False]


@pytest.mark.gpu
def test_gpu_tr_prove_generations_at_reference_counts():
    """Generations 0..3 of the tight_rectangle proof at the reference's piece
    counts (128^4, then 128,128,128,64) on the first two initial lambda
    launches, against the oracle's records in tests/golden/tr_generations.json
    (the oracle needs minutes per generation launch at these counts;
    tests/golden/make_tr_generations.py writes the file)."""
    import json

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tr_generations.json")
    if not os.path.exists(path):
        pytest.skip("tests/golden/tr_generations.json not generated (tests/golden/make_tr_generations.py)")
    gold = json.load(open(path))
    first, count = gold["lambda_launches"]
    for g, recs in enumerate(gold["generations"]):
        s = np.array([[float.fromhex(x) for x in r] for r in recs]).reshape(-1, 10)
        print(f"tr generations {g} at reference counts", flush=True)
        assert not check_tr_run(["--only", "tr", "--tr-la-range", f"{first},{count}", "--tr-generations", str(g)], s)


@pytest.mark.gpu
def test_gpu_rsb_prove_generations_match_oracle():
    """Generations 0..3 of the rectangle_size_bound proof on lambda piece 80
    of 256 with a 4x4 initial grid and 16^4 pieces, a 2x2x2 generation grid
    and 2,2,2,4 generation pieces (critical sets of 1, 10, 8 and 14
    records): the merges <0>, <2>, <1>, <3> of the reference's generations
    all run, and each report equals the oracle's byte for byte."""
    sets = _rsb_oracle_generations(80, 4, 4, (16, 16, 16, 16), (2, 2, 2), (2, 2, 2, 4), 3)
    assert [len(s) for s in sets] == [1, 10, 8, 14]
    base = ["--only", "rsb", "--rsb-grid", "256,4,4", "--rsb-la-range", "80,1", "--rsb-counts", "16,16,16,16",
            "--rsb-gen-grid", "2,2,2", "--rsb-gen-counts", "2,2,2,4"]
    for g, s in enumerate(sets):
        print(f"rsb generations {g}", flush=True)
        assert run_prover([*base, "--rsb-generations", str(g)]) == rsb_report(s)
    # a generation that leaves nothing: 4x4x4 grid, 4,4,4,8 pieces
    sets = _rsb_oracle_generations(80, 4, 4, (16, 16, 16, 16), (4, 4, 4), (4, 4, 4, 8), 3)
    assert [len(s) for s in sets] == [1, 16, 0]
    base[-3:] = ["4,4,4", "--rsb-gen-counts", "4,4,4,8"]
    assert run_prover([*base, "--rsb-generations", "3"]) == rsb_report(sets[-1])


@pytest.mark.gpu
def test_gpu_rsb_prove_resumes_inside_a_generation(tmp_path):
    """Stopped after every batch (the initial cells, then each critical cell
    of generations 1 and 2) and resumed each time, the rectangle_size_bound
    proof writes the uninterrupted run's report."""
    args = ["--only", "rsb", "--rsb-grid", "256,4,4", "--rsb-la-range", "80,1", "--rsb-counts", "16,16,16,16",
            "--rsb-gen-grid", "2,2,2", "--rsb-gen-counts", "2,2,2,4", "--rsb-generations", "2"]
    full = run_prover(args)
    ck = str(tmp_path / "ck")
    stops = 0
    while True:
        code, out = _run_stop([*args, "--checkpoint", ck, "--max-batches", "1"])
        if code == 0:
            break
        assert code == 3 and out == "" and os.path.exists(ck + ".rsb")
        stops += 1
        assert stops < 64
    assert stops >= 3 and out == full and not os.path.exists(ck + ".rsb")


@pytest.mark.gpu
def test_gpu_prove_refuses_a_corrupt_checkpoint(tmp_path):
    """A checkpoint whose record counts do not match its length (truncated,
    extended, or a count of 2^39 records) is refused with an error before
    anything is sized from it."""
    args = ["--only", "rsb", "--rsb-grid", "256,8,8", "--rsb-la-range", "40,70", "--rsb-generations", "0"]
    ck = str(tmp_path / "ck")
    code, _ = _run_stop([*args, "--checkpoint", ck, "--max-batches", "1"])
    assert code == 3
    data = open(ck + ".rsb", "rb").read()
    for bad in (data[:-5], data + b"\0" * 112, data[:128] + (1 << 39).to_bytes(8, "little") + data[136:]):
        open(ck + ".rsb", "wb").write(bad)
        r = subprocess.run([EXE, *args, "--checkpoint", ck], capture_output=True, text=True, timeout=600)
        assert r.returncode == 1 and "checkpoint" in r.stderr
