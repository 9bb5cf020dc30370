#!/usr/bin/env python3
"""Benchmark: interval cell-evaluations per second on one MI355X.

Workload (batched criterion throughput): 2^22 seeded synthetic 4-D boxes
(lambda, r1, r2, r3), FP64 outward-rounded intervals, ten covering criteria
per box evaluated without short-circuit; one step = one pass over all boxes
= 41,943,040 cell-evals (one cell-eval = one (box, criterion) predicate).
Boxes are generated on the host and uploaded before the timed region.

Prints ONE JSON line.  `roofline.achieved` counts the algorithmic directed
FP64 operations (reference op sequence, counted by the CPU oracle's counting
build on a seeded sample and scaled to the batch) per batch, divided by the
batch's duration measured with HIP events on the launch stream.  A batch is
ccrd_c2_eval plus its two small companions (the counter reset and the
second pass over the boxes whose cover-strip search it deferred).
`cpu_baseline` times the CPU oracle on the host cores on the same boxes.

Unless --core-only is given, the line also carries `workloads`: the other
configurations measured in the same run (not part of `value`):
  rsb_bnb      branch-and-bound over the 256x64x64 rectangle_size_bound grid,
               lambda in [1, 8], depth 24: wall time of the level-synchronous
               form (reference operation sequence, and with the batched
               form's exact shortcuts), the batched and the learned+batched
               forms; identical unproven set required;
  tr_small     tight_rectangle: 64 lambda cells of [1, 1.5], r1..r4 in
               64/32/32/32 pieces, three bisection levels;
  fp32_prefilter  the FP32 pre-filter on the 2^22 boxes (first-success order,
               per-box and cascade forms, each against its FP64 form), each
               criterion's FP32 shadow against its FP64 form (proofs decided
               in FP32, FP32-accept with FP64-reject count, weight-budget
               width inflation median/p99) and the ten-criterion batch with
               the shadows first;
  ranker       the learned criterion ranker trained on 2^20 boxes of another
               seed; first-success evaluation of the 2^22 boxes in plain order,
               plain cascade, ranked and ranked cascade (same proven set);
  full_proof   both of the reference's proofs end to end through the drop-in
               executable (`circlecover_prover --only tr|rsb`), wall time and
               number of critical intervals left; then tight_rectangle in full
               and rectangle_size_bound on lambda pieces 24..27 in the three
               enumeration forms (batched, plain = the reference's kernel
               shape and operation sequence, plain-exact), same reports.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED = 0xC1C1EC0000000002
N_BOXES = 1 << 22
N_CRIT = 10
FP64_PEAK_TFLOPS = 78.6
METRIC = "full-proof wall time & interval cell-evals/s: learned+batched vs plain GPU"


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cpu() -> dict:
    """The host as lscpu describes it, and the CPUs this process may use."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass

    def num(key):
        try:
            return int(info[key])
        except (KeyError, ValueError):
            return None

    return {"model": info.get("Model name", cpu_model()), "sockets": num("Socket(s)"),
            "cores_per_socket": num("Core(s) per socket"), "threads_per_core": num("Thread(s) per core"),
            "cpus_online": num("CPU(s)"), "cpus_in_affinity": len(os.sched_getaffinity(0))}


C3_SAMPLE = 32768  # cells of the 256x64x64 grid in the CPU leg of the branch-and-bound


def c3_sample_cells():
    import circlecover_amd as cc

    g = cc.rsb_grid(1.0, 8.0, 256, 64, 64)
    return g[np.sort(np.random.default_rng(3).choice(g.shape[0], C3_SAMPLE, replace=False))]


def cpu_baseline(soa: np.ndarray, threads: int) -> dict:
    """The CPU oracle on this host's cores, on bounded samples of three
    workloads: the ten-criterion batch (all boxes on `threads` threads, and
    2^16 boxes on one thread), the small tight_rectangle branch-and-bound
    (all of it) and the rectangle_size_bound branch-and-bound on a fixed
    sample of its grid's cells."""
    import circlecover_amd as cc
    import oracle

    n = soa.shape[1]
    t0 = time.perf_counter()
    oracle.c2_eval(soa, threads=threads)
    dt = time.perf_counter() - t0
    one = soa[:, : 1 << 16].copy()
    t0 = time.perf_counter()
    oracle.c2_eval(one, threads=1)
    dt1 = time.perf_counter() - t0
    # tight_rectangle small branch-and-bound: 64 lambda cells, 64/32/32/32
    # pieces, three bisection levels
    t0 = time.perf_counter()
    leaves, _ = oracle.tr_find_critical(cc.tr_initial_cells(1.0, 1.5, 64), (64, 32, 32, 32), threads=threads)
    for _ in range(3):
        leaves, _ = oracle.tr_find_critical(oracle.tr_bisect(leaves), (1, 1, 1, 1), threads=threads)
    c1_s = time.perf_counter() - t0
    cells = c3_sample_cells()
    t0 = time.perf_counter()
    unproven, per, _, _ = oracle.rsb_bnb(cells, 24, threads=threads)
    c3_s = time.perf_counter() - t0
    return {"value": n * N_CRIT / dt, "unit": "cell-evals/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of the {N_BOXES} seeded boxes, all ten criteria, CPU oracle "
                      f"(x86 MXCSR directed rounding), {threads} threads, {dt:.2f} s",
            "single_thread_value": one.shape[1] * N_CRIT / dt1,
            "single_thread_sample": f"first {one.shape[1]} boxes, 1 thread, {dt1:.2f} s",
            "c1_s": c1_s, "c1_unproven": int(leaves.shape[0]),
            "c1_sample": "the whole small tight_rectangle branch-and-bound (64 lambda cells of [1, 1.5], "
                         f"64/32/32/32 pieces, 3 bisection levels), {threads} threads",
            "c3_sample_s": c3_s, "c3_sample_cells_evaluated": int(per.sum()),
            "c3_sample_unproven": int(unproven.shape[0]),
            "c3_sample": f"{C3_SAMPLE} seeded cells of the 256x64x64 rectangle_size_bound grid, depth 24, "
                         f"{threads} threads",
            "host": host_cpu(), "cpu_model": cpu_model()}


def ops_per_box(soa: np.ndarray) -> float:
    import oracle

    ops = oracle.c2_count_ops(soa)
    return float(ops.sum()) / soa.shape[1]


COPY_BYTES = 256 << 20  # calibration copy: 256 MiB read + 256 MiB written


def traffic_probe() -> int:
    """One calibration copy and one batch launch, for the HBM counter runs."""
    from circlecover_amd import Prover, gen_c2_boxes

    n = N_BOXES
    soa = gen_c2_boxes(SEED, n)
    with Prover(0) as p:
        a, b = p.alloc(COPY_BYTES), p.alloc(COPY_BYTES)
        d_soa, d_mask, d_first = p.alloc(soa.nbytes), p.alloc(4 * n), p.alloc(n)
        d_soa.upload(soa)
        p.copy_async(b, a, COPY_BYTES)
        p.eval_batch_async(d_soa, n, n, d_mask, d_first)
        p.synchronize()
        for x in (a, b, d_soa, d_mask, d_first):
            x.free()
    return 0


def measured_traffic() -> dict | None:
    """HBM bytes of one batch launch from two rocprofv3 counter runs
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by the
    calibration copy's reading (expected COPY_BYTES read / written).  None
    when rocprofv3 is unavailable or a run fails."""
    import csv
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        with tempfile.TemporaryDirectory(dir="/tmp") as d:
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "t", "--",
                   sys.executable, os.path.abspath(__file__), "--traffic-probe"]
            try:
                subprocess.run(cmd, cwd="/tmp", capture_output=True, timeout=120, check=True,
                               env=dict(os.environ, TMPDIR="/tmp"))
            except (subprocess.SubprocessError, OSError):
                return None
            rows = []
            for root, _, files in os.walk(d):
                for f in files:
                    if f.endswith("counter_collection.csv"):
                        rows += list(csv.DictReader(open(os.path.join(root, f))))
            vals = {}
            for r in rows:
                name = r.get("Kernel_Name", "")
                batch = ("ccrd_c2_eval", "ccrd_c2_strip_deferred", "ccrd_c2_r1_pass", "cc_zero_u64")  # one batch's launches
                key = "copy" if "cc_copy_u4" in name else "c2" if any(k in name for k in batch) else None
                if key and r.get("Counter_Name", "") == counter:
                    vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"])
            if "copy" not in vals or "c2" not in vals or vals["copy"] <= 0:
                return None
            scale = COPY_BYTES / vals["copy"]
            got[counter] = {"kib": vals["c2"], "copy_kib": vals["copy"], "scale": scale, "bytes": vals["c2"] * scale}
    total = got["FETCH_SIZE"]["bytes"] + got["WRITE_SIZE"]["bytes"]
    return {"bytes": total, "fetch": got["FETCH_SIZE"], "write": got["WRITE_SIZE"]}


def extended(p, soa: np.ndarray, mask: np.ndarray) -> dict:
    """The other configurations, each timed once after a warm-up run."""
    import circlecover_amd as cc

    out = {}
    # rectangle_size_bound branch-and-bound: plain, batched, learned+batched
    out["rsb_bnb"] = c3_forms(p)
    # the CPU leg's sample of the same grid, on the GPU (batched form)
    sample = c3_sample_cells()
    p.rsb_bnb_batched(sample, 24)
    t = time.perf_counter()
    us, ss = p.rsb_bnb_batched(sample, 24)
    out["rsb_bnb"]["sample"] = {"cells": int(sample.shape[0]), "batched_s": time.perf_counter() - t,
                                "unproven": int(us.shape[0]),
                                "cells_evaluated": int(sum(ss.cells_per_depth[: ss.depth_reached + 1]))}
    # tight_rectangle small branch-and-bound
    tc = cc.tr_initial_cells(1.0, 1.5, 64)
    p.tr_bnb(tc, (64, 32, 32, 32), 3)
    t = time.perf_counter()
    lv, per = p.tr_bnb(tc, (64, 32, 32, 32), 3)
    out["tr_small"] = {"wall_s": time.perf_counter() - t, "unproven": int(lv.shape[0]),
                       "unproven_per_level": [int(x) for x in per]}
    # FP32 pre-filter vs FP64 first-success, per-box and cascade forms
    p.eval_first(soa)
    f64_first, _, _, st64 = p.eval_first(soa)
    p.eval_first(soa, cascade=True)
    _, _, _, st64c = p.eval_first(soa, cascade=True)
    _, _, _, _, w, _ = p.eval_first_f32(soa, with_widths=True)  # widths for the inflation report, untimed
    first, evals, trap, dec, _, st32 = p.eval_first_f32(soa)
    p.eval_first_f32(soa, cascade=True)
    first_c, _, _, dec_c, _, st32c = p.eval_first_f32(soa, cascade=True)
    bad = int(((dec == 1) & ((mask & 1) == 0)).sum())
    r_cr = w[:, 0] / np.maximum(w[:, 1], 1e-300)
    r_rem = w[:, 2] / np.maximum(w[:, 3], 1e-300)
    out["fp32_prefilter"] = {"fraction_decided_fp32": float(dec.mean()), "fp32_accept_fp64_reject": bad,
                             "identical_first": bool((first == f64_first).all() and (first_c == f64_first).all()
                                                     and (dec_c == dec).all()),
                             "eval_ms_fp32_then_fp64": st32.eval_ms, "eval_ms_fp64": st64.eval_ms,
                             "eval_ms_fp32_then_fp64_cascade": st32c.eval_ms, "eval_ms_fp64_cascade": st64c.eval_ms,
                             "width_inflation": {
                                 "r1_strategies.critical_ratio": [float(np.median(r_cr)),
                                                                  float(np.percentile(r_cr, 99))],
                                 "r1_strategies.remaining_weight": [float(np.median(r_rem)),
                                                                    float(np.percentile(r_rem, 99))]}}
    # every criterion's FP32 shadow against its FP64 form on the 2^22 boxes,
    # and the ten-criterion batch with the shadows first (CC_PREC_FP32_PRE)
    m32, infl = p.f32_report(soa)
    per = {}
    for k in range(N_CRIT):
        p64 = ((mask >> k) & 1).astype(bool)
        if k == 4:
            per[str(k)] = {"shadow": None, "fp64_proves": int(p64.sum())}
            continue
        p32 = ((m32 >> k) & 1).astype(bool)
        r = infl[:, k][infl[:, k] >= 0]
        per[str(k)] = {"fp64_proves": int(p64.sum()), "fp32_proves": int(p32.sum()),
                       "fraction_of_fp64_proofs_decided_fp32": float(p32.sum() / max(1, p64.sum())),
                       "fp32_accept_fp64_reject": int((p32 & ~p64).sum()),
                       "budget_width_inflation_median": float(np.median(r)) if r.size else None,
                       "budget_width_inflation_p99": float(np.percentile(r, 99)) if r.size else None}
    del infl
    n = soa.shape[1]
    d_soa, d_mask, d_first = p.alloc(soa.nbytes), p.alloc(4 * n), p.alloc(n)
    d_soa.upload(soa)
    batch_ms = {}
    for name, prec in (("fp64", 0), ("fp32_pre", 1)):
        p.eval_batch_async(d_soa, n, n, d_mask, d_first, precision=prec)
        p.synchronize()
        p.timer_start()
        for _ in range(5):
            p.eval_batch_async(d_soa, n, n, d_mask, d_first, precision=prec)
        batch_ms[name] = p.timer_stop() / 5
    pre_mask = d_mask.download(np.empty(n, dtype=np.uint32))
    for b in (d_soa, d_mask, d_first):
        b.free()
    out["fp32_prefilter"]["per_criterion"] = per
    # the shortcuts S1..S5 (criteria 5..9, on a cell's lambda, r1, r2, r3) on
    # the cells the rectangle_size_bound branch-and-bound evaluates
    # (64x32x32 grid over lambda in [1, 8], depth 24, every depth logged)
    _, _, rec = p.rsb_bnb_log(cc.rsb_grid(1.0, 8.0, 64, 32, 32), 24)
    cs = rec["cell"]
    csoa = np.ascontiguousarray(cs[:, :8].T)
    c64, _ = p.eval_batch(csoa)
    c32, _ = p.f32_report(csoa, with_inflation=False)
    s15 = {}
    for k in range(5, 10):
        a64 = ((c64 >> k) & 1).astype(bool)
        a32 = ((c32 >> k) & 1).astype(bool)
        s15[f"S{k - 4}"] = {"fp64_proves": int(a64.sum()), "fp32_proves": int(a32.sum()),
                            "fraction_of_fp64_proofs_decided_fp32": float(a32.sum() / max(1, a64.sum())),
                            "fp32_accept_fp64_reject": int((a32 & ~a64).sum())}
    out["fp32_prefilter"]["rsb_bnb_cells"] = {"cells": int(cs.shape[0]), "shortcuts": s15}
    out["fp32_prefilter"]["batch"] = {"ms_fp64": batch_ms["fp64"], "ms_fp32_pre": batch_ms["fp32_pre"],
                                      "identical_masks": bool((pre_mask == mask).all()),
                                      "fp32_accept_fp64_reject": int((m32 & ~(mask & 0x3FF)).astype(bool).sum())}
    # learned ranker: first-success evaluation of the 2^22 boxes in plain
    # order, plain cascade (criterion 0 on every box, the rest on the boxes
    # it leaves open), ranked, and ranked cascade (the ranker trained on, and
    # applied to, only the boxes criterion 0 leaves open)
    train = cc.gen_c2_boxes(SEED + 1, 1 << 20)
    tf, _, _, _ = p.eval_first(train)
    # labels: first proving criterion in plain order, 10 where none proves
    tl = np.where(tf == 255, 10, tf).astype(np.uint8)
    t = time.perf_counter()
    model, losses = p.ranker_train(train, tl, epochs=20, batch=4096, seed=1)
    t_train = time.perf_counter() - t
    tf_c = np.where(tl == 0, 255, tl).astype(np.uint8)
    t = time.perf_counter()
    model_c, losses_c = p.ranker_train(train, tf_c, epochs=20, batch=4096, seed=1)
    t_train_c = time.perf_counter() - t
    variants = {}
    proven_sets = []
    for name, m, casc in (("plain", None, False), ("plain_cascade", None, True), ("ranked", model, False),
                          ("ranked_cascade", model_c, True)):
        p.eval_first(soa, m, cascade=casc)
        t = time.perf_counter()
        f, e, _, st = p.eval_first(soa, m, cascade=casc)
        wall = time.perf_counter() - t
        proven_sets.append(f < 10)
        variants[name] = {"evals_per_proven": st.evals_proven / st.proven, "device_ms": st.rank_ms + st.eval_ms,
                          "rank_ms": st.rank_ms, "eval_ms": st.eval_ms, "second_stage_boxes": int(st.second_stage),
                          "wall_s": wall}
    out["ranker"] = {"train_s": t_train, "final_loss": float(losses[-1]), "train_boxes": int((tl <= 10).sum()),
                     "train_s_cascade": t_train_c, "final_loss_cascade": float(losses_c[-1]),
                     "train_boxes_cascade": int((tf_c <= 10).sum()), "variants": variants,
                     "evals_per_proven_plain": variants["plain"]["evals_per_proven"],
                     "evals_per_proven_ranked": variants["ranked"]["evals_per_proven"],
                     "fastest": min(variants, key=lambda k: variants[k]["device_ms"]),
                     "same_proven_set": bool(all((x == proven_sets[0]).all() for x in proven_sets))}
    return out


C3_FORMS = ("plain", "plain_exact", "batched", "learned_batched")


C3_REPEATS = 3


def c3_forms(p, forms=C3_FORMS) -> dict:
    """The rectangle_size_bound branch-and-bound on the 256x64x64 grid over
    lambda in [1, 8], depth 24, in its forms, each timed three times after a
    warm-up run, the median reported (wall time of the call, cells uploaded
    from the host; the runs of one form differ by a few percent):
    level-synchronous with the reference's operation sequence (plain), the
    same with the batched form's exact shortcuts (plain_exact), batched, and
    learned+batched.  The ranker of the learned form is trained on the GPU
    on the records the prover logs on another grid of the same domain
    (128x48x48)."""
    import circlecover_amd as cc

    cells = cc.rsb_grid(1.0, 8.0, 256, 64, 64)
    out = {"cells": int(cells.shape[0]), "max_depth": 24}
    model = None
    if "learned_batched" in forms:
        log("C3: ranker training data (level-synchronous run on a 128x48x48 grid, logged)")
        t = time.perf_counter()
        _, _, rec = p.rsb_bnb_log(cc.rsb_grid(1.0, 8.0, 128, 48, 48), 24)
        t_log = time.perf_counter() - t
        t = time.perf_counter()
        model, losses = p.rsb_ranker_train(rec, epochs=20, batch=4096, seed=1)
        out.update({"train_records": int(len(rec)), "train_depths": int(rec["depth"].max()) + 1, "log_s": t_log,
                    "train_s": time.perf_counter() - t, "final_loss": float(losses[-1]),
                    "class_step": list(model.class_step)})
    calls = {"batched": lambda: p.rsb_bnb_batched(cells, 24),
             "learned_batched": lambda: p.rsb_bnb_learned(cells, 24, model),
             "plain": lambda: p.rsb_bnb(cells, 24),
             "plain_exact": lambda: p.rsb_bnb(cells, 24, exact=True)}
    runs = {}
    for name in ("batched", "learned_batched", "plain", "plain_exact"):
        if name not in forms:
            continue
        log(f"C3: {name} form")
        if name in ("batched", "learned_batched"):
            calls[name]()  # warm-up: graph capture
        reps = []
        for _ in range(C3_REPEATS):
            t = time.perf_counter()
            u, st = calls[name]()
            reps.append((time.perf_counter() - t, u, st))
        reps.sort(key=lambda r: r[0])
        runs[name] = reps[len(reps) // 2]
        out.setdefault("runs_s", {})[name] = [r[0] for r in reps]
    out.update({f"{k}_s": v[0] for k, v in runs.items()})
    out["device_ms"] = {k: v[2].device_ms for k, v in runs.items()}
    for base in ("plain", "plain_exact"):
        for fast in ("batched", "learned_batched"):
            if base in runs and fast in runs:
                out[f"speedup_{fast}" + ("" if base == "plain" else "_vs_plain_exact")] = runs[base][0] / runs[fast][0]
    ref = runs.get("plain") or next(iter(runs.values()))
    _, u0, s0 = ref
    d = s0.depth_reached
    canon = lambda u: u[np.lexsort(u.T[::-1])] if len(u) else u
    out["unproven"] = int(u0.shape[0])
    # every form: the same unproven cells and cells per depth; the forms in
    # the reference's step order also the same verdict histogram
    out["identical"] = bool(all(
        np.array_equal(canon(u0), canon(u)) and list(st.cells_per_depth[: d + 1]) == list(s0.cells_per_depth[: d + 1])
        and (name == "learned_batched" or list(st.verdict_hist) == list(s0.verdict_hist))
        for name, (_, u, st) in runs.items()))
    out["cells_evaluated"] = int(sum(s0.cells_per_depth[: d + 1]))
    out["cell_evals"] = int(s0.cell_evals)
    for name, (t, _, st) in runs.items():
        out[f"cell_evals_per_s_{name}"] = st.cell_evals / t
    if "learned_batched" in runs:
        out["cell_evals_learned"] = int(runs["learned_batched"][2].cell_evals)
        out["verdict_hist_learned"] = list(runs["learned_batched"][2].verdict_hist)
    out["trap_cells"] = int(s0.trap_cells)
    return out


def c3_only(forms) -> int:
    """The rectangle_size_bound branch-and-bound alone, in the given forms."""
    import circlecover_amd as cc

    with cc.Prover(0) as p:
        print(json.dumps(c3_forms(p, forms)), flush=True)
    return 0


def log(msg: str) -> None:
    """A progress line on stderr (the JSON result alone goes to stdout)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


RSB_SLICE = "24,4"  # lambda pieces 24..27 of the rectangle_size_bound proof grid (65,536 cells)


def full_proof() -> dict:
    """Both of the reference's proofs, end to end, through the drop-in
    executable (reference constants: tight_rectangle 256 lambda launches of
    128^4 pieces plus its refinement generations; rectangle_size_bound the
    256x128x128 grid with 64^4 pieces per cell plus one generation).  Wall
    time from process start to exit, including context set-up; the critical
    intervals are counted from the stdout report, whose rows sit between the
    banner and the closing rule.

    Then the same procedure in the enumeration's three forms
    (`--enum-form`): batched (the default), plain (the reference's kernel
    shape and operation sequence: one block per outer piece pair, no
    compaction) and plain-exact (that shape with the exact shortcuts):
    tight_rectangle in full, rectangle_size_bound on lambda pieces 24..27 of
    its grid (the plain form would take about 25 minutes for the whole
    grid); the reports must be identical."""
    import subprocess

    exe = os.path.join(ROOT, "circlecover_amd", "circlecover_prover")

    def run(args):
        t = time.perf_counter()
        r = subprocess.run([exe, *args], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, timeout=600)
        return time.perf_counter() - t, r

    out = {}
    for proof in ("tr", "rsb"):
        log(f"full proof {proof}")
        wall, r = run(["--only", proof])
        lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
        rows = max(0, len(lines) - 2) if r.returncode == 0 else None
        out[proof] = {"wall_s": wall, "exit_status": r.returncode, "critical_intervals": rows}
    forms = {}
    for name, args in (("tr", ["--only", "tr"]),
                       ("rsb_slice", ["--only", "rsb", "--rsb-la-range", RSB_SLICE, "--rsb-generations", "0"])):
        times, reports = {}, {}
        for form in ("batched", "plain_exact", "plain"):
            log(f"{name} in the {form} enumeration form")
            wall, r = run(args + ["--enum-form", form.replace("_", "-")])
            times[form + "_s"] = wall if r.returncode == 0 else None
            reports[form] = r.stdout
        forms[name] = {**times, "args": " ".join(args),
                       "identical_reports": len(set(reports.values())) == 1,
                       "batched_vs_plain": times["plain_s"] / times["batched_s"],
                       "batched_vs_plain_exact": times["plain_exact_s"] / times["batched_s"]}
    out["forms"] = forms
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--boxes", type=int, default=N_BOXES)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline leg")
    ap.add_argument("--core-only", action="store_true", help="skip the other configurations (workloads)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the HBM counter runs (traffic = null)")
    ap.add_argument("--no-full-proof", action="store_true", help="skip the end-to-end proofs (workloads.full_proof)")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--c3-only", action="store_true",
                    help="run only the rectangle_size_bound branch-and-bound (256x64x64 grid, depth 24) in both "
                         "forms, once each after a warm-up, and print its timings (for profiling)")
    ap.add_argument("--c3-forms", default=",".join(C3_FORMS),
                    help="with --c3-only: the forms to run (comma-separated)")
    args = ap.parse_args()
    if args.traffic_probe:
        return traffic_probe()
    if args.c3_only:
        return c3_only(tuple(args.c3_forms.split(",")))
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(cells are independent: each GPU would take a slice of the boxes or initial cells)", file=sys.stderr)
        return 2

    from circlecover_amd import Prover, gen_c2_boxes

    n = args.boxes
    soa = gen_c2_boxes(SEED, n)
    with Prover(0) as p:
        d_soa = p.alloc(soa.nbytes)
        d_mask = p.alloc(4 * n)
        d_first = p.alloc(n)
        d_soa.upload(soa)
        log(f"batch of {n} boxes: {args.warmup} warm-up and {args.steps} timed steps")
        for _ in range(args.warmup):
            p.eval_batch_async(d_soa, n, n, d_mask, d_first)
        p.synchronize()
        t0 = time.perf_counter()
        p.timer_start()
        for _ in range(args.steps):
            p.eval_batch_async(d_soa, n, n, d_mask, d_first)
        ev_ms = p.timer_stop()
        p.synchronize()
        wall = time.perf_counter() - t0
        mask = d_mask.download(np.empty(n, dtype=np.uint32))
        for b in (d_soa, d_mask, d_first):
            b.free()
        if not args.core_only:
            log("other configurations")
        workloads = None if args.core_only else extended(p, soa, mask)

    ms_per_step = 1e3 * wall / args.steps
    value = n * N_CRIT / (wall / args.steps)
    launch_ms = ev_ms / args.steps

    if not args.no_traffic:
        log("HBM counter runs")
    traffic = None if args.no_traffic else measured_traffic()

    sample = soa[:, : min(n, 1 << 18)]
    opb = ops_per_box(sample)
    ops_launch = opb * n
    achieved = ops_launch / (launch_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched-criteria-4d", "boxes": n, "criteria": N_CRIT, "seed": hex(SEED),
                   "precision": "fp64-outward", "proven_fraction": float(((mask & 0x3FF) != 0).mean())},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": None if traffic is None else traffic["bytes"],
                     "algorithmic_bytes": 69.0 * n, "traffic_detail": traffic,
                     "ops_per_box": opb, "ops_per_launch": ops_launch, "launch_ms": launch_ms,
                     "kernel": "ccrd_c2_eval"},
    }
    if workloads is not None:
        # the metric's comparison: learned+batched against the plain GPU form
        # on the branch-and-bound workload (not part of value)
        c3 = workloads["rsb_bnb"]
        if "learned_batched_s" in c3 and "plain_s" in c3:
            out["comparison"] = {
                "workload": "rsb_bnb 256x64x64 lambda [1, 8] depth 24",
                "plain_s": c3["plain_s"], "plain_exact_s": c3.get("plain_exact_s"),
                "batched_s": c3.get("batched_s"), "learned_batched_s": c3["learned_batched_s"],
                "learned_batched_vs_plain": c3["plain_s"] / c3["learned_batched_s"],
                "learned_batched_vs_plain_exact": (c3["plain_exact_s"] / c3["learned_batched_s"]
                                                   if "plain_exact_s" in c3 else None),
                "identical_unproven": c3["identical"]}
        if not args.no_full_proof:
            fp = full_proof()
            workloads["full_proof"] = fp
            # the metric's full-proof wall time: batched against the plain
            # (reference-shaped) enumeration, same reports
            if "comparison" in out:
                out["comparison"]["full_proof"] = {
                    k: {f: v.get(f) for f in ("batched_s", "plain_s", "plain_exact_s", "batched_vs_plain",
                                              "batched_vs_plain_exact", "identical_reports", "args")}
                    for k, v in fp["forms"].items()}
        out["workloads"] = workloads
    if not args.no_cpu:
        # every CPU this process may use, up to the 16 a one-GPU job is
        # granted on the measurement host (more threads would only share them)
        threads = max(1, min(len(os.sched_getaffinity(0)), 16))
        log(f"CPU oracle baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(soa, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
