#!/usr/bin/env python3
"""Benchmark: interval cell-evaluations per second on one MI355X.

Workload (batched criterion throughput): 2^22 seeded synthetic 4-D boxes
(lambda, r1, r2, r3), FP64 outward-rounded intervals, ten covering criteria
per box evaluated without short-circuit; one step = one pass over all boxes
= 41,943,040 cell-evals (one cell-eval = one (box, criterion) predicate).
Boxes are generated on the host and uploaded before the timed region.

Prints ONE JSON line.  `roofline.achieved` counts the algorithmic directed
FP64 operations (reference op sequence, counted by the CPU oracle's counting
build on a seeded sample and scaled to the batch) per kernel launch, divided
by the launch duration measured with HIP events on the launch stream.
`cpu_baseline` times the CPU oracle on the host cores on the same boxes.
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


def cpu_baseline(soa: np.ndarray, threads: int) -> dict:
    import oracle

    n = soa.shape[1]
    t0 = time.perf_counter()
    oracle.c2_eval(soa, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": n * N_CRIT / dt, "unit": "cell-evals/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of the {N_BOXES} seeded boxes, all ten criteria, CPU oracle "
                      f"(x86 MXCSR directed rounding), {threads} threads, {dt:.2f} s"}


def ops_per_box(soa: np.ndarray) -> float:
    import oracle

    ops = oracle.c2_count_ops(soa)
    return float(ops.sum()) / soa.shape[1]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--boxes", type=int, default=N_BOXES)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the workload shards per GPU as described in DESIGN.md)", file=sys.stderr)
        return 2

    from circlecover_amd import Prover, gen_c2_boxes

    n = args.boxes
    soa = gen_c2_boxes(SEED, n)
    with Prover(0) as p:
        d_soa = p.alloc(soa.nbytes)
        d_mask = p.alloc(4 * n)
        d_first = p.alloc(n)
        d_soa.upload(soa)
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

    ms_per_step = 1e3 * wall / args.steps
    value = n * N_CRIT / (wall / args.steps)
    launch_ms = ev_ms / args.steps

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
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "ops_per_box": opb, "ops_per_launch": ops_launch, "launch_ms": launch_ms,
                     "kernel": "ccrd_c2_eval"},
    }
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        threads = max(1, min(threads, 16))
        out["cpu_baseline"] = cpu_baseline(soa, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
