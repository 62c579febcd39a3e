"""Throughput benchmark of the in-core eq15 depth march on one MI355X.

Workload (one beam, resident in HBM): water, 230 MeV Gaussian pencil
(sigma_yz 0.3 cm), 21x21 transverse nodes over 8x8 cm, 20x20 directions,
500 energy groups on [1, 260] MeV, dt = 0.02 cm (Nx = 2000, 40 cm),
eq15 with straggling, primary + secondary with the catastrophic source.
FP32 state, FP64 energy recurrence and IDD accumulation.

A "step" is one Strang depth step of both species (E/2, T/2, A, T/2, E/2 and
the source), i.e. Ng*(Ny+1)*(Nz+1)*Nmu*Nom phase-space cell-updates.  The
warmup steps include step 0 (no secondary), so every timed step carries the
full secondary chain.

Prints ONE JSON line (rank 0).  Fields beyond the common contract:
  roofline      dominant operator: algorithmic HBM bytes per call / mean call
                time from HIP events on the solver's stream in an
                instrumented copy of the timed region; peak 8 TB/s
  phases        per-operator seconds/calls/GB/s of that instrumented run
  cpu_baseline  the CPU FP64 oracle on this host, one full depth step of the
                same workload (bounded sample)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "phase-space cell-updates/s; beamlet solves/s + p99 latency at IDD err ≤0.5%"


def workload_case():
    from bfpn_gpu_amd.abi import Case
    from bfpn_gpu_amd.tables import material_dir

    return Case(Nx=2000, Ny=20, Nz=20, Ng=500, Nmu=20, Nom=20, Ly=8, Lz=8, Lu=2, Lv=2,
                energy=230.0, sigma_yz=0.3, material="water", data=material_dir("water"),
                time=40.0, flags=dict(eq15_straggling=1))


def cpu_baseline(case, threads: int | None):
    """One full depth step (step 1: primary + secondary) of the workload with the
    CPU FP64 oracle, after an untimed step 0."""
    if threads:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle

    oracle.build()
    s = oracle.Solver(case)
    s.advance(1)
    t0 = time.perf_counter()
    s.advance(1)
    dt = time.perf_counter() - t0
    cells = s.cells
    n = oracle.threads()
    s.close()
    return {"value": cells / dt, "unit": "cell-updates/s", "cores": n, "kind": "port",
            "sample": f"CPU FP64 oracle (C++/OpenMP), 1 full depth step (primary+secondary) of the "
                      f"same workload, {cells} cells, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the solver runs one beam per GPU and has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)
    if args.warmup < 1:
        args.warmup = 1

    import numpy as np

    import bfpn_gpu_amd as bg

    case = workload_case()
    prec = bg.FP32 if args.precision == "fp32" else bg.FP64
    solver = bg.Solver(case, prec)
    cells = solver.cells
    print(f"[bench] workload ready: {cells} cells, {solver.device_bytes / 2**30:.2f} GiB resident",
          file=sys.stderr, flush=True)

    solver.advance(args.warmup)
    solver.history()  # synchronizes the stream
    t0 = time.perf_counter()
    solver.advance(args.steps)
    idd, _ = solver.history()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = cells * args.steps / elapsed
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.3f} ms/step", file=sys.stderr,
          flush=True)
    assert np.all(np.isfinite(idd)) and idd[-1] > 0

    # instrumented copy of the timed region: per-operator HIP events
    solver.set_profiling(True)
    solver.advance(args.steps)
    stats = solver.phase_stats()
    solver.set_profiling(False)
    solver.close()
    phases = {}
    for name, (sec, calls, byt) in stats.items():
        if calls:
            phases[name] = {"seconds": round(sec, 6), "calls": calls,
                            "ms_per_call": round(sec / calls * 1e3, 4),
                            "GB_per_call": round(byt / calls / 1e9, 4),
                            "GBps": round(byt / sec / 1e9, 1) if sec > 0 else None}
    dom = max(phases, key=lambda k: phases[k]["seconds"])
    ach = phases[dom]["GBps"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None}
    step_bytes = sum(v[2] for v in stats.values()) / args.steps
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == bg.FP32 else "f64",
        "data": "synthetic",
        "config": {"workload": "single beam, water 230 MeV, 21x21 nodes x 20x20 directions x 500 "
                               "groups, dt 0.02 cm, eq15 + straggling + secondary source",
                   "cells": cells, "precision": "FP32 state; FP64 energy recurrence and IDD"
                   if prec == bg.FP32 else "FP64",
                   "step_fraction_of_hbm_roofline": round(step_bytes / (ms_per_step * 1e-3) / 1e9
                                                          / HBM_PEAK_GBS, 4)},
        "roofline": roofline,
        "phases": phases,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
