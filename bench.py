"""Throughput benchmark of the in-core eq15 depth march on one MI355X.

Default workload ("paper", the paper grid with its whole state resident in
HBM): water, 230 MeV Gaussian pencil (sigma_yz 0.3 cm), 81x81 transverse nodes
over 8x8 cm, 20x20 directions, 500 energy groups on [1, 260] MeV, dt = 0.01 cm
(Nx = 4000, 40 cm), eq15 with straggling, primary + secondary with the
catastrophic source.  FP32 state, FP64 energy recurrence and IDD accumulation.
The beam is synthetic (the Eq.23 Gaussian pencil); the material tables are the
reference's water tables.

A "step" is one Strang depth step of both species (E/2, T/2, A, T/2, E/2 and
the source), i.e. Ng*(Ny+1)*(Nz+1)*Nmu*Nom phase-space cell-updates.  The
warmup steps include step 0 (no secondary), so every timed step carries the
full secondary chain.

Prints ONE JSON line.  Fields beyond the common contract:
  roofline      dominant operator: algorithmic HBM bytes per launch / mean
                launch time from HIP events on the solver's stream in an
                instrumented copy of the timed region; peak 8 TB/s.  traffic:
                HBM bytes per launch of that operator measured with rocprofv3
                FETCH_SIZE / WRITE_SIZE (calibrated on a plain copy), read from
                the committed summaries profiles/hbm_traffic_single_beam.md and
                profiles/hbm_traffic_paper_grid.md (FP32 state only)
  phases        per-operator seconds/calls/GB/s of that instrumented run
  accuracy      the same steps with the FP64 instantiation: max relative IDD
                error of the timed FP32 run over all its steps
  cpu_baseline  (kind "port") the CPU FP64 oracle on this host (all cores), a
                bounded sample of the same workload: the same physics on a
                7x7-node transverse slab at the same node spacing, full depth
                steps after an untimed one (cell-updates/s does not depend on
                the slab width); "reference": the reference's own CPU
                prototype serial_bfpn (single-threaded, not eq15: its grid and
                physics are fixed in its source), one depth step

Other workloads (--workload), each one JSON line:
  beam      21x21 nodes, Nx = 2000 (the single-beam configuration)
  beamlets  a seeded sample of the 64-beam batch (E ~ U[70,230] MeV, sigma
            ~ U[0.2,0.5] cm, 21x21 or 41x41 nodes, Ng = 300, water/bone, depth
            1.2 x CSDA range): one-at-a-time solver vs the batching server
  serve     seeded Poisson arrivals (rate, first N requests) into the server:
            solves/s and arrival->IDD latency p50/p99
  accuracy  one case (--grid beam|paper, --material water|bone) to full depth
            with FP32 state and with the FP64 instantiation: IDD error above 10 %
            of the maximum and BP/D90/D20 shifts (gate 0.5 %, 0.1 mm); with
            --slab A:B:MATERIAL[:RHO] the beam crosses that slab in water
            (depth layers), the metrics are taken behind it and its D90 shift
            against the homogeneous water run is reported; with
            --insert Y0:Y1:Z0:Z1:X0:X1:MATERIAL[:RHO] a box of that material
            sits in water (a material map), and the mapped step's time is
            reported beside the homogeneous one
  copy      the plain device copy used to calibrate the traffic counters
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import re
import subprocess
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STEP_BYTES_PER_CELL = 176.0  # algorithmic HBM bytes per cell-update of the step model
METRIC = "phase-space cell-updates/s; beamlet solves/s + p99 latency at IDD err ≤0.5%"
TRAFFIC_FILES = {"beam": os.path.join(ROOT, "profiles", "hbm_traffic_single_beam.md"),
                 "paper": os.path.join(ROOT, "profiles", "hbm_traffic_paper_grid.md")}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def beam_case(Ny=20, Nx=2000, time_cm=40.0, material="water"):
    from bfpn_gpu_amd.abi import Case
    from bfpn_gpu_amd.tables import material_dir

    return Case(Nx=Nx, Ny=Ny, Nz=Ny, Ng=500, Nmu=20, Nom=20, Ly=8, Lz=8, Lu=2, Lv=2,
                energy=230.0, sigma_yz=0.3, material=material, data=material_dir(material),
                time=time_cm, flags=dict(eq15_straggling=1))


def cpu_baseline(case, threads: int | None, nodes: int = 7, steps: int = 3):
    """The CPU FP64 oracle on the same physics over a nodes x nodes transverse
    slab with the workload's node spacing: one untimed step, `steps` timed
    full steps."""
    if threads:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle

    oracle.build()
    dy = case.Ly / case.Ny
    w = nodes - 1
    sample = dataclasses.replace(case, Ny=w, Nz=w, Ly=w * dy, Lz=w * dy)
    s = oracle.Solver(sample)
    s.advance(1)
    t0 = time.perf_counter()
    s.advance(steps)
    dt = time.perf_counter() - t0
    cells = s.cells
    n = oracle.threads()
    s.close()
    return {"value": steps * cells / dt, "unit": "cell-updates/s", "cores": n, "kind": "port",
            "sample": f"CPU FP64 oracle (C++/OpenMP, {n} threads): same physics on a "
                      f"{nodes}x{nodes}-node transverse slab at {dy:.2f} cm spacing ({cells} "
                      f"cells), {steps} full depth steps (primary+secondary) after an untimed "
                      f"one, {dt:.2f} s"}


def reference_cpu_baseline():
    """The reference's CPU prototype serial_bfpn (BFPn_CPU_Solver/serial_bfpn.cpp,
    built into oracle/_ref by oracle/Makefile), one depth step.  It is single
    threaded and not eq15 (periodic DFT angles, first-order upwind transport,
    exponential attenuation; serial_bfpn.cpp:92-219); its grid is fixed in its
    source: 11x11 nodes, 11x11 directions, 500 groups, 230 MeV
    (serial_bfpn.cpp:15-43, 289).  Nx = 100 with --time 0.4 is one step."""
    from bfpn_gpu_amd.tables import material_dir

    exe = os.path.join(ROOT, "oracle", "_ref", "serial_bfpn")
    if not os.path.exists(exe):
        return {"skipped": "oracle/_ref/serial_bfpn not built"}
    # the synthetic table has the 500 rows the prototype reads (the shipped
    # water table has 499 and is refused by it)
    proc = subprocess.run([exe, "100", "--data", material_dir("synthetic"), "--time", "0.4"],
                          capture_output=True, text=True, timeout=900)
    if proc.returncode != 0:
        return {"skipped": f"serial_bfpn exited {proc.returncode}: {proc.stderr.strip()[:200]}"}
    sec = float(re.search(r"Elapsed seconds:\s*([0-9.eE+-]+)", proc.stdout).group(1))
    steps = int(re.search(r"Steps:\s*(\d+)", proc.stdout).group(1))
    cells = 500 * 11 * 11 * 11 * 11
    host = os.cpu_count() or 0
    return {"value": steps * cells / sec, "unit": "cell-updates/s", "cores": 1,
            "kind": "reference", "seconds_per_step": sec / steps,
            "sample": f"reference CPU prototype serial_bfpn (non-eq15; primary only), 1 of "
                      f"{host} host cores, its fixed grid of 11x11 nodes x 11x11 directions x "
                      f"500 groups ({cells} cells), {steps} depth step(s) in {sec:.1f} s"}


def load_traffic(grid: str, phase: str):
    """HBM bytes per launch of `phase` on `grid` from the committed rocprofv3 summary."""
    try:
        text = open(TRAFFIC_FILES[grid]).read()
    except OSError:
        return None
    m = re.search(r"```json\s*(\{.*?\})\s*```", text, re.S)
    if not m:
        return None
    data = json.loads(m.group(1))
    return data.get("bytes_per_launch", {}).get(phase)


def timed_steps(solver, steps, warmup):
    solver.advance(warmup)
    solver.history()  # synchronizes the stream
    t0 = time.perf_counter()
    solver.advance(steps)
    idd, _ = solver.history()
    return time.perf_counter() - t0, idd


def phases_of(solver, steps):
    """Per-kernel device time of `steps` more steps, issued in order on one
    stream (overlap off) so that each interval holds one kernel alone."""
    solver.set_overlap(False)
    solver.set_profiling(True)
    solver.advance(steps)
    stats = solver.phase_stats()
    solver.set_profiling(False)
    solver.set_overlap(True)
    phases = {}
    for name, (sec, calls, byt) in stats.items():
        if calls:
            phases[name] = {"seconds": round(sec, 6), "calls": calls,
                            "ms_per_call": round(sec / calls * 1e3, 4),
                            "GB_per_call": round(byt / calls / 1e9, 4),
                            "GBps": round(byt / sec / 1e9, 1) if sec > 0 else None}
    step_bytes = sum(v[2] for v in stats.values()) / steps
    return phases, step_bytes


def run_resident(args, case, workload, grid):
    import numpy as np

    import bfpn_gpu_amd as bg

    prec = bg.FP32 if args.precision == "fp32" else bg.FP64
    solver = bg.Solver(case, prec)
    # the headline computes every group of every step (active-group skipping
    # off); the same steps with skipping on are timed separately below
    solver.set_group_skip(False)
    cells = solver.cells
    log(f"workload ready: {cells} cells, {solver.device_bytes / 2**30:.2f} GiB resident")
    elapsed, idd = timed_steps(solver, args.steps, args.warmup)
    ms_per_step = elapsed / args.steps * 1e3
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step")
    assert np.all(np.isfinite(idd)) and idd[-1] > 0
    phases, step_bytes = phases_of(solver, args.steps)
    solver.close()
    skip = None
    if not args.no_skip_figure:
        with bg.Solver(case, prec) as s2:  # skipping on (the library default)
            e2, idd2 = timed_steps(s2, args.steps, args.warmup)
            top = s2.active_groups()
        skip = {"ms_per_step": round(e2 / args.steps * 1e3, 4),
                "value": cells * args.steps / e2,
                "bit_identical_idd": bool(np.array_equal(idd2, idd[:len(idd2)])),
                "primary_top_group": top[0], "secondary_top_group": top[1], "Ng": case.Ng,
                "note": "the groups above each species' top nonzero group are exact zeros and "
                        "are skipped; same steps, same results"}
        log(f"with active-group skipping: {skip['ms_per_step']:.3f} ms/step, top {top}")
    accuracy = None
    if prec == bg.FP32 and not args.no_fp64_check:
        # the same steps with the FP64 instantiation (after the timed region)
        n = len(idd)
        t0 = time.perf_counter()
        with bg.Solver(case, bg.FP64) as ref:
            ref.advance(n)
            ref_idd, _ = ref.history()
        rel = float(np.max(np.abs(idd - ref_idd[:n]) / np.abs(ref_idd[:n])))
        accuracy = {"vs": "FP64 instantiation, same case and steps", "steps": n,
                    "max_rel_idd_err": rel, "seconds": round(time.perf_counter() - t0, 1)}
        log(f"FP64 check over {n} steps: max rel IDD error {rel:.2e}")
    dom = max(phases, key=lambda k: phases[k]["seconds"])
    ach = phases[dom]["GBps"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": load_traffic(grid, dom) if args.precision == "fp32" else None,
                "measured": "HIP events around each launch, one stream, after the timed steps"}
    return {
        "metric": METRIC,
        "value": cells * args.steps / elapsed,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == bg.FP32 else "f64",
        "data": "synthetic Gaussian pencil beam; reference material tables",
        "config": {"workload": workload, "cells": cells,
                   "precision": "FP32 state; FP64 energy recurrence and IDD"
                   if prec == bg.FP32 else "FP64",
                   # the step model: 176 B of HBM traffic per cell-update (both
                   # species, FP32 state) at 8 TB/s is 45.5 G cell-updates/s
                   "step_fraction_of_hbm_roofline":
                       round(cells / (ms_per_step * 1e-3) * STEP_BYTES_PER_CELL
                             * (2 if prec == bg.FP64 else 1) / 1e9 / HBM_PEAK_GBS, 4),
                   "moved_bytes_per_step_GB": round(step_bytes / 1e9, 3)},
        "roofline": roofline,
        "phases": phases,
        "accuracy": accuracy,
        "active_group_skip": skip,
    }


def idd_error(idd, ref, dt):
    import numpy as np

    from bfpn_gpu_amd import metrics

    c = metrics.compare(idd, ref)
    x = (np.arange(len(ref)) + 1) * dt
    a, b = metrics.analyze(x, idd), metrics.analyze(x, ref)
    shift = max(abs(a[k] - b[k]) for k in ("BP", "D90", "D20"))
    return {"max_rel": float(c["max_rel"]), "max_shift_mm": float(10 * shift)}


def solve_each(cases, skip):
    """One fresh solver per beam, one after another (the reference's
    one-process-per-beam model): (seconds, IDDs)."""
    import bfpn_gpu_amd as bg

    t0 = time.perf_counter()
    out = []
    for k, c in enumerate(cases):
        with bg.Solver(c, bg.FP32) as s:
            s.set_group_skip(skip)
            out.append(s.run()[0])
        if (k + 1) % 8 == 0:
            log(f"  one at a time (skipping {'on' if skip else 'off'}): {k + 1}/{len(cases)}, "
                f"{time.perf_counter() - t0:.1f} s")
    return time.perf_counter() - t0, out


def run_beamlets(args):
    import numpy as np

    import bfpn_gpu_amd as bg
    from bfpn_gpu_amd.workloads import beamlet_batch

    cases = beamlet_batch(64)[: args.beams]
    log(f"{len(cases)} beams: " + ", ".join(f"{c.energy:.0f}MeV/{c.Ny}/{c.material}" for c in cases))
    # the plain solver computes every energy group (the reference's in-core
    # arithmetic); the same solver with active-group skipping; then the server
    t_full, plain = solve_each(cases, skip=False)
    log(f"one at a time, every group: {t_full:.2f} s")
    t_skip, plain_skip = solve_each(cases, skip=True)
    log(f"one at a time, active groups only: {t_skip:.2f} s")
    t0 = time.perf_counter()
    with bg.Server(max_active=args.max_active, slice_steps=args.slice_steps or 32) as srv:
        ids = [srv.submit(c) for c in cases]
        served = []
        for k, i in enumerate(ids):
            served.append(srv.wait(i, timeout=3600)[0])
            if (k + 1) % 8 == 0:
                log(f"  server: {k + 1}/{len(ids)} results, {time.perf_counter() - t0:.1f} s")
        st = srv.stats()
    t_srv = time.perf_counter() - t0
    log(f"server: {t_srv:.2f} s, peak active {st['peak_active']}")
    same = all(np.array_equal(a, b) and np.array_equal(a, c)
               for a, b, c in zip(plain, plain_skip, served))
    steps = sum(len(i) for i in plain)
    cell_steps = sum(len(i) * c.cells for i, c in zip(plain, cases))
    out = {"metric": "beamlet solves/s", "value": len(cases) / t_srv, "unit": "solves/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
           "data": "synthetic Gaussian pencil beams; reference material tables",
           "config": {"workload": f"beamlets: {len(cases)} of the seeded 64-beam batch (E ~ "
                                  f"U[70,230] MeV, sigma ~ U[0.2,0.5] cm, 21x21 or 41x41 nodes, "
                                  f"20x20 directions, Ng 300, water/bone, depth 1.2 x CSDA range, "
                                  f"dt 0.02 cm)", "depth_steps": steps,
                      "cell_steps": cell_steps, "max_active": args.max_active,
                      "slice_steps": args.slice_steps or 32},
           "one_at_a_time_every_group": {"value": len(cases) / t_full, "seconds": t_full},
           "one_at_a_time_active_groups": {"value": len(cases) / t_skip, "seconds": t_skip},
           "server": {"seconds": t_srv, "peak_active": st["peak_active"],
                      "cached_tables": st["cached_tables"]},
           "speedup_vs_every_group": t_full / t_srv,
           "speedup_vs_active_groups": t_skip / t_srv,
           "served_equals_alone": same}
    if args.fp64_beams:
        errs = []
        for c, idd in list(zip(cases, served))[: args.fp64_beams]:
            with bg.Solver(c, bg.FP64) as s:
                ref = s.run()[0]
            errs.append(idd_error(idd, ref, c.dt))
        out["accuracy_vs_fp64"] = {"beams": len(errs),
                                   "max_rel": max(e["max_rel"] for e in errs),
                                   "max_shift_mm": max(e["max_shift_mm"] for e in errs),
                                   "each_max_rel": [e["max_rel"] for e in errs]}
    return out


def run_serve(args):
    import numpy as np

    import bfpn_gpu_amd as bg
    from bfpn_gpu_amd.workloads import serving_requests

    stream = serving_requests(args.rate, 60.0)
    reqs = stream[: args.requests]
    log(f"{len(reqs)} of {len(stream)} requests at {args.rate}/s: " +
        ", ".join(f"{c.energy:.0f}MeV/{c.Ny}/{c.material}" for _, c in reqs))
    lat = [None] * len(reqs)
    served = [None] * len(reqs)
    with bg.Server(max_active=args.max_active, slice_steps=args.slice_steps or 8) as srv:
        t0 = time.perf_counter()
        ids = [None] * len(reqs)
        late = [0.0] * len(reqs)  # submit time past the scheduled arrival
        done = threading.Event()

        def submitter():
            for k, (at, c) in enumerate(reqs):
                dt = t0 + at - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                late[k] = max(0.0, time.perf_counter() - (t0 + at))
                ids[k] = srv.submit(c)
            done.set()

        th = threading.Thread(target=submitter)
        th.start()
        done.wait()
        th.join()
        # latency = arrival -> IDD ready, from the server's own submit/done times
        # (waiting here in submission order must not delay the measurement)
        for k, rid in enumerate(ids):
            served[k], times = srv.wait(rid, timeout=7200)
            lat[k] = late[k] + times["total"]
            log(f"  request {k}: {lat[k]:.1f} s")
        t_end = time.perf_counter()
        st = srv.stats()
    lat = np.array(lat)
    out = {"metric": "beamlet solves/s", "value": len(reqs) / (t_end - t0), "unit": "solves/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
           "data": "synthetic Gaussian pencil beams; reference material tables",
           "config": {"workload": f"serve: the first {len(reqs)} of {len(stream)} seeded Poisson "
                                  f"arrivals at {args.rate}/s over 60 s (E ~ U[70,230] MeV, "
                                  f"21x21/41x41/81x81 nodes, 20x20 directions, Ng 300, water/bone,"
                                  f" full depth 1.2 x CSDA range)",
                      "max_active": args.max_active, "slice_steps": args.slice_steps or 8,
                      "requests": len(reqs), "offered_requests": len(stream)},
           "latency_s": {"p50": float(np.percentile(lat, 50)),
                         "p99": float(np.percentile(lat, 99)), "max": float(lat.max()),
                         "samples": len(lat), "each": [round(float(v), 3) for v in lat]},
           "capacity_note": "the GPU stays busy from the first arrival to the last result, so "
                            "solves/s here is this mix's capacity; the offered rate exceeds it "
                            "and latency is mostly queueing",
           "server": {"peak_active": st["peak_active"], "cached_tables": st["cached_tables"]}}
    errs = []
    for (_, c), idd in zip(reqs, served):
        if c.Ny > args.fp64_max_ny:
            continue
        with bg.Solver(c, bg.FP64) as s:
            ref = s.run()[0]
        errs.append(dict(idd_error(idd, ref, c.dt), Ny=c.Ny, energy=round(c.energy, 1),
                         material=c.material))
    if errs:
        out["accuracy_vs_fp64"] = {
            "results": len(errs), "of": len(reqs),
            "which": f"every served result with at most {args.fp64_max_ny + 1}x"
                     f"{args.fp64_max_ny + 1} nodes",
            "max_rel": max(e["max_rel"] for e in errs),
            "max_shift_mm": max(e["max_shift_mm"] for e in errs), "each": errs}
    return out


def run_accuracy(args):
    """FP32-state solver against the FP64 instantiation on the same case, full depth."""
    import numpy as np

    import bfpn_gpu_amd as bg

    from bfpn_gpu_amd.tables import material_dir

    ny, nx = (20, 2000) if args.grid == "beam" else (80, 4000)
    depth = args.depth or (40.0 if args.material == "water" else 22.0)
    case = beam_case(Ny=ny, Nx=nx, time_cm=depth, material=args.material)
    if args.material == "bone":
        case = dataclasses.replace(case, density=1.85)
    layers, slab = None, None
    if args.slab:  # A:B:MATERIAL[:RHO], in water
        f = args.slab.split(":")
        presets = {"water": 1.0, "bone": 1.85, "air": 0.001205}
        slab = (float(f[0]), float(f[1]), f[2], float(f[3]) if len(f) > 3 else presets[f[2]])
        water = ("water", 1.0, material_dir("water"))
        layers = [(slab[0], *water), (slab[1], slab[2], slab[3], material_dir(slab[2])),
                  (depth + 100.0, *water)]
    mapped, insert = None, None
    if args.insert:  # Y0:Y1:Z0:Z1:X0:X1:MATERIAL[:RHO], a box of MATERIAL in water
        f = args.insert.split(":")
        presets = {"water": 1.0, "bone": 1.85, "air": 0.001205}
        insert = tuple(float(v) for v in f[:6]) + (f[6], float(f[7]) if len(f) > 7 else presets[f[6]])
        y0, y1, z0, z1, x0, x1, mat, rho = insert
        yv = -0.5 * case.Ly + np.arange(ny + 1) * case.Ly / ny  # node coordinates
        zv = -0.5 * case.Lz + np.arange(ny + 1) * case.Lz / ny
        inside = ((yv[None, :] >= y0) & (yv[None, :] < y1) & (zv[:, None] >= z0) &
                  (zv[:, None] < z1)).ravel()  # node s = iz (Ny+1) + iy
        mp = np.zeros((3, inside.size), dtype=np.uint8)
        mp[1][inside] = 1
        water = ("water", 1.0, material_dir("water"))
        mapped = ([water, (mat, rho, material_dir(mat))], [x0, x1, depth + 100.0], mp)
        log(f"insert: {int(inside.sum())} of {inside.size} nodes {mat} from {x0} to {x1} cm")
    res = {}
    runs = (("fp32", bg.FP32), ("fp64", bg.FP64))
    if args.only:
        runs = tuple(r for r in runs if r[0] == args.only)
    for name, prec in runs:
        t0 = time.perf_counter()
        with bg.Solver(case, prec, layers=layers, mapped=mapped) as s:
            total = case.num_steps()
            done = 0
            while done < total:  # in chunks, with a progress line per chunk
                n = min(200, total - done)
                s.advance(n)
                done += n
                idd, _ = s.history()  # synchronizes
                log(f"{name}: step {done}/{total}, {time.perf_counter() - t0:.0f} s, "
                    f"top groups {s.active_groups()}")
                if args.save:  # the IDD so far, rewritten after every chunk
                    os.makedirs(args.save, exist_ok=True)
                    tag = f"{args.grid}_{args.material}" + ("_slab" if slab else "")
                    np.save(os.path.join(args.save, f"idd_{tag}_{name}.npy"), idd)
        res[name] = (idd, time.perf_counter() - t0)
        log(f"{name}: {len(idd)} steps in {res[name][1]:.1f} s")
    from bfpn_gpu_amd import metrics

    if args.only:  # one precision per run: the IDD is saved for a later --compare
        idd, sec = res[args.only]
        x = (np.arange(len(idd)) + 1) * case.dt
        return {"metric": f"{args.only} full-depth run (IDD saved for comparison)",
                "config": {"workload": f"{args.grid} grid ({ny + 1}x{ny + 1} nodes), "
                                       f"{args.material}, depth {depth} cm", "cells": case.cells},
                "steps": len(idd), "seconds": sec, "metrics_cm": metrics.analyze(x, idd)}
    if slab:
        return slab_result(args, case, slab, res, ny, nx, depth)
    if insert:
        return insert_result(args, case, insert, res, ny, nx, depth)
    err = idd_error(res["fp32"][0], res["fp64"][0], case.dt)

    x = (np.arange(len(res["fp64"][0])) + 1) * case.dt
    return {"metric": "IDD error of the FP32-state solver vs the FP64 instantiation",
            "config": {"workload": f"{args.grid} grid ({ny + 1}x{ny + 1} nodes), {args.material} "
                                   f"rho {case.density or 1.0}, 230 MeV, depth {depth} cm, "
                                   f"Nx {nx}, Ng 500, 20x20 directions, eq15 + straggling + "
                                   f"secondary source", "cells": case.cells},
            "max_rel_err_above_10pct": err["max_rel"], "max_shift_mm_BP_D90_D20": err["max_shift_mm"],
            "fp64_metrics_cm": metrics.analyze(x, res["fp64"][0]),
            "seconds": {"fp32": res["fp32"][1], "fp64": res["fp64"][1]},
            "gate": {"max_rel_err": 0.005, "max_shift_mm": 0.1,
                     "pass": bool(err["max_rel"] <= 0.005 and err["max_shift_mm"] <= 0.1)}}


def insert_result(args, case, insert, res, ny, nx, depth):
    """FP32 vs FP64 with a material insert (a material map), the cost of a
    mapped step against the homogeneous one, and the D90 shift the insert
    causes."""
    import numpy as np

    import bfpn_gpu_amd as bg
    from bfpn_gpu_amd import metrics

    y0, y1, z0, z1, x0, x1, mat, rho = insert
    err = idd_error(res["fp32"][0], res["fp64"][0], case.dt)
    n = case.num_steps()
    x = (np.arange(n) + 1) * case.dt
    times = {}
    for label, skip in (("default", True), ("every group, one stream", False)):
        with bg.Solver(case, bg.FP32) as s:
            s.set_group_skip(skip)
            s.set_overlap(skip)
            t0 = time.perf_counter()
            s.advance(n)
            water, _ = s.history()
            times[label] = time.perf_counter() - t0
    log(f"homogeneous water: {times}")
    mw, m64 = metrics.analyze(x, water), metrics.analyze(x, res["fp64"][0])
    return {"metric": "IDD error of the FP32-state solver vs the FP64 instantiation with a "
                      "material insert",
            "config": {"workload": f"{args.grid} grid ({ny + 1}x{ny + 1} nodes), water 230 MeV "
                                   f"with a {mat} box (rho {rho}) over y [{y0}, {y1}), "
                                   f"z [{z0}, {z1}) cm from depth {x0} to {x1} cm, depth {depth} "
                                   f"cm, Nx {nx}, Ng 500, 20x20 directions, eq15 + straggling + "
                                   f"secondary source", "cells": case.cells},
            "max_rel_err_above_10pct": err["max_rel"], "max_shift_mm_BP_D90_D20": err["max_shift_mm"],
            "fp64_metrics_cm": m64,
            "d90_shift_vs_homogeneous_water_cm": float(m64["D90"] - mw["D90"]),
            "seconds": {"fp32 mapped": res["fp32"][1], "fp64 mapped": res["fp64"][1],
                        **{f"fp32 homogeneous, {k}": v for k, v in times.items()}},
            "ms_per_step": {"fp32 mapped": 1e3 * res["fp32"][1] / n,
                            **{f"fp32 homogeneous, {k}": 1e3 * v / n for k, v in times.items()}},
            "gate": {"max_rel_err": 0.005, "max_shift_mm": 0.1,
                     "pass": bool(err["max_rel"] <= 0.005 and err["max_shift_mm"] <= 0.1)}}


def slab_result(args, case, slab, res, ny, nx, depth):
    """FP32 vs FP64 through a slab, with the metrics behind it; the D90 shift
    against the same beam in homogeneous water (an FP32 run)."""
    import numpy as np

    import bfpn_gpu_amd as bg
    from bfpn_gpu_amd import metrics

    a, b, mat, rho = slab
    idd32, idd64 = res["fp32"][0], res["fp64"][0]
    x = (np.arange(len(idd64)) + 1) * case.dt
    behind = x > b  # the IDD scales by 1/density: the slab's own values are not comparable
    c = metrics.compare(idd32[behind], idd64[behind])
    m32, m64 = metrics.analyze(x[behind], idd32[behind]), metrics.analyze(x[behind], idd64[behind])
    shift = 10 * max(abs(m32[k] - m64[k]) for k in ("BP", "D90", "D20"))
    t0 = time.perf_counter()
    with bg.Solver(case, bg.FP32) as s:
        s.advance(case.num_steps())
        water, _ = s.history()
    mw = metrics.analyze(x, water)
    log(f"homogeneous water: {time.perf_counter() - t0:.1f} s, D90 {mw['D90']:.4f} cm")
    # CSDA prediction from the solver's own stopping powers: (b + R_w(E_out)) - R_w(E0)
    Ng, Lg = case.Ng, case.Lg
    E = np.concatenate(([1.0], (np.arange(1, Ng) - 0.5) * Lg / Ng + 1.0, [1.0 + Lg]))
    Sw, Sm = bg.stopping_power("water", Ng, Lg), bg.stopping_power(mat, Ng, Lg)
    e_fine = np.linspace(1.0, case.energy, 200001)
    inv_w = 1.0 / np.interp(e_fine, E, Sw)
    range_w = np.concatenate(([0.0], np.cumsum(0.5 * (inv_w[1:] + inv_w[:-1]) * np.diff(e_fine))))

    def slow(e0, t, S, dens):  # energy after t cm (RK4 on dE/dx = -dens S(E))
        h = t / 2000
        e = e0
        for _ in range(2000):
            k1 = -dens * np.interp(e, E, S)
            k2 = -dens * np.interp(e + 0.5 * h * k1, E, S)
            k3 = -dens * np.interp(e + 0.5 * h * k2, E, S)
            k4 = -dens * np.interp(e + h * k3, E, S)
            e += h * (k1 + 2 * k2 + 2 * k3 + k4) / 6
        return e

    e_out = slow(slow(case.energy, a, Sw, 1.0), b - a, Sm, rho)
    predicted = b + float(np.interp(e_out, e_fine, range_w)) - float(range_w[-1])
    return {"metric": "IDD error of the FP32-state solver vs the FP64 instantiation through a slab",
            "config": {"workload": f"{args.grid} grid ({ny + 1}x{ny + 1} nodes), water 230 MeV "
                                   f"with a {mat} slab (rho {rho}) from {a} to {b} cm, depth "
                                   f"{depth} cm, Nx {nx}, Ng 500, 20x20 directions, eq15 + "
                                   f"straggling + secondary source", "cells": case.cells},
            "max_rel_err_above_10pct_behind_slab": float(c["max_rel"]),
            "max_shift_mm_BP_D90_D20": float(shift),
            "fp64_metrics_cm_behind_slab": m64,
            "d90_shift_vs_homogeneous_water_cm": float(m64["D90"] - mw["D90"]),
            "d90_shift_csda_prediction_cm": predicted,
            "seconds": {"fp32": res["fp32"][1], "fp64": res["fp64"][1]},
            "gate": {"max_rel_err": 0.005, "max_shift_mm": 0.1,
                     "pass": bool(c["max_rel"] <= 0.005 and shift <= 0.1)}}


def compare_saved(args):
    """FP32 vs FP64 from IDDs saved by two --only runs (no GPU needed)."""
    import numpy as np

    nx = 2000 if args.grid == "beam" else 4000
    tag = f"{args.grid}_{args.material}"
    a = np.load(os.path.join(args.compare, f"idd_{tag}_fp32.npy"))
    b = np.load(os.path.join(args.compare, f"idd_{tag}_fp64.npy"))
    n = min(len(a), len(b))  # the common depth steps of the two runs
    a, b = a[:n], b[:n]
    from bfpn_gpu_amd import metrics

    dt = 40.0 / nx
    err = idd_error(a, b, dt)
    x = (np.arange(n) + 1) * dt
    return {"metric": "IDD error of the FP32-state solver vs the FP64 instantiation",
            "config": {"workload": tag, "steps": n, "depth_cm": n * dt},
            "fp64_metrics_cm": metrics.analyze(x, b), "fp32_metrics_cm": metrics.analyze(x, a),
            "max_rel_err_above_10pct": err["max_rel"], "max_shift_mm_BP_D90_D20": err["max_shift_mm"],
            "gate": {"max_rel_err": 0.005, "max_shift_mm": 0.1,
                     "pass": bool(err["max_rel"] <= 0.005 and err["max_shift_mm"] <= 0.1)}}


def run_copy(args):
    import numpy as np

    import bfpn_gpu_amd as bg

    n = int(args.copy_gib * (1 << 28))  # GiB of float32 per buffer
    a = bg.DeviceArray(shape=(n,), dtype=np.float32)
    b = bg.DeviceArray(shape=(n,), dtype=np.float32)
    bg.lib().bfpn_memset(a.ptr, 0, n * 4)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        bg.check(bg.lib().bfpn_device_copy(b.ptr, a.ptr, n * 4))
    dt = (time.perf_counter() - t0) / reps
    return {"metric": "device copy", "bytes_read_per_launch": n * 4,
            "bytes_written_per_launch": n * 4, "seconds_per_launch": dt,
            "GBps": 2 * n * 4 / dt / 1e9, "launches": reps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--workload", choices=("beam", "paper", "beamlets", "serve", "accuracy",
                                           "copy"), default="paper")
    ap.add_argument("--grid", choices=("beam", "paper"), default="beam")
    ap.add_argument("--material", choices=("water", "bone"), default="water")
    ap.add_argument("--depth", type=float, default=None,
                    help="accuracy: depth in cm (default 40 water, 22 bone)")
    ap.add_argument("--save", default=None, help="directory for the accuracy IDD arrays")
    ap.add_argument("--insert", default=None,
                    help="accuracy: Y0:Y1:Z0:Z1:X0:X1:MATERIAL[:RHO], a box of MATERIAL in water "
                         "(transverse node coordinates in cm, centred on the beam; depth X0 to "
                         "X1; write --insert=... when Y0 is negative)")
    ap.add_argument("--slab", default=None,
                    help="accuracy: A:B:MATERIAL[:RHO], a slab from A to B cm in water")
    ap.add_argument("--only", choices=("fp32", "fp64"), default=None,
                    help="accuracy: run one precision and save its IDD (needs --save)")
    ap.add_argument("--compare", default=None,
                    help="accuracy: compare the fp32/fp64 IDDs saved in this directory")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-fp64-check", action="store_true",
                    help="skip the FP64 rerun of the timed steps")
    ap.add_argument("--no-skip-figure", action="store_true",
                    help="skip the rerun of the timed steps with active-group skipping")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--copy-gib", type=float, default=1.0, help="copy: GiB per buffer")
    ap.add_argument("--beams", type=int, default=64)
    ap.add_argument("--fp64-beams", type=int, default=0,
                    help="beamlets: FP64 reruns of the first N served beams")
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--requests", type=int, default=12)
    ap.add_argument("--fp64-max-ny", type=int, default=40,
                    help="serve: FP64 reruns of the served results up to this Ny")
    ap.add_argument("--max-active", type=int, default=8)
    ap.add_argument("--slice-steps", type=int, default=None,
                    help="server: depth steps per beam per turn (default 32 for beamlets, where "
                         "throughput counts, 8 for serve, where a new arrival's first turn does)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the solver runs one beam per GPU and has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)
    if args.warmup < 1:
        args.warmup = 1

    if args.workload in ("beam", "paper"):
        if args.workload == "beam":
            case = beam_case()
            name = ("single beam, water 230 MeV, 21x21 nodes x 20x20 directions x 500 groups, "
                    "dt 0.02 cm, eq15 + straggling + secondary source")
        else:
            case = beam_case(Ny=80, Nx=4000)
            name = ("paper grid, water 230 MeV, 81x81 nodes x 20x20 directions x 500 groups, "
                    "dt 0.01 cm, eq15 + straggling + secondary source, state resident in HBM")
        out = run_resident(args, case, name, grid=args.workload)
        if not args.no_cpu:
            log("CPU baseline: oracle, bounded sample")
            out["cpu_baseline"] = cpu_baseline(case, args.cpu_threads)
            log("CPU baseline: reference prototype serial_bfpn, one step")
            out["cpu_baseline"]["reference"] = reference_cpu_baseline()
    elif args.workload == "beamlets":
        out = run_beamlets(args)
    elif args.workload == "serve":
        out = run_serve(args)
    elif args.workload == "accuracy" and args.compare:
        out = compare_saved(args)
    elif args.workload == "accuracy":
        out = run_accuracy(args)
    else:
        out = run_copy(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
