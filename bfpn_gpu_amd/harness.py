"""Harness runners over the drop-in ``bfp_solver``: paper Table 2 cases and
depth-step convergence.

Both shell out to the executable exactly as the reference's scripts do, one
working directory per case (the solver writes ``idd_output.txt`` where it
runs), and evaluate the IDD files on the host:

* ``table2``: the six cases of scripts/run_table2.py:12-19 (water and bone at
  50, 100 and 230 MeV) with its command line (positional Nx, ``--time``,
  ``--material``, ``--energy``; ``--paper-grid`` adds the 80x80 space and
  20x20 angle grid, run_table2.py:62-71), the Bragg peak, P90, D90 and D20 of
  each IDD (:mod:`bfpn_gpu_amd.metrics`, restating scripts/analyze_idd.py) and
  their distance from the paper's deterministic column.
* ``grid_sensitivity``: scripts/idd_grid_sensitivity.py run against the
  binary: a baseline grid and variants that each change exactly one grid
  family (depth Nx, transverse Ny x Nz, angular Nmu x Nom), each IDD's
  entrance dose, Dmax, BP/P90/D90/D20, and the relative L2 / max-norm of its
  Dmax-normalised curve against the baseline's.
* ``convergence``: the depth-step refinement of
  scripts/convergence_order.py:127-196, 251-333: the same case at each Nx,
  every coarser IDD compared with the finest one on their common depth range
  (relative L2 by the trapezoid rule and relative max-norm, both over the
  finest curve interpolated to the coarse samples), and the observed order
  between neighbouring levels, log(e0/e1) / log(h0/h1) with h = Lx / Nx.

The material tables default to the ones bundled under ``bfpn_gpu_amd/tables``.
``python -m bfpn_gpu_amd.harness table2 --help`` / ``convergence --help`` /
``grid-sensitivity --help``.
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

from . import metrics

PKG = Path(__file__).resolve().parent
DEFAULT_EXE = PKG / "bin" / "bfp_solver"

# scripts/run_table2.py:12-19
TABLE2_CASES = (("water", 50), ("water", 100), ("water", 230),
                ("bone", 50), ("bone", 100), ("bone", 230))


def data_dir(material: str) -> str:
    """Bundled material tables (stopping power and cross sections)."""
    return str(PKG / "tables" / material)


def run_solver(exe, nx: int, args, case_dir: Path, timeout: float = 3600.0) -> Path:
    """Runs ``exe nx args...`` inside case_dir; returns its idd_output.txt."""
    case_dir.mkdir(parents=True, exist_ok=True)
    cmd = [str(exe), str(nx), *[str(a) for a in args]]
    (case_dir / "command.txt").write_text(" ".join(cmd) + "\n", encoding="utf-8")
    with open(case_dir / "run.log", "w", encoding="utf-8") as log:
        r = subprocess.run(cmd, cwd=case_dir, stdout=log, stderr=subprocess.STDOUT,
                           timeout=timeout)
    if r.returncode != 0:
        tail = (case_dir / "run.log").read_text(encoding="utf-8").splitlines()[-5:]
        raise RuntimeError(f"bfp_solver exited with {r.returncode}: " + " | ".join(tail))
    out = case_dir / "idd_output.txt"
    if not out.exists():
        raise FileNotFoundError(f"{out} was not produced")
    return out


def load_curve(path):
    """(x, value) columns of an IDD file, finite samples, sorted by x."""
    data = np.loadtxt(path, comments="#")
    if data.ndim != 2 or data.shape[1] < 2:
        raise ValueError(f"{path}: expected at least two columns (x value)")
    x, y = data[:, 0], data[:, 1]
    keep = np.isfinite(x) & np.isfinite(y)
    x, y = x[keep], y[keep]
    o = np.argsort(x, kind="stable")
    return x[o], y[o]


def _overlap(x, y, x_ref, y_ref):
    lo = max(float(np.min(x)), float(np.min(x_ref)))
    hi = min(float(np.max(x)), float(np.max(x_ref)))
    m = (x >= lo) & (x <= hi)
    if np.count_nonzero(m) < 2:
        raise ValueError("curves overlap on fewer than two samples")
    return x[m], y[m], np.interp(x[m], x_ref, y_ref)


def _trapezoid(f, x):
    if f.size < 2:
        return 0.0
    return float(np.sum(0.5 * (f[1:] + f[:-1]) * np.diff(x)))


def relative_l2(x, y, x_ref, y_ref) -> float:
    """||y - y_ref|| / ||y_ref|| in L2 over the common depth range."""
    xc, yc, yr = _overlap(x, y, x_ref, y_ref)
    den = _trapezoid(yr * yr, xc)
    if den <= 0:
        raise ValueError("reference curve has zero L2 norm")
    return math.sqrt(_trapezoid((yc - yr) ** 2, xc) / den)


def relative_linf(x, y, x_ref, y_ref) -> float:
    """max |y - y_ref| / max |y_ref| over the common depth range."""
    xc, yc, yr = _overlap(x, y, x_ref, y_ref)
    return float(np.max(np.abs(yc - yr)) / max(float(np.max(np.abs(yr))), 1e-300))


def observed_orders(rows) -> None:
    """rows[i]["order_to_next"] = log(e_i / e_{i+1}) / log(h_i / h_{i+1}) (NaN last)."""
    for a, b in zip(rows, rows[1:]):
        e0, e1, h0, h1 = a["rel_l2"], b["rel_l2"], a["h"], b["h"]
        a["order_to_next"] = (math.log(e0 / e1) / math.log(h0 / h1)
                              if e0 > 0 and e1 > 0 and h0 != h1 else math.nan)
    if rows:
        rows[-1]["order_to_next"] = math.nan


def table2(exe=DEFAULT_EXE, outdir="table2_runs", nx: int = 4000, time: float = 40.0,
           paper_grid: bool = False, extra=(), skip_run: bool = False, cases=TABLE2_CASES):
    """Runs (or re-reads) the Table 2 cases; per case the metrics in cm and their
    differences from the paper's deterministic column in mm."""
    out = []
    for material, energy in cases:
        case_dir = Path(outdir) / f"{material}_{energy}"
        if skip_run:
            idd = case_dir / "idd_output.txt"
        else:
            args = ["--time", time, "--material", material, "--energy", energy,
                    "--data", data_dir(material)]
            if paper_grid:
                args += ["--ny", 80, "--nz", 80, "--nmu", 20, "--nom", 20]
            idd = run_solver(exe, nx, [*args, *extra], case_dir)
        x, y = load_curve(idd)
        got = metrics.analyze(x, y)
        ref = metrics.TABLE2[(material, energy)]
        out.append({"material": material, "energy": energy, "metrics_cm": got,
                    "paper_cm": ref,
                    "diff_mm": {k: 10.0 * (got[k] - ref[k]) for k in ("BP", "P90", "D90", "D20")}})
    return out


def convergence(exe=DEFAULT_EXE, outdir="convergence", nxs=(500, 1000, 2000, 4000),
                time: float = 40.0, lx: float = 40.0, material: str = "water",
                energy: float = 230.0, ny: int = 20, nz: int = 20, ng: int = 500, nmu: int = 20,
                nom: int = 20, flags=(), extra=(), skip_run: bool = False):
    """Runs the case at each Nx (coarse to fine) and measures every coarser IDD
    against the finest; returns the rows with rel_l2, rel_linf, order_to_next."""
    nxs = list(nxs)
    if len(nxs) < 3 or nxs != sorted(nxs) or min(nxs) <= 0:
        raise ValueError("need at least three positive Nx values, coarse to fine")
    curves = {}
    for nx in nxs:
        case_dir = Path(outdir) / f"nx_{nx}"
        if skip_run:
            path = case_dir / "idd_output.txt"
        else:
            args = ["--time", time, "--material", material, "--energy", energy,
                    "--ny", ny, "--nz", nz, "--ng", ng, "--nmu", nmu, "--nom", nom,
                    "--energy-model", "eq15", "--data", data_dir(material), *flags, *extra]
            path = run_solver(exe, nx, args, case_dir)
        curves[nx] = load_curve(path)
    x_ref, y_ref = curves[nxs[-1]]
    rows = [{"nx": nx, "h": lx / nx,
             "rel_l2": relative_l2(*curves[nx], x_ref, y_ref),
             "rel_linf": relative_linf(*curves[nx], x_ref, y_ref)} for nx in nxs[:-1]]
    observed_orders(rows)
    return rows


# scripts/idd_grid_sensitivity.py:44-77: a grid is (Nx, Ny, Nz, Nmu, Nom) and
# a variant may change only the family it names
GRID_AXES = ("depth", "transverse", "angular")


def grid_families_changed(base, var) -> dict:
    """Which grid families differ between two (Nx, Ny, Nz, Nmu, Nom) grids."""
    return {"depth": var[0] != base[0],
            "transverse": (var[1], var[2]) != (base[1], base[2]),
            "angular": (var[3], var[4]) != (base[3], base[4])}


def check_variant(axis: str, label: str, base, var) -> None:
    """ValueError unless `var` changes the `axis` family of `base` and nothing else."""
    if axis not in GRID_AXES:
        raise ValueError(f"unknown variant axis: {axis}")
    if len(var) != 5 or min(var) <= 0:
        raise ValueError(f"{label}: grid sizes must be five positive integers")
    changed = grid_families_changed(base, var)
    if not changed[axis]:
        raise ValueError(f"{label}: the requested {axis} grid did not change")
    other = [a for a in GRID_AXES if a != axis and changed[a]]
    if other:
        raise ValueError(f"{label}: changes {axis} together with {', '.join(other)}")


def sensitivity_rows(baseline, variants):
    """Rows of scripts/idd_grid_sensitivity.py:119-145.  baseline = (label,
    grid, idd path); variants = [(axis, label, grid, idd path)]."""
    b_label, b_grid, b_path = baseline
    for axis, label, grid, _ in variants:
        check_variant(axis, label, b_grid, grid)

    def curve(path):
        x, dose = metrics.load_idd(path)
        got = metrics.analyze(x, dose)
        return got, x, dose / got["Dmax"], float(dose[0])

    _, bx, bnorm, _ = curve(b_path)
    rows = []
    for axis, label, grid, path in [("baseline", b_label, b_grid, b_path), *variants]:
        got, x, norm, entrance = curve(path)
        base = axis == "baseline"
        rows.append({"axis": axis, "label": label, "nx": grid[0], "ny": grid[1], "nz": grid[2],
                     "nmu": grid[3], "nom": grid[4], "entrance": entrance, "dmax": got["Dmax"],
                     "norm0": float(norm[0]), "bp": got["BP"], "p90": got["P90"],
                     "d90": got["D90"], "d20": got["D20"],
                     "norm_l2_vs_baseline": 0.0 if base else relative_l2(x, norm, bx, bnorm),
                     "norm_linf_vs_baseline": 0.0 if base else relative_linf(x, norm, bx, bnorm)})
    return rows


def sensitivity_markdown(rows) -> str:
    """The table of scripts/idd_grid_sensitivity.py:157-169."""
    lines = ["| axis | label | grid | norm0 | BP cm | normalized L2 | normalized Linf |",
             "|---|---|---|---:|---:|---:|---:|"]
    for r in rows:
        grid = f"{r['nx']} / {r['ny']}x{r['nz']} / {r['nmu']}x{r['nom']}"
        lines.append(f"| {r['axis']} | {r['label']} | {grid} | {r['norm0']:.9f} | {r['bp']:.4f} "
                     f"| {r['norm_l2_vs_baseline']:.6g} | {r['norm_linf_vs_baseline']:.6g} |")
    return "\n".join(lines) + "\n"


def grid_sensitivity(exe=DEFAULT_EXE, outdir="grid_sensitivity", baseline=None, variants=(),
                     time: float = 40.0, material: str = "water", energy: float = 230.0,
                     extra=(), skip_run: bool = False):
    """Runs the baseline grid ("label", (Nx, Ny, Nz, Nmu, Nom)) and each variant
    (axis, label, grid) through the solver, one directory per label (the
    variants are checked before anything runs), and returns the rows of
    sensitivity_rows; writes grid_sensitivity.csv and .md into outdir."""
    if baseline is None:
        raise ValueError("a baseline grid is required")
    b_label, b_grid = baseline[0], tuple(int(v) for v in baseline[1])
    if len(b_grid) != 5 or min(b_grid) <= 0:
        raise ValueError(f"{b_label}: grid sizes must be five positive integers")
    variants = [(a, l, tuple(int(v) for v in g)) for a, l, g in variants]
    labels = [b_label] + [l for _, l, _ in variants]
    if len(set(labels)) != len(labels):
        raise ValueError("case labels must be unique")
    for axis, label, grid in variants:
        check_variant(axis, label, b_grid, grid)
    paths = {}
    for label, grid in [(b_label, b_grid)] + [(l, g) for _, l, g in variants]:
        case_dir = Path(outdir) / label
        if skip_run:
            paths[label] = case_dir / "idd_output.txt"
            continue
        nx, ny, nz, nmu, nom = grid
        args = ["--time", time, "--material", material, "--energy", energy,
                "--ny", ny, "--nz", nz, "--nmu", nmu, "--nom", nom, "--data", data_dir(material),
                *extra]
        paths[label] = run_solver(exe, nx, args, case_dir)
    rows = sensitivity_rows((b_label, b_grid, paths[b_label]),
                            [(a, l, g, paths[l]) for a, l, g in variants])
    out = Path(outdir)
    out.mkdir(parents=True, exist_ok=True)
    with open(out / "grid_sensitivity.csv", "w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0]), lineterminator="\n")
        w.writeheader()
        w.writerows(rows)
    (out / "grid_sensitivity.md").write_text(sensitivity_markdown(rows), encoding="utf-8")
    return rows


def _write_csv(path: Path, rows) -> None:
    with open(path, "w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, fieldnames=["nx", "h", "rel_l2", "rel_linf", "order_to_next"])
        w.writeheader()
        w.writerows(rows)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m bfpn_gpu_amd.harness")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("table2", help="paper Table 2 cases")
    t.add_argument("--exe", default=str(DEFAULT_EXE))
    t.add_argument("--nx", type=int, default=4000)
    t.add_argument("--time", type=float, default=40.0)
    t.add_argument("--outdir", default="table2_runs")
    t.add_argument("--paper-grid", action="store_true")
    t.add_argument("--materials", default="water,bone", help="subset of the cases, e.g. water")
    t.add_argument("--skip-run", action="store_true")
    t.add_argument("extra", nargs=argparse.REMAINDER,
                   help="further solver arguments after --, e.g. -- --ly 8 --lz 8 --sigma-yz 0.3")
    s = sub.add_parser("grid-sensitivity", help="one grid family changed at a time")
    s.add_argument("--exe", default=str(DEFAULT_EXE))
    s.add_argument("--outdir", default="grid_sensitivity")
    s.add_argument("--baseline", nargs=6, required=True,
                   metavar=("LABEL", "NX", "NY", "NZ", "NMU", "NOM"))
    s.add_argument("--variant", nargs=7, action="append", required=True,
                   metavar=("AXIS", "LABEL", "NX", "NY", "NZ", "NMU", "NOM"),
                   help="AXIS is depth, transverse or angular")
    s.add_argument("--time", type=float, default=40.0)
    s.add_argument("--material", default="water", choices=["water", "bone", "air"])
    s.add_argument("--energy", type=float, default=230.0)
    s.add_argument("--skip-run", action="store_true")
    s.add_argument("extra", nargs=argparse.REMAINDER,
                   help="further solver arguments after --, e.g. -- --ly 8 --lz 8 --ng 300")
    c = sub.add_parser("convergence", help="depth-step convergence and observed order")
    c.add_argument("--exe", default=str(DEFAULT_EXE))
    c.add_argument("--out", default="convergence")
    c.add_argument("--nxs", default="500,1000,2000,4000")
    c.add_argument("--time", type=float, default=40.0)
    c.add_argument("--lx", type=float, default=40.0)
    c.add_argument("--material", default="water", choices=["water", "bone", "air"])
    c.add_argument("--energy", type=float, default=230.0)
    for k, v in (("--ny", 20), ("--nz", 20), ("--ng", 500), ("--nmu", 20), ("--nom", 20)):
        c.add_argument(k, type=int, default=v)
    for k in ("--primary-only", "--eq15-straggling", "--energy-only", "--no-transport",
              "--no-angle", "--no-spatial-clipping"):
        c.add_argument(k, action="store_true")
    c.add_argument("--skip-run", action="store_true")
    c.add_argument("extra", nargs=argparse.REMAINDER,
                   help="further solver arguments after --, e.g. -- --sigma-yz 0.3")
    a = ap.parse_args(argv)
    extra = [e for e in getattr(a, "extra", []) if e != "--"]
    if a.cmd == "grid-sensitivity":
        def grid(vals):
            try:
                return tuple(int(v) for v in vals)
            except ValueError:
                ap.error(f"grid sizes must be integers: {' '.join(vals)}")
        variants = [(v[0], v[1], grid(v[2:])) for v in a.variant]
        try:
            rows = grid_sensitivity(a.exe, a.outdir, (a.baseline[0], grid(a.baseline[1:])),
                                    variants, a.time, a.material, a.energy, extra, a.skip_run)
        except ValueError as exc:
            ap.error(str(exc))
        print(sensitivity_markdown(rows), end="")
        return 0
    if a.cmd == "table2":
        mats = set(a.materials.split(","))
        rows = table2(a.exe, a.outdir, a.nx, a.time, a.paper_grid, extra, a.skip_run,
                      cases=tuple(c for c in TABLE2_CASES if c[0] in mats))
        for r in rows:
            d = r["diff_mm"]
            print(f"{r['material']:5s} {r['energy']:3d} MeV: " + "  ".join(
                f"{k} {r['metrics_cm'][k]:.3f} cm ({d[k]:+.2f} mm)" for k in ("BP", "P90", "D90", "D20")))
        Path(a.outdir).mkdir(parents=True, exist_ok=True)
        (Path(a.outdir) / "summary.json").write_text(json.dumps(rows, indent=1), encoding="utf-8")
        return 0
    flags = [f"--{k.replace('_', '-')}" for k in ("primary_only", "eq15_straggling", "energy_only",
                                                  "no_transport", "no_angle", "no_spatial_clipping")
             if getattr(a, k)]
    rows = convergence(a.exe, a.out, [int(v) for v in a.nxs.split(",") if v], a.time, a.lx,
                       a.material, a.energy, a.ny, a.nz, a.ng, a.nmu, a.nom, flags, extra,
                       a.skip_run)
    Path(a.out).mkdir(parents=True, exist_ok=True)
    _write_csv(Path(a.out) / "convergence.csv", rows)
    print("nx,h,rel_l2,rel_linf,order_to_next")
    for r in rows:
        o = r["order_to_next"]
        print(f"{r['nx']},{r['h']:.6g},{r['rel_l2']:.6e},{r['rel_linf']:.6e},"
              f"{'nan' if math.isnan(o) else f'{o:.6g}'}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
