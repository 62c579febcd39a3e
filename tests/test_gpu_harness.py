"""The harness runners driving the real bfp_solver on the GPU (small case)."""
import math
import os

import pytest

from bfpn_gpu_amd import harness

pytestmark = pytest.mark.gpu


def test_convergence_runner_on_the_solver(tmp_path, gpu):
    """Depth-step refinement of a small water case through the CLI: every
    coarser IDD is closer to the finest as Nx grows, with a positive observed
    order (the measured orders on the full single-beam case are stated with
    the design's measurements)."""
    assert os.path.exists(harness.DEFAULT_EXE)
    rows = harness.convergence(harness.DEFAULT_EXE, tmp_path, nxs=(200, 400, 800, 1600),
                               time=8.0, lx=8.0, material="water", energy=100, ny=4, nz=4,
                               ng=100, nmu=4, nom=4, flags=("--eq15-straggling",),
                               extra=("--ly", "4", "--lz", "4"))
    errs = [r["rel_l2"] for r in rows]
    assert all(math.isfinite(e) and e > 0 for e in errs), errs
    assert errs[0] > errs[1] > errs[2], errs
    assert all(r["order_to_next"] > 0.5 for r in rows[:-1]), rows


def test_table2_runner_on_one_case(tmp_path, gpu):
    """One Table 2 case at a coarse depth grid: the runner produces all four
    metrics and their distance from the paper's values."""
    rows = harness.table2(harness.DEFAULT_EXE, tmp_path, nx=500, time=10.0,
                          cases=(("water", 100),))
    (r,) = rows
    assert all(math.isfinite(r["metrics_cm"][k]) for k in ("BP", "P90", "D90", "D20")), r
    assert 6.0 < r["metrics_cm"]["BP"] < 9.0, r


def test_grid_sensitivity_runner_on_the_solver(tmp_path, gpu):
    """A small water case with each grid family refined on its own through
    the CLI: every variant's normalised curve stays close to the baseline's
    and keeps its Bragg peak, and the baseline row is exactly zero."""
    rows = harness.grid_sensitivity(
        harness.DEFAULT_EXE, tmp_path, ("base", (400, 4, 4, 4, 4)),
        [("depth", "fine_dx", (800, 4, 4, 4, 4)),
         ("transverse", "fine_yz", (400, 8, 8, 4, 4)),
         ("angular", "fine_angle", (400, 4, 4, 8, 8))],
        time=10.0, material="water", energy=100,  # 10 cm: past the distal 20 % point
        extra=("--ng", "100", "--ly", "4", "--lz", "4", "--eq15-straggling"))
    base = rows[0]
    assert base["norm_l2_vs_baseline"] == 0.0
    for r in rows:
        assert all(math.isfinite(r[k]) for k in ("entrance", "dmax", "bp", "d90", "d20")), r
    for r in rows[1:]:
        assert 0.0 < r["norm_l2_vs_baseline"] < 0.2, r
        assert abs(r["bp"] - base["bp"]) < 0.5, r
    assert (tmp_path / "grid_sensitivity.md").exists()


def test_grid_sensitivity_reproduces_the_reference_stage3_csv(tmp_path, gpu):
    """The reference's grid-sensitivity study (results/stage3_grid_sensitivity:
    water 230 MeV, baseline Nx 500 / 8x8 cells / 7x7 directions, one family
    refined per variant) through the CLI in FP64 with the settings that
    reproduce its entrance values (tests/test_oracle_golden.py): every column
    of every row equals the committed CSV -- entrance, Dmax and norm0 to 1e-12,
    BP/P90/D90/D20 to 1e-10 cm, the normalised L2/Linf differences to 1e-9."""
    import csv

    from conftest import GOLDEN

    rows = harness.grid_sensitivity(
        harness.DEFAULT_EXE, tmp_path, ("base", (500, 8, 8, 7, 7)),
        [("depth", "nx1000", (1000, 8, 8, 7, 7)),
         ("transverse", "yz12", (500, 12, 12, 7, 7)),
         ("angular", "uv9", (500, 8, 8, 9, 9))],
        time=40.0, material="water", energy=230,
        extra=("--ly", "8", "--lz", "8", "--lu", "2", "--lv", "2", "--sigma-yz", "0.3",
               "--normalize-initial-mass", "--legacy-cross-section-calibration",
               "--stopping-power-dose-only", "--precision", "fp64"))
    with open(os.path.join(GOLDEN, "grid_sensitivity.csv"), newline="") as f:
        gold = list(csv.DictReader(f))
    assert [r["label"] for r in rows] == [g["label"] for g in gold]
    for r, g in zip(rows, gold):
        for k in ("nx", "ny", "nz", "nmu", "nom"):
            assert r[k] == int(g[k])
        for k in ("entrance", "dmax", "norm0"):
            assert abs(r[k] - float(g[k])) <= 1e-12 * abs(float(g[k])), (r["label"], k, r[k], g[k])
        for k in ("bp", "p90", "d90", "d20"):
            assert abs(r[k] - float(g[k])) <= 1e-10, (r["label"], k, r[k], g[k])
        for k in ("norm_l2_vs_baseline", "norm_linf_vs_baseline"):
            assert abs(r[k] - float(g[k])) <= 1e-9, (r["label"], k, r[k], g[k])
