"""Pins the CPU FP64 oracle on the reference's own golden IDD files.

Fixtures (copied verbatim from the reference, data only):
  tests/golden/stage6_water230/8_7_nx1000/idd_output.txt    <- results/stage6_final_water230/8_7_nx1000
  tests/golden/stage6_water230/12_9_nx1000/idd_output.txt   <- results/stage6_final_water230/12_9_nx1000
Settings (results/stage6_final_water230/summary.md:5-8): water 230 MeV,
Nx=1000, Ly=Lz=8, Lu=Lv=2, sigma_yz=0.3, discrete angular delta, normalised
initial mass, trapezoidal y/z weights, fitted sigma_c, full Eq.24 dose.  The
unstated settings were fixed by this test: Ng=500 and straggling OFF
reproduce the file to ~1e-15 (straggling ON misses by ~1e-5).

The first steps exercise every operator of the depth march: energy, MUSCL
transport, DST angular diffusion (step 0), and from step 1 on the secondary
source and the secondary energy/transport/angle chain, plus the IDD.

  tests/golden/grid_sensitivity.csv  <- results/stage3_grid_sensitivity/grid_sensitivity.csv
Four water 230 MeV runs (Nx 500/1000, 8x8/12x12 cells, 7x7/9x9 directions)
whose settings the reference does not state.  The absolute entrance value
(the IDD after the first depth step) of every row is reproduced to 1e-14 by
Ly=Lz=8, Lu=Lv=2, sigma_yz=0.3, normalised initial mass, the legacy
cross-section calibration and the stopping-power-only dose, straggling off;
flipping either of the last two, or straggling on, moves it by 1e-8 or more.  This pins the
oracle's calibration (src/bfpm_solver.cu:1488-1493) and stopping-power-only
dose branches; the full runs (Dmax, BP, P90, D90, D20, normalised L2/Linf)
are reproduced on the GPU in tests/test_gpu_harness.py.
"""
import csv
import os

import numpy as np
import pytest

from bfpn_gpu_amd.abi import Case
from conftest import GOLDEN, TABLES


def golden_case(tag):
    ny, nmu = {"8_7": (8, 7), "12_9": (12, 9)}[tag]
    return Case(Nx=1000, Ny=ny, Nz=ny, Ng=500, Nmu=nmu, Nom=nmu, Ly=8, Lz=8, Lu=2, Lv=2,
                energy=230, sigma_yz=0.3, material="water", data=os.path.join(TABLES, "water"),
                time=40, flags=dict(discrete_angle_delta=1, normalize_initial_mass=1,
                                    trapezoidal_yz=1))


def golden_idd(tag):
    return np.loadtxt(os.path.join(GOLDEN, "stage6_water230", f"{tag}_nx1000", "idd_output.txt"))


@pytest.mark.parametrize("tag,steps", [("8_7", 3), ("12_9", 2)])
def test_oracle_reproduces_reference_golden_idd(oracle_mod, tag, steps):
    s = oracle_mod.Solver(golden_case(tag))
    s.advance(steps)
    idd, _ = s.history()
    gold = golden_idd(tag)
    np.testing.assert_allclose(gold[:steps, 0], 0.04 * np.arange(1, steps + 1), rtol=1e-15)
    rel = np.abs(idd - gold[:steps, 1]) / gold[:steps, 1]
    assert rel.max() < 1e-13, rel


def test_straggling_flag_is_off_in_the_fixture(oracle_mod):
    """With --eq15-straggling the first IDD sample moves by ~7e-6: the fixture
    was produced without it, which locks the flag set used everywhere."""
    c = golden_case("8_7")
    c.flags["eq15_straggling"] = 1
    s = oracle_mod.Solver(c)
    s.advance(1)
    idd, _ = s.history()
    gold = golden_idd("8_7")
    assert abs(idd[0] - gold[0, 1]) / gold[0, 1] > 1e-6


def test_oracle_tables_load_like_reference(oracle_mod):
    """Water has 499 explicit rows (the 500th is zero-filled); bone reads the
    first Ng rows of a 1000-row CRLF file (src/bfpm_solver.cu:1455-1466)."""
    mu, sig, rows = oracle_mod.load_tables(os.path.join(TABLES, "water"), 500)
    assert rows == 499 and np.all(mu[499] == 0) and sig.shape == (500,)
    mu, sig, rows = oracle_mod.load_tables(os.path.join(TABLES, "bone"), 500)
    assert rows == 500
    with pytest.raises(OSError, match="need at least"):
        oracle_mod.load_tables(os.path.join(TABLES, "water"), 502)
    with pytest.raises(OSError, match="Cannot open"):
        oracle_mod.load_tables("/nonexistent/dir", 10)


def stage3_rows():
    with open(os.path.join(GOLDEN, "grid_sensitivity.csv"), newline="") as f:
        return list(csv.DictReader(f))


def stage3_case(nx, ny, nmu, **flags):
    f = dict(normalize_initial_mass=1, calibrate_cross_sections=1, dose_catastrophic_loss=0)
    f.update(flags)
    return Case(Nx=nx, Ny=ny, Nz=ny, Ng=500, Nmu=nmu, Nom=nmu, Ly=8, Lz=8, Lu=2, Lv=2,
                energy=230, sigma_yz=0.3, material="water", data=os.path.join(TABLES, "water"),
                time=40, flags=f)


def _entrance(oracle_mod, case):
    s = oracle_mod.Solver(case)
    s.advance(1)
    return s.history()[0][0]


@pytest.mark.parametrize("row", range(4))
def test_oracle_reproduces_stage3_entrance(oracle_mod, row):
    r = stage3_rows()[row]
    case = stage3_case(int(r["nx"]), int(r["ny"]), int(r["nmu"]))
    assert int(r["nz"]) == case.Nz and int(r["nom"]) == case.Nom
    got = _entrance(oracle_mod, case)
    assert abs(got - float(r["entrance"])) / float(r["entrance"]) < 1e-14, (got, r["entrance"])


@pytest.mark.parametrize("flip", [dict(calibrate_cross_sections=0),
                                  dict(dose_catastrophic_loss=1),
                                  dict(eq15_straggling=1)], ids=lambda f: next(iter(f)))
def test_stage3_flags_are_the_ones_that_reproduce_it(oracle_mod, flip):
    r = stage3_rows()[0]
    got = _entrance(oracle_mod, stage3_case(500, 8, 7, **flip))
    assert abs(got - float(r["entrance"])) / float(r["entrance"]) > 1e-9
