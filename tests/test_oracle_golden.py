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
"""
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
