"""Depth-layered media on the GPU (bfpn_create_layered; no reference
counterpart: the reference solver is homogeneous).  Each depth step uses the
tables of the layer holding its midpoint; a paired energy sweep never spans
two layers, so the steps on either side of a boundary close and open with
single half steps.  Checked against:
- the homogeneous run, bit for bit, for a one-layer stack and for a stack of
  identical layers (boundaries inside a call, skipping on and off, both
  precisions);
- the layered oracle (pinned on the CPU by tests/test_oracle_layers.py):
  FP64 to 1e-10, FP32 to 2e-4 (its DST and source run on the matrix cores;
  the sign-changing proxy to 1e-3 of its maximum),
  for water/bone/water and water/air/water stacks, the eq15, legacy and
  lite-memory energy models, and through the CLI's --layer flag.
"""
import os
import subprocess

import numpy as np
import pytest

from bfpn_gpu_amd.abi import Case
from bfpn_gpu_amd.tables import material_dir
from conftest import ROOT

pytestmark = pytest.mark.gpu

WATER = ("water", 1.0, material_dir("water"))
BONE = ("bone", 1.85, material_dir("bone"))
AIR = ("air", 0.001205, material_dir("air"))


def case_of(**over):
    """100 MeV beam, 5x4 cells x 4x3 directions, 200 groups, dt 0.04 cm,
    every operator and the secondary active."""
    kw = dict(Nx=1000, Ny=5, Nz=4, Ng=200, Nmu=4, Nom=3, Ly=4, Lz=4, Lu=2, Lv=2, energy=100.0,
              sigma_yz=0.5, material="water", data=material_dir("water"), time=2.4,
              flags=dict(eq15_straggling=1, trapezoidal_yz=1))
    flags = over.pop("flags", None)
    kw.update(over)
    c = Case(**kw)
    if flags is not None:
        c.flags = flags
    return c


def run_gpu(gpu, case, precision, layers=None, chunks=None, skip=True):
    with gpu.Solver(case, precision, layers=layers) as s:
        s.set_group_skip(skip)
        for n in chunks or [case.num_steps()]:
            s.advance(n)
        idd, proxy = s.history()
        return idd, proxy, [s.state(w) for w in range(4)]


def assert_identical(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    for w in range(4):
        np.testing.assert_array_equal(a[2][w], b[2][w], err_msg=f"state {w}")


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
def test_one_layer_is_the_homogeneous_run(gpu, precision):
    case = case_of()
    p = getattr(gpu, precision)
    assert_identical(run_gpu(gpu, case, p, [(100.0, *WATER)]), run_gpu(gpu, case, p))


@pytest.mark.parametrize("precision,skip", [("FP32", True), ("FP32", False), ("FP64", True)])
def test_identical_layers_are_the_homogeneous_run(gpu, precision, skip):
    """Boundaries at 0.33, 0.71 and 1.5 cm force single energy halves at
    steps 8/9, 17/18 and 37/38 inside one advance() call (and inside the
    second of two calls); the paired and single sweeps agree bit for bit."""
    case = case_of()
    p = getattr(gpu, precision)
    stack = [(0.33, *WATER), (0.71, *WATER), (1.5, *WATER), (9.0, *WATER)]
    ref = run_gpu(gpu, case, p, skip=skip)
    assert_identical(run_gpu(gpu, case, p, stack, skip=skip), ref)
    assert_identical(run_gpu(gpu, case, p, stack, chunks=[13, case.num_steps() - 13], skip=skip),
                     ref)


STACKS = {
    "water-bone-water": [(0.6, *WATER), (1.4, *BONE), (9.0, *WATER)],
    "water-air-water": [(0.5, *WATER), (1.3, *AIR), (9.0, *WATER)],
    "bone-water": [(1.0, *BONE), (9.0, *WATER)],
}


@pytest.mark.parametrize("precision,tol", [("FP64", 1e-10), ("FP32", 2e-4)])
@pytest.mark.parametrize("stack", list(STACKS))
def test_layered_run_matches_oracle(gpu, oracle_mod, precision, tol, stack):
    case = case_of()
    n = case.num_steps()
    o = oracle_mod.Solver(case, layers=STACKS[stack])
    o.advance(n)
    g = run_gpu(gpu, case, getattr(gpu, precision), STACKS[stack], chunks=[7, n - 7])
    oi, op = o.history()
    np.testing.assert_allclose(g[0], oi, rtol=tol)
    # the proxy changes sign (in the air gap the FP32 state's rounding reaches
    # 5e-4 of its maximum): its error is measured against that maximum
    ptol = tol if precision == "FP64" else 1e-3
    np.testing.assert_allclose(g[1], op, rtol=ptol, atol=ptol * np.abs(op).max())
    scale = np.abs(o.state(0)).max()  # every state against the primary's scale
    for w in range(4):
        assert np.abs(g[2][w] - o.state(w)).max() / scale < tol, w


@pytest.mark.parametrize("flags", [{"legacy_energy": 1, "eq15_straggling": 1},
                                   {"lite_memory": 1}], ids=lambda f: next(iter(f)))
def test_layered_legacy_and_lite_match_oracle(gpu, oracle_mod, flags):
    case = case_of(flags=dict(trapezoidal_yz=1, **flags))
    stack = STACKS["water-bone-water"]
    o = oracle_mod.Solver(case, layers=stack)
    o.advance(case.num_steps())
    g = run_gpu(gpu, case, gpu.FP64, stack)
    np.testing.assert_allclose(g[0], o.history()[0], rtol=1e-10)


def test_cli_layers_write_the_oracle_values(tmp_path, gpu, oracle_mod):
    cli = os.path.join(ROOT, "bfpn_gpu_amd", "bin", "bfp_solver")
    case = case_of()
    w, b = material_dir("water"), material_dir("bone")
    cmd = [cli, str(case.Nx), "--data", case.data, "--material", "water", "--energy", "100",
           "--ny", str(case.Ny), "--nz", str(case.Nz), "--nmu", str(case.Nmu), "--nom",
           str(case.Nom), "--ly", "4", "--lz", "4", "--lu", "2", "--lv", "2", "--ng", "200",
           "--sigma-yz", "0.5", "--eq15-straggling", "--trapezoidal-yz", "--time", "2.4",
           "--precision", "fp64", "--layer", f"0.6:water:1.0:{w}", "--layer", f"1.4:bone:1.85:{b}",
           "--layer", f"9:water::{w}"]
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Layer: to x=1.4 cm, bone, rho=1.85" in r.stdout
    out = np.loadtxt(tmp_path / "idd_output.txt")
    o = oracle_mod.Solver(case, layers=STACKS["water-bone-water"])
    o.advance(case.num_steps())
    np.testing.assert_allclose(out[:, 1], o.history()[0], rtol=1e-10)
