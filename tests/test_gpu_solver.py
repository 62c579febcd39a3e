"""Whole depth march on the GPU through the C ABI vs the CPU FP64 oracle,
and against the reference's own golden IDD files.

Tolerances:
  FP64 state vs oracle (same steps):  states and IDD within 1e-10 relative to
      the array maximum (the only differences are FMA contraction and
      summation order, compounded over the steps);
  FP64 full golden run (1000 steps):  IDD within 1e-8 relative pointwise where
      IDD >= 1e-6 of its max, metrics within 1e-4 cm;
  FP32 state (the accuracy gate):     IDD within 0.5 % pointwise where
      IDD >= 10 % of max; BP, D90, D20 within 0.01 cm (0.1 mm).
"""
import os
import subprocess

import numpy as np
import pytest

from bfpn_gpu_amd import metrics
from conftest import GOLDEN, ROOT
from gpu_cases import golden_case, small_case

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _run_both(gpu, oracle_mod, case, steps, precision):
    o = oracle_mod.Solver(case)
    o.advance(steps)
    g = gpu.Solver(case, precision)
    g.advance(steps)
    return o, g


@pytest.mark.parametrize("flags", [
    {},
    {"eq15_straggling": 1},
    {"primary_only": 1},
    {"energy_only": 1},
    {"no_transport": 1, "no_angle": 1},
    {"no_spatial_clipping": 1, "dose_catastrophic_loss": 0},
    {"calibrate_cross_sections": 1, "trapezoidal_yz": 0, "discrete_angle_delta": 0,
     "normalize_initial_mass": 0},
], ids=lambda f: "-".join(sorted(f)) or "default")
def test_fp64_solver_matches_oracle(gpu, oracle_mod, flags):
    case = small_case(flags=flags)
    o, g = _run_both(gpu, oracle_mod, case, 4, gpu.FP64)
    for k in range(4):
        assert _rel(g.state(k), o.state(k)) < 1e-10 or np.abs(o.state(k)).max() == 0, k
    gi, gp = g.history()
    oi, op = o.history()
    assert len(gi) == 4
    np.testing.assert_allclose(gi, oi, rtol=1e-11)
    np.testing.assert_allclose(gp, op, rtol=1e-10, atol=1e-12 * np.abs(op).max())


@pytest.mark.parametrize("material,rho", [("bone", 1.85), ("air", None)])
def test_fp64_materials(gpu, oracle_mod, material, rho):
    from conftest import TABLES

    case = small_case(material=material, density=rho, data=os.path.join(TABLES, material),
                      energy=150.0, Ny=3, Nz=3)
    o, g = _run_both(gpu, oracle_mod, case, 3, gpu.FP64)
    np.testing.assert_allclose(g.history()[0], o.history()[0], rtol=1e-11)
    for k in range(4):
        assert _rel(g.state(k), o.state(k)) < 1e-10 or np.abs(o.state(k)).max() == 0


@pytest.mark.parametrize("precision,tol", [("FP64", 1e-10), ("FP32", 2e-4)])
@pytest.mark.parametrize("flags", [
    {"legacy_energy": 1, "eq15_straggling": 1},
    {"legacy_energy": 1},
    {"legacy_energy": 1, "primary_only": 1, "eq15_straggling": 1},
    {"lite_memory": 1},
    {"lite_memory": 1, "eq15_straggling": 1, "trapezoidal_yz": 0},
], ids=lambda f: "-".join(sorted(f)))
def test_legacy_energy_model_and_lite_memory_match_oracle(gpu, oracle_mod, flags, precision, tol):
    """--energy-model legacy (src/energy_solver.cu:88-106) and --lite-memory
    (src/bfpm_solver.cu:2191-2212): FP64 to 1e-10, FP32 state to 2e-4 (its
    DST and source run on the matrix cores); lite keeps psi1 only, reports
    zero proxies and ignores straggling."""
    case = small_case(flags=flags)
    o, g = _run_both(gpu, oracle_mod, case, 5, getattr(gpu, precision))
    scale = np.abs(o.state(0)).max()
    for k in range(4):
        assert np.abs(g.state(k) - o.state(k)).max() / scale < tol, k
    gi, gp = g.history()
    oi, op = o.history()
    assert len(gi) == 5
    np.testing.assert_allclose(gi, oi, rtol=tol)
    np.testing.assert_allclose(gp, op, rtol=tol, atol=tol * np.abs(op).max())
    if flags.get("lite_memory"):
        assert np.all(gp == 0) and all(np.all(g.state(k) == 0) for k in (1, 2, 3))


def test_tables_bit_identical_to_oracle(gpu, oracle_mod):
    case = small_case()
    o = oracle_mod.Solver(case)
    g = gpu.Solver(case, gpu.FP32)
    for name in ("S", "sigma", "T", "sig_tr", "en", "ker_e1", "ker_e2", "ker_v", "omend", "mu",
                 "ker_beg", "ker_end"):
        np.testing.assert_array_equal(g.table(name), o.table(name), err_msg=name)
    F = g.state(0)
    np.testing.assert_allclose(F, o.state(0).astype(np.float32), rtol=0, atol=0)


def test_fp32_solver_tracks_oracle(gpu, oracle_mod):
    """FP32 state after 6 steps: IDD within 2e-5; primary psi^1 within 1e-4 of
    its maximum; secondary psi^1 within 5e-4 (its source runs in split-BF16,
    ~2^-16 per product, and feeds every secondary step)."""
    case = small_case()
    o, g = _run_both(gpu, oracle_mod, case, 6, gpu.FP32)
    np.testing.assert_allclose(g.history()[0], o.history()[0], rtol=2e-5)
    assert _rel(g.state(0), o.state(0)) < 1e-4
    assert _rel(g.state(2), o.state(2)) < 5e-4


def test_fp32_wide_grid_tracks_oracle(gpu, oracle_mod):
    """128x128 nodes x 4x4 directions (2^18 lanes, the four-lanes-per-thread
    energy sweep with the fused dose sum), 40 groups: IDD and states after 3
    steps within the FP32 bounds of the small-grid test."""
    case = small_case(Ny=128, Nz=128, Nmu=4, Nom=4, Ng=40, Ly=8, Lz=8)
    o, g = _run_both(gpu, oracle_mod, case, 3, gpu.FP32)
    np.testing.assert_allclose(g.history()[0], o.history()[0], rtol=2e-5)
    assert _rel(g.state(0), o.state(0)) < 1e-4
    assert _rel(g.state(2), o.state(2)) < 5e-4


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
def test_overlap_gives_identical_results(gpu, precision):
    """The secondary march on its own stream (default) and every operator on
    one stream produce the same bits: only the schedule differs."""
    case = small_case(flags={"eq15_straggling": 1})
    out = []
    for overlap in (True, False):
        with gpu.Solver(case, getattr(gpu, precision)) as s:
            s.set_overlap(overlap)
            s.advance(3)
            s.set_overlap(not overlap)  # switching between calls is allowed
            s.advance(3)
            out.append((s.history(), [s.state(k) for k in range(4)]))
    (h1, st1), (h2, st2) = out
    np.testing.assert_array_equal(h1[0], h2[0])
    np.testing.assert_array_equal(h1[1], h2[1])
    for a, b in zip(st1, st2):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
@pytest.mark.parametrize("flags", [{"eq15_straggling": 1}, {}, {"energy_only": 1},
                                   {"primary_only": 1, "eq15_straggling": 1}],
                         ids=lambda f: "-".join(sorted(f)) or "default")
def test_paired_energy_sweeps_are_bit_identical(gpu, precision, flags):
    """Inside one advance() call the end half of step n and the start half of
    step n+1 run as one paired sweep; one step per call runs every half step
    alone.  States, IDD and proxy agree bit for bit."""
    case = small_case(Ny=3, Nz=4, flags=flags)
    prec = getattr(gpu, precision)
    with gpu.Solver(case, prec) as one, gpu.Solver(case, prec) as many:
        for _ in range(5):
            one.advance(1)
        many.advance(5)
        for a, b in zip(one.history(), many.history()):
            np.testing.assert_array_equal(a, b)
        for k in range(4):
            np.testing.assert_array_equal(one.state(k), many.state(k))


@pytest.mark.parametrize("flags", [{"eq15_straggling": 1}, {"primary_only": 1}],
                         ids=lambda f: "-".join(sorted(f)))
def test_paired_energy_sweeps_are_bit_identical_wide_lanes(gpu, flags):
    """2^18 lanes (32 x 32 nodes x 16 x 16 directions): the FP32 sweeps run
    four adjacent lanes per thread, paired and alone; bit for bit the same."""
    case = small_case(Ny=31, Nz=31, Nmu=16, Nom=16, Ng=40, flags=flags)
    assert (case.Ny + 1) * (case.Nz + 1) * case.Nmu * case.Nom == 1 << 18
    with gpu.Solver(case, gpu.FP32) as one, gpu.Solver(case, gpu.FP32) as many:
        for _ in range(3):
            one.advance(1)
        many.advance(3)
        for a, b in zip(one.history(), many.history()):
            np.testing.assert_array_equal(a, b)
        for k in range(4):
            np.testing.assert_array_equal(one.state(k), many.state(k))


def test_history_and_restart_are_consistent(gpu):
    """advance(a)+advance(b) equals advance(a+b); write_state round-trips."""
    case = small_case(Ny=3, Nz=3)
    a = gpu.Solver(case, gpu.FP64)
    a.advance(2)
    a.advance(3)
    b = gpu.Solver(case, gpu.FP64)
    b.advance(5)
    np.testing.assert_array_equal(a.history()[0], b.history()[0])
    st = b.state(2)
    b.set_state(2, st)
    np.testing.assert_array_equal(b.state(2), st)


def test_invalid_data_path_is_an_error(gpu):
    case = small_case(data="/nonexistent")
    with pytest.raises(gpu.BfpnError, match="Cannot open scattering data"):
        gpu.Solver(case, gpu.FP32)


# ---------------------------------------------------------------- golden ---
@pytest.fixture(scope="module")
def golden():
    return np.loadtxt(os.path.join(GOLDEN, "stage6_water230", "8_7_nx1000", "idd_output.txt"))


def test_fp64_full_run_reproduces_reference_golden(gpu, golden):
    g = gpu.Solver(golden_case("8_7"), gpu.FP64)
    idd, _ = g.run()
    assert len(idd) == 1000
    ref = golden[:, 1]
    m = ref >= 1e-6 * ref.max()
    rel = np.abs(idd[m] - ref[m]) / ref[m]
    assert rel.max() < 1e-8, rel.max()
    mg, mr = metrics.analyze(golden[:, 0], idd), metrics.analyze(golden[:, 0], ref)
    for k in ("BP", "P90", "D90", "D20"):
        assert abs(mg[k] - mr[k]) < 1e-4


def test_fp32_full_run_meets_accuracy_gate(gpu, golden):
    g = gpu.Solver(golden_case("8_7"), gpu.FP32)
    idd, _ = g.run()
    ref = golden[:, 1]
    cmp = metrics.compare(idd, ref, band=0.10)
    assert cmp["max_rel"] <= 5e-3, cmp
    mg, mr = metrics.analyze(golden[:, 0], idd), metrics.analyze(golden[:, 0], ref)
    for k in ("BP", "D90", "D20"):
        assert abs(mg[k] - mr[k]) <= 0.01, (k, mg[k], mr[k])


def test_fp32_long_run_tracks_fp64(gpu):
    """800 depth steps of a 230 MeV beam with 20x20 directions: the FP32-state
    IDD stays within 1e-3 of the FP64 instantiation everywhere above 10 % of
    its maximum (compounding operator errors would show as a steady drift)."""
    case = small_case(Ny=4, Nz=4, Nmu=20, Nom=20, Ng=200, Nx=2000, time=16.0, Ly=4, Lz=4,
                      flags={"eq15_straggling": 1})
    with gpu.Solver(case, gpu.FP32) as a, gpu.Solver(case, gpu.FP64) as b:
        ia, ib = a.run()[0], b.run()[0]
    assert len(ia) == 800
    cmp = metrics.compare(ia, ib, band=0.10)
    assert cmp["max_rel"] <= 1e-3, cmp


def test_cli_fp64_writes_the_reference_file(tmp_path, golden, gpu):
    """bfp_solver with the fixture's flags writes idd_output.txt in the
    reference's format (17 digits, x = (i+1) dt) with the golden values."""
    cli = os.path.join(ROOT, "bfpn_gpu_amd", "bin", "bfp_solver")
    data = os.path.join(ROOT, "bfpn_gpu_amd", "tables", "water")
    cmd = [cli, "1000", "--data", data, "--material", "water", "--energy", "230", "--ny", "8",
           "--nz", "8", "--nmu", "7", "--nom", "7", "--ly", "8", "--lz", "8", "--lu", "2", "--lv", "2",
           "--sigma-yz", "0.3", "--discrete-angle-delta", "--normalize-initial-mass",
           "--trapezoidal-yz", "--time", "40", "--precision", "fp64"]
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Step 1000, t = 40" in r.stdout
    lines = open(tmp_path / "idd_output.txt").read().splitlines()
    assert lines[0] == "# x_cm value" and len(lines) == 1001
    ref_lines = open(os.path.join(GOLDEN, "stage6_water230", "8_7_nx1000",
                                  "idd_output.txt")).read().splitlines()
    assert [ln.split()[0] for ln in lines] == [ln.split()[0] for ln in ref_lines]
    out = np.loadtxt(tmp_path / "idd_output.txt")
    m = golden[:, 1] >= 1e-6 * golden[:, 1].max()
    assert np.max(np.abs(out[m, 1] - golden[m, 1]) / golden[m, 1]) < 1e-8
    assert os.path.exists(tmp_path / "dose_output.txt")


@pytest.mark.parametrize("extra", [["--lite-memory"], ["--energy-model", "legacy",
                                                      "--eq15-straggling"]])
def test_cli_legacy_and_lite_runs(tmp_path, gpu, oracle_mod, extra):
    """The CLI accepts --lite-memory and --energy-model legacy and writes the
    reference's files with the oracle's values."""
    cli = os.path.join(ROOT, "bfpn_gpu_amd", "bin", "bfp_solver")
    case = small_case(Nx=200, time=1.0)
    cmd = [cli, "200", "--data", case.data, "--material", "water", "--energy", "230",
           "--ny", str(case.Ny), "--nz", str(case.Nz), "--nmu", str(case.Nmu), "--nom",
           str(case.Nom), "--ly", str(case.Ly), "--lz", str(case.Lz), "--lu", "2", "--lv", "2",
           "--sigma-yz", str(case.sigma_yz), "--discrete-angle-delta", "--normalize-initial-mass",
           "--trapezoidal-yz", "--time", "1.0", "--precision", "fp64", *extra]
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = np.loadtxt(tmp_path / "idd_output.txt")
    proxy = np.loadtxt(tmp_path / "dose_output.txt")
    assert out.shape == (5, 2)
    flags = {"lite_memory": 1} if extra[0] == "--lite-memory" else {"legacy_energy": 1,
                                                                     "eq15_straggling": 1}
    case.flags.update(flags)
    o = oracle_mod.Solver(case)
    o.advance(5)
    np.testing.assert_allclose(out[:, 1], o.history()[0], rtol=1e-10)
    if extra[0] == "--lite-memory":
        assert np.all(proxy[:, 1] == 0)


# ----------------------------------------------------------- diagnostics ---
@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_energy_moments_and_spot_plane_match_oracle(gpu, oracle_mod, precision):
    """Eq.25 moments and the Y-Z dose plane of the GPU's own state after 3
    steps (primary and secondary) equal the oracle's functions on that state
    (1e-12: FP64 sums of the same values in another order); the plane
    integrated with the trapezoid weights is the step's IDD (FP32: to the
    rounding of the stored state, which the IDD does not see)."""
    case = small_case(flags={"eq15_straggling": 1})
    g = gpu.Solver(case, getattr(gpu, precision))
    g.advance(3)
    st = [g.state(k) for k in range(4)]
    d = oracle_mod.deltas(case)
    nyz = (case.Ny + 1) * (case.Nz + 1)
    nang = case.Nmu * case.Nom
    m_ref = oracle_mod.energy_moments(*st, case.Ng, nyz, nang, d.du, d.dv, d.dy, d.dz)
    m = g.energy_moments()
    assert m.shape == (case.Ng, 2)
    np.testing.assert_allclose(m, m_ref, rtol=1e-12, atol=1e-12 * np.abs(m_ref).max())
    S, sig, en = (g.table(k) for k in ("S", "sigma", "en"))
    p_ref = oracle_mod.spot_plane(*st, S, sig, en, case.Ng, nyz, nang, d.dg, d.du, d.dv,
                                  case.phys().density)
    p = g.spot_plane()
    np.testing.assert_allclose(p, p_ref, rtol=1e-12, atol=1e-12 * np.abs(p_ref).max())
    iy, iz = np.arange(nyz) % (case.Ny + 1), np.arange(nyz) // (case.Ny + 1)
    w = np.where((iy == 0) | (iy == case.Ny), 0.5, 1.0) * np.where((iz == 0) | (iz == case.Nz), 0.5, 1.0)
    # the IDD sums the energy sweep's FP64 outputs, the plane the stored state
    # (rounded to FP32 in the FP32 instantiation)
    tol = 1e-11 if precision == "FP64" else 1e-6
    assert p @ w * d.dy * d.dz == pytest.approx(g.history()[0][-1], rel=tol)


def test_cli_writes_energy_moments_and_spot_planes(tmp_path, gpu):
    """--save-energy-moments and --spot-depths/--spot-prefix write the
    reference's files (names, headers, layout) with the API's values."""
    cli = os.path.join(ROOT, "bfpn_gpu_amd", "bin", "bfp_solver")
    data = os.path.join(ROOT, "bfpn_gpu_amd", "tables", "water")
    cmd = [cli, "1000", "--data", data, "--material", "water", "--energy", "100", "--ny", "4",
           "--nz", "3", "--nmu", "4", "--nom", "5", "--ly", "4", "--lz", "3", "--lu", "2",
           "--lv", "2", "--ng", "60", "--sigma-yz", "0.3", "--trapezoidal-yz", "--time", "0.2",
           "--precision", "fp64", "--save-energy-moments", "--spot-depths", "0.05,,0.11",
           "--spot-prefix", "plane"]
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    idd = np.loadtxt(tmp_path / "idd_output.txt")
    steps, dt = len(idd), idd[0, 0]
    lines = open(tmp_path / "energy_moments.txt").read().splitlines()
    assert lines[0] == "# step x_cm g psi1_integral psi2_integral"
    mom = np.loadtxt(tmp_path / "energy_moments.txt")
    assert mom.shape == (steps * 60, 5)
    np.testing.assert_array_equal(mom[:, 0], np.repeat(np.arange(1, steps + 1), 60))
    np.testing.assert_array_equal(mom[:, 2], np.tile(np.arange(60), steps))
    # spot planes: after the first step with (k dt) + dt/2 >= depth
    files = sorted(p.name for p in tmp_path.glob("plane_req*.txt"))
    k1, k2 = int(np.ceil(0.05 / dt - 0.5 - 1e-9)), int(np.ceil(0.11 / dt - 0.5 - 1e-9))
    assert files == sorted([f"plane_req0.05_x{k1 * dt:.2f}.txt", f"plane_req0.11_x{k2 * dt:.2f}.txt"])
    body = open(tmp_path / files[0]).read().splitlines()
    assert body[0].startswith("# requested_x_cm 0.05") and body[2] == "# y_cm z_cm dose"
    blocks = [b for b in "\n".join(body[3:]).split("\n\n") if b.strip()]
    assert len(blocks) == 4 and all(len(b.splitlines()) == 5 for b in blocks)
    y0, z0, _ = map(float, blocks[0].splitlines()[0].split())
    assert (y0, z0) == (-2.0, -1.5)


def test_dst_takes_the_vector_path_beyond_32_bit_plane_offsets(gpu):
    """104x104 nodes x 20x20 directions x 500 groups: 2.16e9 cells, past the
    32-bit plane offsets of the MFMA DST, so the FP32 solver takes the vector
    DST (FP32) and still agrees with the FP64 instantiation
    after one full depth step (86 GB of FP32 state, 173 GB of FP64)."""
    from bfpn_gpu_amd.abi import Case
    from bfpn_gpu_amd.tables import material_dir

    case = Case(Nx=4000, Ny=103, Nz=103, Ng=500, Nmu=20, Nom=20, Ly=8, Lz=8, Lu=2, Lv=2,
                energy=230.0, sigma_yz=0.3, material="water", data=material_dir("water"),
                time=40.0, flags=dict(eq15_straggling=1))
    assert case.cells * 1 >= 2**31
    with gpu.Solver(case, gpu.FP32) as a:
        a.advance(2)
        ia = a.history()[0]
    with gpu.Solver(case, gpu.FP64) as b:
        b.advance(2)
        ib = b.history()[0]
    np.testing.assert_allclose(ia, ib, rtol=1e-4)
