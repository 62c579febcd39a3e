"""Material maps on the GPU (bfpn_create_mapped; no reference counterpart:
the reference solver is homogeneous, its material-map item, TODO.md:114-118,
is open).  Node s of depth slab j uses the tables and density of material
map[j][s] in every node-local operator (energy sweeps, angular diffusion,
secondary source, dose); transport is material-independent.  Checked against:
- the homogeneous run, bit for bit, for a one-material map (both precisions)
  and, for the states, for two identical materials on a checkerboard;
- the layered run, bit for bit, for maps uniform within each slab;
- the homogeneous run of each node's own material, bit for bit, for the
  primary with its transport off (every primary node then evolves alone);
- the mapped oracle (pinned on the CPU by tests/test_oracle_maps.py) for a
  lateral water/bone split: FP64 to 1e-10, FP32 to 2e-4 (the proxy, which
  changes sign, to 1e-3 of its maximum).
Group skipping and the secondary's second stream (both on by default, maps
included) are off in the reference comparisons and checked separately: on or
off, a mapped run is the same bit for bit.
"""
import numpy as np
import pytest

from bfpn_gpu_amd.abi import Case
from bfpn_gpu_amd.tables import material_dir

pytestmark = pytest.mark.gpu

WATER = ("water", 1.0, material_dir("water"))
BONE = ("bone", 1.85, material_dir("bone"))


def case_of(**over):
    """100 MeV beam, 6x5 nodes x 4x3 directions, 200 groups, dt 0.04 cm,
    every operator and the secondary active."""
    kw = dict(Nx=1000, Ny=5, Nz=4, Ng=200, Nmu=4, Nom=3, Ly=4, Lz=4, Lu=2, Lv=2, energy=100.0,
              sigma_yz=0.5, material="water", data=material_dir("water"), time=1.2,
              flags=dict(eq15_straggling=1, trapezoidal_yz=1))
    flags = over.pop("flags", None)
    kw.update(over)
    c = Case(**kw)
    if flags is not None:
        c.flags = flags
    return c


def run_gpu(gpu, case, precision, chunks=None, skip=False, overlap=False, **kw):
    with gpu.Solver(case, precision, **kw) as s:
        s.set_group_skip(skip)
        s.set_overlap(overlap)
        for n in chunks or [case.num_steps()]:
            s.advance(n)
        idd, proxy = s.history()
        return idd, proxy, [s.state(w) for w in range(4)]


def assert_identical(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    for w in range(4):
        np.testing.assert_array_equal(a[2][w], b[2][w], err_msg=f"state {w}")


def split_map(case):
    """Bone for the nodes with iy below the middle column, water elsewhere."""
    left = (np.arange(case.nyz) % (case.Ny + 1)) < (case.Ny + 1) // 2
    return left, np.where(left, 1, 0).astype(np.uint8)[None, :]


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
def test_one_material_map_is_the_homogeneous_run(gpu, precision):
    case = case_of()
    p = getattr(gpu, precision)
    mapped = ([WATER], [100.0], np.zeros((1, case.nyz)))
    assert_identical(run_gpu(gpu, case, p, mapped=mapped, chunks=[7, case.num_steps() - 7]),
                     run_gpu(gpu, case, p))


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
def test_identical_materials_on_a_checkerboard_are_the_homogeneous_run(gpu, precision):
    case = case_of()
    p = getattr(gpu, precision)
    board = (np.arange(case.nyz) + np.arange(case.nyz) // (case.Ny + 1)) % 2
    got = run_gpu(gpu, case, p, mapped=([WATER, WATER], [100.0], board[None, :]))
    ref = run_gpu(gpu, case, p)
    for w in range(4):
        np.testing.assert_array_equal(got[2][w], ref[2][w], err_msg=f"state {w}")
    np.testing.assert_allclose(got[0], ref[0], rtol=1e-13)  # dose summed per material
    np.testing.assert_allclose(got[1], ref[1], rtol=1e-13, atol=1e-13 * np.abs(ref[1]).max())


def test_slab_uniform_maps_are_the_layered_run(gpu):
    case = case_of()
    layers = [(0.33, *WATER), (0.71, *BONE), (9.0, *WATER)]
    mp = np.zeros((3, case.nyz), dtype=np.uint8)
    mp[1] = 1
    got = run_gpu(gpu, case, gpu.FP32, mapped=([WATER, BONE], [0.33, 0.71, 9.0], mp),
                  chunks=[13, case.num_steps() - 13])
    assert_identical(got, run_gpu(gpu, case, gpu.FP32, layers=layers))


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
def test_without_transport_each_primary_node_is_its_materials_run(gpu, precision):
    flags = dict(eq15_straggling=1, trapezoidal_yz=1, no_transport=1)
    case = case_of(flags=flags)
    p = getattr(gpu, precision)
    left, mp = split_map(case)
    got = run_gpu(gpu, case, p, mapped=([WATER, BONE], [100.0], mp))[2]
    water = run_gpu(gpu, case, p)[2]
    bone = run_gpu(gpu, case_of(material="bone", density=1.85, data=material_dir("bone"),
                                flags=flags), p)[2]
    shape = (case.Ng, case.nyz, case.nang)
    for w in range(2):
        g = got[w].reshape(shape)
        np.testing.assert_array_equal(g[:, ~left], water[w].reshape(shape)[:, ~left])
        np.testing.assert_array_equal(g[:, left], bone[w].reshape(shape)[:, left])


@pytest.mark.parametrize("precision,tol", [("FP64", 1e-10), ("FP32", 2e-4)])
def test_lateral_map_matches_oracle(gpu, oracle_mod, precision, tol):
    case = case_of(time=2.4)
    _, mp = split_map(case)
    slabs = np.vstack([np.zeros_like(mp), mp])  # water to 0.5 cm, then the split
    mapped = ([WATER, BONE], [0.5, 100.0], slabs)
    o = oracle_mod.Solver(case, mapped=mapped)
    n = case.num_steps()
    o.advance(n)
    g = run_gpu(gpu, case, getattr(gpu, precision), mapped=mapped, chunks=[7, n - 7])
    oi, op = o.history()
    np.testing.assert_allclose(g[0], oi, rtol=tol)
    ptol = tol if precision == "FP64" else 1e-3
    np.testing.assert_allclose(g[1], op, rtol=ptol, atol=ptol * np.abs(op).max())
    scale = np.abs(o.state(0)).max()
    for w in range(4):
        assert np.abs(g[2][w] - o.state(w)).max() / scale < tol, w
    bone_only = run_gpu(gpu, case_of(time=2.4, material="bone", density=1.85,
                                     data=material_dir("bone")), getattr(gpu, precision))
    assert not np.allclose(g[0], bone_only[0], rtol=1e-3)  # the water half counts


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
def test_group_skipping_under_a_map_changes_nothing(gpu, precision):
    case = case_of(time=2.4)
    _, mp = split_map(case)
    mapped = ([WATER, BONE], [0.5, 100.0], np.vstack([np.zeros_like(mp), mp]))
    p = getattr(gpu, precision)
    assert_identical(run_gpu(gpu, case, p, mapped=mapped, skip=True, chunks=[9, case.num_steps() - 9]),
                     run_gpu(gpu, case, p, mapped=mapped, skip=False))


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
def test_second_stream_under_a_map_changes_nothing(gpu, precision):
    case = case_of(time=2.4)
    _, mp = split_map(case)
    mapped = ([WATER, BONE], [0.5, 100.0], np.vstack([np.zeros_like(mp), mp]))
    p = getattr(gpu, precision)
    n = case.num_steps()
    assert_identical(run_gpu(gpu, case, p, mapped=mapped, skip=True, overlap=True, chunks=[11, n - 11]),
                     run_gpu(gpu, case, p, mapped=mapped, skip=True, overlap=False))


def test_bad_maps_are_refused(gpu):
    case = case_of()
    with pytest.raises(gpu.BfpnError, match="names no material"):
        gpu.Solver(case, gpu.FP32, mapped=([WATER], [100.0], np.ones((1, case.nyz))))
    with pytest.raises(gpu.BfpnError, match="eq15 energy model"):
        gpu.Solver(case_of(flags=dict(legacy_energy=1)), gpu.FP32,
                   mapped=([WATER], [100.0], np.zeros((1, case.nyz))))
    with pytest.raises(gpu.BfpnError, match="slab depths must increase"):
        gpu.Solver(case, gpu.FP32, mapped=([WATER], [0.5, 0.4, 9.0], np.zeros((3, case.nyz))))


def test_cli_material_map_writes_the_oracle_values(tmp_path, gpu, oracle_mod):
    import os
    import subprocess

    from conftest import ROOT

    case = case_of(time=2.4)
    _, mp = split_map(case)
    slabs = np.vstack([np.zeros_like(mp), mp])
    w, b = material_dir("water"), material_dir("bone")

    def rows(slab):
        return "".join(" ".join(str(int(v)) for v in slab[iz * (case.Ny + 1):(iz + 1) * (case.Ny + 1)])
                       + "\n" for iz in range(case.Nz + 1))

    path = tmp_path / "phantom.txt"
    path.write_text(f"# water, then a bone half from 0.5 cm\nmaterial water 1.0 {w}\n"
                    f"material bone 1.85 {b}\nslab 0.5\n{rows(slabs[0])}slab inf\n{rows(slabs[1])}")
    cli = os.path.join(ROOT, "bfpn_gpu_amd", "bin", "bfp_solver")
    cmd = [cli, str(case.Nx), "--data", case.data, "--material", "water", "--energy", "100",
           "--ny", str(case.Ny), "--nz", str(case.Nz), "--nmu", str(case.Nmu), "--nom",
           str(case.Nom), "--ly", "4", "--lz", "4", "--lu", "2", "--lv", "2", "--ng", "200",
           "--sigma-yz", "0.5", "--eq15-straggling", "--trapezoidal-yz", "--time", "2.4",
           "--precision", "fp64", "--material-map", str(path)]
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Material map: 2 depth slab(s)" in r.stdout
    out = np.loadtxt(tmp_path / "idd_output.txt")
    o = oracle_mod.Solver(case, mapped=([WATER, BONE], [0.5, 100.0], slabs))
    o.advance(case.num_steps())
    np.testing.assert_allclose(out[:, 1], o.history()[0], rtol=1e-10)
