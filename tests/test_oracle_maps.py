"""Material maps in the oracle (no counterpart in the reference: its
material-map item, TODO.md:114-118, is open).  Every operator but transport
is node-local, so the mapped oracle is pinned by what the homogeneous and
layered ones fix:
- a map of one material is the homogeneous run bit for bit;
- two materials with identical tables on a checkerboard of nodes give the
  homogeneous states bit for bit (the dose is summed per material, so the IDD
  agrees to rounding);
- maps uniform within each depth slab are the layered run bit for bit;
- with the primary's transport off (the reference's isolation flags leave the
  secondary's march as it is, src/bfpm_solver.cu:3335-3394) every primary node
  evolves alone: each node of a water/bone map holds, bit for bit, the
  primary state the homogeneous run of its own material gives it.
"""
import numpy as np
import pytest

from bfpn_gpu_amd.abi import Case
from bfpn_gpu_amd.tables import material_dir

WATER = ("water", 1.0, material_dir("water"))
BONE = ("bone", 1.85, material_dir("bone"))


def small_case(**over):
    """100 MeV beam, 4x3 nodes x 3x2 directions, 200 groups, dt 0.04 cm."""
    kw = dict(Nx=1000, Ny=3, Nz=2, Ng=200, Nmu=3, Nom=2, Ly=4, Lz=4, Lu=2, Lv=2, energy=100.0,
              sigma_yz=0.5, material="water", data=material_dir("water"), time=0.8,
              flags=dict(eq15_straggling=1, trapezoidal_yz=1))
    flags = over.pop("flags", None)
    kw.update(over)
    c = Case(**kw)
    if flags is not None:
        c.flags = flags
    return c


def nyz(case):
    return (case.Ny + 1) * (case.Nz + 1)


def run(oracle_mod, case, steps=None, **kw):
    s = oracle_mod.Solver(case, **kw)
    s.advance(case.num_steps() if steps is None else steps)
    return s


def states(s):
    return [s.state(w).copy() for w in range(4)]


def test_one_material_map_is_the_homogeneous_run(oracle_mod):
    case = small_case()
    ref = run(oracle_mod, case)
    m = run(oracle_mod, case, mapped=([WATER], [100.0], np.zeros((1, nyz(case)))))
    np.testing.assert_array_equal(m.history()[0], ref.history()[0])
    np.testing.assert_array_equal(m.history()[1], ref.history()[1])
    for a, b in zip(states(m), states(ref)):
        np.testing.assert_array_equal(a, b)


def test_identical_materials_on_a_checkerboard_are_the_homogeneous_run(oracle_mod):
    case = small_case()
    ref = run(oracle_mod, case)
    board = (np.arange(nyz(case)) + np.arange(nyz(case)) // (case.Ny + 1)) % 2
    m = run(oracle_mod, case, mapped=([WATER, WATER], [100.0], board[None, :]))
    for a, b in zip(states(m), states(ref)):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(m.history()[0], ref.history()[0], rtol=1e-13)
    np.testing.assert_array_equal(m.history()[1], ref.history()[1])


def test_slab_uniform_maps_are_the_layered_run(oracle_mod):
    case = small_case()
    layers = [(0.33, *WATER), (0.55, *BONE), (9.0, *WATER)]
    ref = run(oracle_mod, case, layers=layers)
    mp = np.zeros((3, nyz(case)), dtype=np.uint8)
    mp[1] = 1
    m = run(oracle_mod, case, mapped=([WATER, BONE], [0.33, 0.55, 9.0], mp))
    np.testing.assert_array_equal(m.history()[0], ref.history()[0])
    for a, b in zip(states(m), states(ref)):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("flags", [dict(eq15_straggling=1, no_transport=1),
                                   dict(eq15_straggling=1, no_transport=1, no_angle=1)],
                         ids=["angle", "no-angle"])
def test_without_transport_each_primary_node_is_its_materials_homogeneous_run(oracle_mod, flags):
    case = small_case(flags=dict(trapezoidal_yz=1, **flags))
    left = (np.arange(nyz(case)) % (case.Ny + 1)) < (case.Ny + 1) // 2  # bone for small iy
    mp = np.where(left, 1, 0).astype(np.uint8)[None, :]
    m = states(run(oracle_mod, case, mapped=([WATER, BONE], [100.0], mp)))
    water = states(run(oracle_mod, case))
    bone = states(run(oracle_mod, small_case(material="bone", density=1.85,
                                             data=material_dir("bone"),
                                             flags=dict(trapezoidal_yz=1, **flags))))
    shape = (case.Ng, nyz(case), case.nang)
    for w in range(2):
        got = m[w].reshape(shape)
        np.testing.assert_array_equal(got[:, ~left], water[w].reshape(shape)[:, ~left])
        np.testing.assert_array_equal(got[:, left], bone[w].reshape(shape)[:, left])
    assert not np.array_equal(water[0], bone[0])  # the materials differ


def test_map_entry_beyond_the_materials_is_refused(oracle_mod):
    case = small_case()
    mp = np.full((1, nyz(case)), 2, dtype=np.uint8)
    with pytest.raises(OSError, match="names no material"):
        oracle_mod.Solver(case, mapped=([WATER, BONE], [100.0], mp))
