"""Depth-layered media in the oracle.  The reference solver is homogeneous
(its material-map item, TODO.md:114-118, and validation gap, README.md:30-37,
are open), so no reference output covers a layered run; the layered oracle is
pinned here by what the homogeneous one and the physics fix:
- a one-layer stack, and a stack of identical layers, are the homogeneous run
  bit for bit;
- the layer of depth step n is the one holding its midpoint (n + 1/2) dt: the
  steps before a boundary are the first material's homogeneous run, the first
  step past it is not;
- a slab shifts the distal dose fall-off (D90) by its water-equivalent
  thickness minus its thickness, as a CSDA integration of the layer's own
  stopping-power table (scipy, independent of the solver) predicts; a wrong
  density, table or layer order moves it by millimetres to centimetres.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from bfpn_gpu_amd import metrics
from bfpn_gpu_amd.abi import Case
from bfpn_gpu_amd.tables import material_dir

WATER = ("water", 1.0, material_dir("water"))


def line_case(**over):
    """Energy sweep only, 2x2 nodes x 1 direction: a 100 MeV beam's CSDA
    slowing-down to 11 cm, dt 0.01 cm, 300 groups on [1, 260] MeV."""
    kw = dict(Nx=4000, Ny=1, Nz=1, Ng=300, Nmu=1, Nom=1, Ly=4, Lz=4, Lu=2, Lv=2, energy=100.0,
              sigma_yz=0.5, material="water", data=material_dir("water"), time=11.0,
              flags=dict(energy_only=1))
    kw.update(over)
    return Case(**kw)


def small_case(**over):
    kw = dict(Nx=1000, Ny=3, Nz=2, Ng=200, Nmu=3, Nom=2, Ly=4, Lz=4, Lu=2, Lv=2, energy=100.0,
              sigma_yz=0.5, material="water", data=material_dir("water"), time=1.2,
              flags=dict(eq15_straggling=1))
    kw.update(over)
    return Case(**kw)


def run(oracle_mod, case, layers=None, steps=None):
    s = oracle_mod.Solver(case, layers=layers)
    s.advance(case.num_steps() if steps is None else steps)
    return s


@pytest.mark.parametrize("flags", [{"eq15_straggling": 1}, {"legacy_energy": 1},
                                   {"lite_memory": 1}], ids=lambda f: next(iter(f)))
def test_one_layer_and_identical_layers_are_the_homogeneous_run(oracle_mod, flags):
    case = small_case(flags=flags)
    ref = run(oracle_mod, case)
    for layers in ([(100.0, *WATER)], [(0.33, *WATER), (0.71, *WATER), (5.0, *WATER)]):
        s = run(oracle_mod, case, layers)
        np.testing.assert_array_equal(s.history()[0], ref.history()[0])
        np.testing.assert_array_equal(s.history()[1], ref.history()[1])
        for w in range(4):
            np.testing.assert_array_equal(s.state(w), ref.state(w))


def test_layer_of_a_step_is_the_one_holding_its_midpoint(oracle_mod):
    # dt 0.04 cm: step 4 spans [0.16, 0.20] with midpoint 0.18.  A boundary at
    # 0.17 puts it in the second layer (a start-of-step rule would not), one
    # at 0.19 in the first (an end-of-step rule would not).
    case = small_case()
    bone = ("bone", 1.85, material_dir("bone"))
    water = run(oracle_mod, case).history()[0]
    for boundary, first_layer_steps in ((0.17, 4), (0.19, 5)):
        s = run(oracle_mod, case, [(boundary, *WATER), (9.0, *bone)]).history()[0]
        k = first_layer_steps
        np.testing.assert_array_equal(s[:k], water[:k])
        assert abs(s[k] / water[k] - 1) > 1e-3, boundary


def csda_shift(oracle_mod, case, slab, material, density):
    """D90 shift of a slab [a, b] of `material` in water, from the CSDA:
    (b + R_w(E_out)) - R_w(E0), with E_out the energy leaving the slab."""
    w = oracle_mod.Solver(case)
    m = oracle_mod.Solver(line_case(material=material, data=material_dir(material)))
    Sw, Sm = w.table("S"), m.table("S")
    Ng, Lg = case.Ng, case.Lg  # the table's points: 1, (i - 1/2) dg + 1, 1 + Lg
    en = np.concatenate(([1.0], (np.arange(1, Ng) - 0.5) * Lg / Ng + 1.0, [1.0 + Lg]))

    def slow(E, t, S, rho):
        f = lambda x, e: [-rho * np.interp(e[0], en, S)]  # noqa: E731
        return solve_ivp(f, (0, t), [E], rtol=1e-10, atol=1e-10).y[0, -1]

    def Rw(E):  # CSDA range in water from E down to the grid's 1 MeV floor
        e = np.linspace(1.0, E, 400001)
        return np.trapezoid(1.0 / np.interp(e, en, Sw), e)

    a, b = slab
    E_out = slow(slow(case.energy, a, Sw, 1.0), b - a, Sm, density)
    return b + Rw(E_out) - Rw(case.energy)


@pytest.mark.parametrize("material,density,tol", [("bone", 1.85, 0.005), ("bone", 1.0, 0.005),
                                                  ("air", 0.001205, 0.01)])
def test_slab_shifts_the_distal_falloff_by_its_water_equivalent_thickness(
        oracle_mod, material, density, tol):
    case = line_case()
    n = case.num_steps()
    x = (np.arange(n) + 1) * case.dt
    behind = x > 4.0  # the IDD scales by 1/density: in air it dwarfs the water's
    water = metrics.analyze(x[behind], run(oracle_mod, case).history()[0][behind])
    layers = [(2.0, *WATER), (4.0, material, density, material_dir(material)), (100.0, *WATER)]
    slab = metrics.analyze(x[behind], run(oracle_mod, case, layers).history()[0][behind])
    want = csda_shift(oracle_mod, case, (2.0, 4.0), material, density)
    got = slab["D90"] - water["D90"]
    assert abs(got - want) < tol, (got, want)
    if material == "bone" and density > 1:
        assert got < -1.0  # a dense slab pulls the peak in by centimetres
