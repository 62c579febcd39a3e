"""Active-group skipping: every energy sweep measures the highest nonzero
group of what it writes, and the operators skip the groups above it.  The
skipped results are exact zeros, so runs with skipping on and off must agree
bit for bit -- IDD, proxy and all four state arrays -- in both precisions,
across advance() call boundaries, with and without the two-stream overlap,
after a host write of the state, and for beams whose top group falls far
below Ng as they slow down (the case skipping exists for).  The path with
skipping off is the one the oracle parity tests pin.
"""
import os

import numpy as np
import pytest

from bfpn_gpu_amd.abi import Case
from conftest import TABLES
from gpu_cases import small_case

pytestmark = pytest.mark.gpu


def low_energy_case(material="water", Ng=300, energy=70.0, **over):
    """A 70 MeV beam on a 260 MeV group grid, past its range (~4 cm in water)."""
    kw = dict(Nx=2000, Ny=4, Nz=4, Ng=Ng, Nmu=8, Nom=8, Ly=4, Lz=4, Lu=2, Lv=2, energy=energy,
              sigma_yz=0.4, material=material, data=os.path.join(TABLES, material), time=5.0,
              flags=dict(eq15_straggling=1))
    kw.update(over)
    return Case(**kw)


def run(gpu, case, precision, skip, chunks, overlap=True, rewrite_at=None):
    with gpu.Solver(case, precision) as s:
        s.set_group_skip(skip)
        s.set_overlap(overlap)
        tops = []
        for k, n in enumerate(chunks):
            s.advance(n)
            tops.append(s.active_groups())
            if rewrite_at is not None and k == rewrite_at:
                for w in range(4):
                    s.set_state(w, s.state(w))
        idd, proxy = s.history()
        states = [s.state(w) for w in range(4)]
    return idd, proxy, states, tops


def assert_identical(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    for w in range(4):
        np.testing.assert_array_equal(a[2][w], b[2][w], err_msg=f"state {w}")


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
@pytest.mark.parametrize("flags", [
    {"eq15_straggling": 1},
    {},
    {"primary_only": 1, "eq15_straggling": 1},
    {"energy_only": 1},
], ids=lambda f: "-".join(sorted(f)) or "default")
def test_skipping_is_bit_identical_on_the_small_case(gpu, precision, flags):
    case = small_case(flags=flags)
    p = getattr(gpu, precision)
    on = run(gpu, case, p, True, [3, 1, 5])
    off = run(gpu, case, p, False, [3, 1, 5])
    assert_identical(on, off)


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
@pytest.mark.parametrize("material", ["water", "bone"])
def test_skipping_follows_a_slowing_beam_bit_identically(gpu, precision, material):
    case = low_energy_case(material)
    p = getattr(gpu, precision)
    n = case.num_steps()
    chunks = [1, 7, 40, 1, 1] + [50] * ((n - 50) // 50) + [(n - 50) % 50]
    chunks = [c for c in chunks if c > 0]
    assert sum(chunks) == n
    on = run(gpu, case, p, True, chunks)
    off = run(gpu, case, p, False, chunks)
    assert_identical(on, off)
    assert all(t == case.Ng - 1 for t, _ in off[3])
    tops = [t[0] for t in on[3]]
    assert max(tops) < case.Ng - 1, tops
    if precision == "FP32":
        # a 70 MeV beam on a [1, 260] MeV grid never reaches the upper groups
        # (its FP32 tails underflow about 15 MeV above it), and its top falls
        # as it slows; FP64 tails reach much further before they underflow
        assert max(tops) < case.Ng // 2, tops
        assert tops[-1] < tops[1], tops


def test_skipping_with_serial_streams_and_a_state_rewrite(gpu):
    case = low_energy_case(Ng=200, time=2.0)
    n = case.num_steps()
    chunks = [10, 20, n - 30]
    ref = run(gpu, case, gpu.FP32, False, chunks)
    assert_identical(run(gpu, case, gpu.FP32, True, chunks, overlap=False), ref)
    assert_identical(run(gpu, case, gpu.FP32, True, chunks, rewrite_at=0), ref)


def test_skipping_toggled_mid_run(gpu):
    case = low_energy_case(Ng=200, time=2.0)
    ref = run(gpu, case, gpu.FP32, True, [100])
    with gpu.Solver(case, gpu.FP32) as s:
        s.set_group_skip(False)
        s.advance(30)
        s.set_group_skip(True)
        s.advance(40)
        s.set_group_skip(False)
        s.advance(10)
        s.set_group_skip(True)
        s.advance(20)
        idd, proxy = s.history()
        np.testing.assert_array_equal(idd, ref[0])
        np.testing.assert_array_equal(proxy, ref[1])
        for w in range(4):
            np.testing.assert_array_equal(s.state(w), ref[2][w])


@pytest.mark.parametrize("Ny,skip", [(40, True), (40, False), (6, True)])
def test_call_boundaries_do_not_change_results(gpu, Ny, skip):
    """One advance() over all steps and the same steps in slices (as the
    serving loop issues them) give the same bits: paired energy sweeps inside
    a call equal two single sweeps at its ends.  41x41 nodes x 20x20
    directions is the four-lanes-per-thread sweep; 7x7 the one-lane sweep."""
    case = low_energy_case(Ny=Ny, Nz=Ny, Nmu=20, Nom=20, Ng=120, energy=100.0, Ly=8, Lz=8,
                           time=0.6)
    n = case.num_steps()
    whole = run(gpu, case, gpu.FP32, skip, [n])
    sliced = run(gpu, case, gpu.FP32, skip, [1, 2, 8, 3, n - 14])
    assert_identical(whole, sliced)
