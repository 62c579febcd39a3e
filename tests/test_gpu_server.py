"""Continuous-batching server (bfpn_server) on the GPU.

* Each beam's IDD is identical, bit for bit, alone (a plain Solver) and in a
  mix of other beams running concurrently (beams share only read-only tables).
* Admission respects the HBM budget and slot count; a request that can never
  fit fails with a message and does not disturb the others.
* Requests submitted while others are running are admitted and completed.
"""
import os

import numpy as np
import pytest

from conftest import TABLES
from gpu_cases import small_case

pytestmark = pytest.mark.gpu


def _mix():
    return [
        small_case(Ny=3, Nz=3, Nmu=4, Nom=4, energy=120.0, time=0.12, flags={"eq15_straggling": 1}),
        small_case(Ny=5, Nz=4, Nmu=5, Nom=4, energy=200.0, time=0.16),
        small_case(Ny=3, Nz=3, Nmu=4, Nom=4, energy=90.0, time=0.08, material="bone", density=1.85,
                   data=small_case().data.replace("water", "bone")),
        small_case(Ny=4, Nz=4, Nmu=4, Nom=4, energy=150.0, time=0.20, flags={"primary_only": 1}),
    ]


def _alone(gpu, case):
    with gpu.Solver(case, gpu.FP32) as s:
        idd, _ = s.run()
    return idd


@pytest.mark.parametrize("max_active,slice_steps", [(4, 1), (2, 3), (1, 2)])
def test_results_are_identical_alone_and_in_the_mix(gpu, max_active, slice_steps):
    cases = _mix()
    ref = [_alone(gpu, c) for c in cases]
    with gpu.Server(max_active=max_active, slice_steps=slice_steps) as srv:
        ids = [srv.submit(c) for c in cases]
        out = [srv.wait(i, timeout=120)[0] for i in ids]
        st = srv.stats()
    for a, b in zip(out, ref):
        assert len(a) == len(b) > 0
        np.testing.assert_array_equal(a, b)
    assert st["completed"] == len(cases) and st["failed"] == 0
    assert st["peak_active"] <= max_active


def test_budget_limits_admission_and_shares_tables(gpu):
    cases = [small_case(Ny=3, Nz=3, Nmu=4, Nom=4, energy=100.0 + 10 * k, time=0.08) for k in range(3)]
    ref = [_alone(gpu, c) for c in cases]
    beam, shared = gpu.Server.estimate(cases[0])
    # room for the shared tables and one and a half beams: admission runs them one at a time
    with gpu.Server(max_active=3, slice_steps=1, hbm_budget_bytes=shared + beam * 3 // 2) as srv:
        ids = [srv.submit(c) for c in cases]
        out = [srv.wait(i, timeout=120)[0] for i in ids]
        st = srv.stats()
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)
    assert st["peak_active"] == 1 and st["completed"] == 3 and st["cached_tables"] == 1


def test_oversized_request_fails_cleanly(gpu):
    small = small_case(Ny=3, Nz=3, Nmu=4, Nom=4, time=0.08)
    huge = small_case(Ny=200, Nz=200, Nmu=20, Nom=20, time=0.08)
    with gpu.Server(max_active=2, slice_steps=2, hbm_budget_bytes=512 * 1024 * 1024) as srv:
        a = srv.submit(small)
        b = srv.submit(huge)
        c = srv.submit(small)
        with pytest.raises(gpu.BfpnError, match="HBM budget"):
            srv.wait(b, timeout=120)
        ra = srv.wait(a, timeout=120)[0]
        rc = srv.wait(c, timeout=120)[0]
    np.testing.assert_array_equal(ra, rc)


def test_requests_join_while_others_run(gpu):
    """A short request arriving while a long one runs is admitted, runs
    concurrently, and (shortest remaining work first) finishes first."""
    import time

    long_case = small_case(Ny=5, Nz=5, Nmu=5, Nom=5, energy=200.0, time=20.0)
    short = small_case(Ny=3, Nz=3, Nmu=4, Nom=4, energy=100.0, time=0.08)
    ref_short = _alone(gpu, short)
    with gpu.Server(max_active=4, slice_steps=2) as srv:
        a = srv.submit(long_case)
        while srv.poll(a) == 0:  # wait until the long beam runs
            time.sleep(0.005)
        b = srv.submit(short)
        rb, tb = srv.wait(b, timeout=120)
        assert srv.poll(a) == 1  # still running: the short one overtook it
        ra, ta = srv.wait(a, timeout=120)
        st = srv.stats()
    np.testing.assert_array_equal(rb, ref_short)
    assert len(ra) == long_case.num_steps()
    assert st["peak_active"] == 2


def test_alternating_table_keys_under_a_tight_budget(gpu):
    """Water and bone requests (two table keys) alternate under a budget
    that holds one set of tables and one beam: admitting one key evicts the
    other's unused tables but never its own, every byte stays charged, and
    the results are those of plain solves."""
    w = small_case(Ny=3, Nz=3, Nmu=4, Nom=4, time=0.08)
    b = small_case(Ny=3, Nz=3, Nmu=4, Nom=4, time=0.08, material="bone", density=1.85,
                   data=os.path.join(TABLES, "bone"))
    cases = [w, b, w, b, w, b]
    ref = {id(c): _alone(gpu, c) for c in (w, b)}
    (beam_w, shared_w), (beam_b, shared_b) = gpu.Server.estimate(w), gpu.Server.estimate(b)
    budget = max(shared_w, shared_b) + max(beam_w, beam_b) + min(shared_w, shared_b) // 2
    with gpu.Server(max_active=1, slice_steps=1, hbm_budget_bytes=budget) as srv:
        ids = [srv.submit(c) for c in cases]
        out = [srv.wait(i, timeout=120)[0] for i in ids]
        st = srv.stats()
    for c, o in zip(cases, out):
        np.testing.assert_array_equal(o, ref[id(c)])
    assert st["completed"] == len(cases) and st["failed"] == 0
    assert st["cached_tables"] == 1
    assert 0 < st["bytes_in_use"] <= budget


def _layered_mix():
    """Two plain water beams and two slab beams (water/bone/water and
    water/air/water) on one grid: the water layers' tables are the plain
    beams' (same key), so the cache holds three table sets."""
    def tables(m):
        return os.path.join(TABLES, m)

    water = ("water", 1.0, tables("water"))
    base = dict(Ny=3, Nz=3, Nmu=4, Nom=4, time=1.2, flags={"eq15_straggling": 1})
    wbw = [(0.31, *water), (0.71, "bone", 1.85, tables("bone")), (9.0, *water)]
    waw = [(0.51, *water), (0.93, "air", 0.001205, tables("air")), (9.0, *water)]
    return [(small_case(energy=100.0, **base), None), (small_case(energy=100.0, **base), wbw),
            (small_case(energy=120.0, **base), None), (small_case(energy=100.0, **base), waw)]


@pytest.mark.parametrize("max_active,slice_steps", [(4, 1), (2, 7)])
def test_layered_requests_match_their_solo_runs(gpu, max_active, slice_steps):
    """Slab beams served beside plain beams (slices ending on both sides of
    the layer boundaries, mid-pair included) give the solo layered Solver's
    IDD bit for bit; each material's tables are built once and shared."""
    mix = _layered_mix()
    ref = []
    for case, layers in mix:
        with gpu.Solver(case, gpu.FP32, layers=layers) as s:
            ref.append(s.run()[0])
    with gpu.Server(max_active=max_active, slice_steps=slice_steps) as srv:
        ids = [srv.submit(case, layers=layers) for case, layers in mix]
        out = [srv.wait(i, timeout=120)[0] for i in ids]
        st = srv.stats()
    for (case, layers), a, b in zip(mix, out, ref):
        assert len(a) == len(b) == case.num_steps()
        np.testing.assert_array_equal(a, b, err_msg=str(layers))
    assert not np.array_equal(ref[0], ref[1])  # the slab changes the dose
    assert st["completed"] == len(mix) and st["failed"] == 0
    assert st["cached_tables"] == 3


def test_layered_request_with_missing_tables_fails_alone(gpu):
    """A slab beam whose middle layer names a table directory that does not
    exist fails with the loader's message; the beams around it, a good slab
    beam among them, complete with their solo results."""
    mix = _layered_mix()
    good_case, good_layers = mix[1]
    bad_layers = [good_layers[0], (0.71, "bone", 1.85, os.path.join(TABLES, "no_such_material")),
                  good_layers[2]]
    water_case = mix[0][0]
    with gpu.Solver(good_case, gpu.FP32, layers=good_layers) as s:
        ref_slab = s.run()[0]
    ref_water = _alone(gpu, water_case)
    with gpu.Server(max_active=2, slice_steps=4) as srv:
        a = srv.submit(water_case)
        b = srv.submit(good_case, layers=bad_layers)
        c = srv.submit(good_case, layers=good_layers)
        with pytest.raises(gpu.BfpnError, match="no_such_material"):
            srv.wait(b, timeout=120)
        ra = srv.wait(a, timeout=120)[0]
        rc = srv.wait(c, timeout=120)[0]
        st = srv.stats()
    np.testing.assert_array_equal(ra, ref_water)
    np.testing.assert_array_equal(rc, ref_slab)
    assert st["failed"] == 1 and st["completed"] == 2


def test_mapped_requests_match_their_solo_runs(gpu):
    """A beam under a material map (water, then bone for the nodes with iy < 2
    from 0.31 cm on) served beside a plain and a slab beam gives its solo
    mapped run bit for bit; its materials' tables are the cached water and
    bone sets the other beams use."""
    mix = _layered_mix()
    water_case, slab_case, slab_layers = mix[0][0], mix[1][0], mix[1][1]
    nyz = water_case.nyz
    split = (np.arange(nyz) % (water_case.Ny + 1) < 2).astype(np.uint8)
    mapped = ([("water", 1.0, os.path.join(TABLES, "water")),
               ("bone", 1.85, os.path.join(TABLES, "bone"))], [0.31, 100.0],
              np.vstack([np.zeros(nyz, np.uint8), split]))
    with gpu.Solver(slab_case, gpu.FP32, mapped=mapped) as s:
        ref_map = s.run()[0]
    with gpu.Solver(slab_case, gpu.FP32, layers=slab_layers) as s:
        ref_slab = s.run()[0]
    ref_water = _alone(gpu, water_case)
    with gpu.Server(max_active=3, slice_steps=5) as srv:
        a = srv.submit(water_case)
        b = srv.submit(slab_case, mapped=mapped)
        c = srv.submit(slab_case, layers=slab_layers)
        out = [srv.wait(i, timeout=120)[0] for i in (a, b, c)]
        st = srv.stats()
    np.testing.assert_array_equal(out[0], ref_water)
    np.testing.assert_array_equal(out[1], ref_map)
    np.testing.assert_array_equal(out[2], ref_slab)
    assert not np.array_equal(ref_map, ref_slab)
    assert st["completed"] == 3 and st["failed"] == 0 and st["cached_tables"] == 2
