"""Each benchmark configuration's own workload on the GPU.

  water 100 MeV, 11x11 nodes over 4x4 cm, 8x8 directions, Ng = 200 on
      [1, 101] MeV (synthetic generator tables at Lg = 100), 8 cm = 400 steps,
      straggling: the FP64 HIP instantiation against the CPU FP64 oracle over
      all 400 steps -- IDD and proxy within 1e-10 relative, the four state
      arrays within 1e-10 of their maxima (FMA contraction and summation
      order are the only differences);
  water 230 MeV, 21x21 nodes over 8x8 cm, 20x20 directions, Ng = 500,
      dt = 0.02 cm, straggling: FP32 state against the FP64 instantiation over
      the first 300 steps (12 cm), IDD within 0.5 % wherever the FP64 IDD is
      above 10 % of its maximum;
  the paper grid (81x81 nodes, 20x20 directions, Ng = 500, dt = 0.01 cm),
      water and bone (rho 1.85): 10 FP32 steps against 10 FP64 steps at full
      size (IDD within 0.5 % pointwise; the primary and secondary psi^1
      within 1e-4 of the FP64 primary's maximum -- in bone the secondary is
      still ~1e-57 of the primary after 10 steps, below the FP32 range);
  the seeded 64-beam batch (E ~ U[70, 230] MeV, sigma ~ U[0.2, 0.5] cm, 21x21
      or 41x41 nodes, Ng = 300, water/bone), each beam to 1 cm: served by the
      batching server, every IDD equals the beam's plain solve bit for bit;
  a seeded Poisson stream at 20 requests/s (21x21 to 81x81 nodes, Ng = 300),
      its first 6 requests submitted at their arrival times, each to 0.5 cm:
      every served IDD within 0.5 % of an FP64 solve of the same request.
"""

import numpy as np
import pytest

from bfpn_gpu_amd.abi import Case
from bfpn_gpu_amd.tables import material_dir, write_synthetic_tables

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def cpu_case(tmp_path):
    """The CPU-runnable configuration (Lu = Lv = 2, Nx = 2000)."""
    data = write_synthetic_tables(200, 100.0, str(tmp_path / "synthetic_ng200_lg100"))
    return Case(Nx=2000, Ny=10, Nz=10, Ng=200, Nmu=8, Nom=8, Ly=4, Lz=4, Lu=2, Lv=2, Lg=100.0,
                energy=100.0, sigma_e=1.0, sigma_yz=0.3, material="water", density=1.0,
                data=data, time=8.0, flags=dict(eq15_straggling=1))


def test_cpu_config_fp64_matches_oracle_over_all_steps(gpu, oracle_mod, tmp_path):
    case = cpu_case(tmp_path)
    o = oracle_mod.Solver(case)
    n = o.num_steps()
    assert n == 400 and o.cells == 200 * 121 * 64
    o.advance(n)
    oi, op = o.history()
    with gpu.Solver(case, gpu.FP64) as g:
        assert g.num_steps() == n
        g.advance(n)
        gi, gp = g.history()
        assert len(gi) == n
        np.testing.assert_allclose(gi, oi, rtol=1e-10)
        np.testing.assert_allclose(gp, op, rtol=1e-10, atol=1e-12 * np.abs(op).max())
        for k in range(4):
            assert _rel(g.state(k), o.state(k)) < 1e-10, k
    o.close()


def beam_case(Ny, Nx, material="water", density=None):
    return Case(Nx=Nx, Ny=Ny, Nz=Ny, Ng=500, Nmu=20, Nom=20, Ly=8, Lz=8, Lu=2, Lv=2,
                energy=230.0, sigma_yz=0.3, material=material, density=density,
                data=material_dir(material), time=40.0, flags=dict(eq15_straggling=1))


def test_single_beam_config_fp32_within_gate_of_fp64(gpu):
    from bfpn_gpu_amd import metrics

    case = beam_case(20, 2000)
    steps = 300
    with gpu.Solver(case, gpu.FP32) as a:
        a.advance(steps)
        ia = a.history()[0]
    with gpu.Solver(case, gpu.FP64) as b:
        b.advance(steps)
        ib = b.history()[0]
    assert len(ia) == len(ib) == steps
    cmp = metrics.compare(ia, ib, band=0.10)
    assert cmp["max_rel"] <= 5e-3, cmp


@pytest.mark.parametrize("material,density", [("water", None), ("bone", 1.85)])
def test_paper_grid_fp32_tracks_fp64_at_full_size(gpu, material, density):
    case = beam_case(80, 4000, material, density)
    steps = 10
    with gpu.Solver(case, gpu.FP32) as a:
        assert a.cells == 1312200000
        a.advance(steps)
        ia = a.history()[0]
        sa = [a.state(k) for k in (0, 2)]
    with gpu.Solver(case, gpu.FP64) as b:
        b.advance(steps)
        ib = b.history()[0]
        ref = [b.state(k) for k in (0, 2)]
    scale = np.abs(ref[0]).max()
    for k, s, r in zip((0, 2), sa, ref):
        assert np.abs(s - r).max() / scale < 1e-4, k
    assert np.all(np.abs(ia - ib) / ib < 5e-3), (ia, ib)


def test_beamlet_batch_served_equals_alone(gpu):
    from bfpn_gpu_amd.workloads import beamlet_batch

    cases = beamlet_batch(64)
    depth = 1.0
    alone = []
    for c in cases:
        with gpu.Solver(c, gpu.FP32) as s:
            alone.append(s.run(depth)[0])
    with gpu.Server(max_active=8, slice_steps=8) as srv:
        ids = [srv.submit(c, depth) for c in cases]
        served = [srv.wait(i, timeout=600)[0] for i in ids]
        st = srv.stats()
    assert st["completed"] == 64 and st["failed"] == 0
    for a, b in zip(alone, served):
        assert len(a) == 50
        np.testing.assert_array_equal(a, b)


def test_served_stream_within_gate_of_fp64(gpu):
    import threading
    import time

    from bfpn_gpu_amd import metrics
    from bfpn_gpu_amd.workloads import serving_requests

    reqs = serving_requests(20.0, 60.0)[:6]
    depth = 0.5
    with gpu.Server(max_active=4, slice_steps=8) as srv:
        ids = [None] * len(reqs)
        t0 = time.perf_counter()

        def submit():
            for k, (at, c) in enumerate(reqs):
                time.sleep(max(0.0, t0 + at - time.perf_counter()))
                ids[k] = srv.submit(c, depth)

        th = threading.Thread(target=submit)
        th.start()
        th.join()
        served = [srv.wait(i, timeout=600)[0] for i in ids]
    for (_, c), idd in zip(reqs, served):
        with gpu.Solver(c, gpu.FP64) as s:
            ref = s.run(depth)[0]
        assert len(idd) == len(ref) == 25
        assert metrics.compare(idd, ref, band=0.10)["max_rel"] <= 5e-3
