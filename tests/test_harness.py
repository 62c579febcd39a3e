"""Harness runners (Table 2 cases, depth-step convergence) on the host.

A stub executable stands in for bfp_solver: it reads the same command line
and writes an idd_output.txt whose depth-step error is known exactly, so the
runners' plumbing, error norms and observed order are checked against the
closed form.  The same runners drive the real binary in
tests/test_gpu_harness.py.
"""
import csv
import math
import os
import stat
import sys

import numpy as np
import pytest

from bfpn_gpu_amd import harness, metrics

STUB = r'''#!{python}
import sys, math
args = sys.argv[1:]
nx = int(args[0])
opt = {{args[i]: args[i + 1] for i in range(1, len(args) - 1) if args[i].startswith("--")}}
time = float(opt.get("--time", 40.0))
energy = float(opt.get("--energy", 230.0))
dt = time / nx
# a smooth depth-dose with a peak near the CSDA-like depth r, plus an error of C*dt^ORDER
r = {{"water": 0.0022, "bone": 0.0013}}[opt.get("--material", "water")] * energy ** 1.77
with open("idd_output.txt", "w") as f:
    f.write("# x_cm value\n")
    for i in range(nx):
        x = (i + 1) * dt
        v = math.exp(-((x - r) / (0.02 * r + 0.05)) ** 2) + 0.3 * math.exp(-x / (2 * r + 1e-9))
        v += {C} * dt ** {ORDER} * math.sin(x)
        # transverse size scales the dose, angular size adds a small wave
        v *= 1 + 0.01 * (int(opt.get("--ny", 20)) - 20)
        v += 1e-3 * (int(opt.get("--nmu", 20)) - 20) * math.sin(3 * x)
        f.write(f"{{x:.17g}} {{v:.17g}}\n")
print(f"Step {{nx}}, t = {{time}}")
'''


def _stub(tmp_path, C=0.0, order=2.0):
    p = tmp_path / "stub_solver"
    p.write_text(STUB.format(python=sys.executable, C=C, ORDER=order))
    p.chmod(p.stat().st_mode | stat.S_IEXEC)
    return p


def test_observed_order_of_a_known_power_law():
    rows = [{"h": h, "rel_l2": 3.0 * h ** 2} for h in (0.08, 0.04, 0.02)]
    harness.observed_orders(rows)
    assert rows[0]["order_to_next"] == pytest.approx(2.0, abs=1e-12)
    assert rows[1]["order_to_next"] == pytest.approx(2.0, abs=1e-12)
    assert math.isnan(rows[2]["order_to_next"])


def test_relative_norms_on_interpolated_curves():
    xf = np.linspace(0, 2, 2001)
    yf = 1.0 + xf
    xc = np.linspace(0.1, 1.9, 10)
    # a constant offset eps: L2 ratio = eps * sqrt(len / int (1+x)^2), max ratio = eps / max(1+x)
    eps = 1e-3
    yc = 1.0 + xc + eps
    lo, hi = 0.1, 1.9
    l2 = eps * math.sqrt((hi - lo) / (((1 + hi) ** 3 - (1 + lo) ** 3) / 3))
    # the trapezoid rule on 10 samples is exact for the constant numerator only
    den = np.sum(0.5 * ((1 + xc[1:]) ** 2 + (1 + xc[:-1]) ** 2) * np.diff(xc))
    assert harness.relative_l2(xc, yc, xf, yf) == pytest.approx(eps * math.sqrt((hi - lo) / den), rel=1e-9)
    assert harness.relative_l2(xc, yc, xf, yf) == pytest.approx(l2, rel=1e-2)
    assert harness.relative_linf(xc, yc, xf, yf) == pytest.approx(eps / (1 + hi), rel=1e-9)
    with pytest.raises(ValueError):
        harness.relative_l2(np.array([5.0, 6.0]), np.array([1.0, 1.0]), xf, yf)


@pytest.mark.parametrize("order", [1.0, 2.0])
def test_convergence_runner_recovers_the_order(tmp_path, order):
    exe = _stub(tmp_path, C=0.5, order=order)
    rows = harness.convergence(exe, tmp_path / "conv", nxs=(400, 800, 1600, 3200), time=8.0,
                               lx=8.0, material="water", energy=100)
    assert [r["nx"] for r in rows] == [400, 800, 1600]
    # every level is measured against the finest, so e_k is proportional to
    # h_k^p - h_f^p and the order between neighbours is
    # log((h0^p - hf^p) / (h1^p - hf^p)) / log 2
    hf = 8.0 / 3200
    for a, b in zip(rows, rows[1:]):
        want = math.log((a["h"] ** order - hf ** order) / (b["h"] ** order - hf ** order)) / math.log(2)
        assert a["order_to_next"] == pytest.approx(want, rel=2e-2)
    assert os.path.exists(tmp_path / "conv" / "nx_400" / "command.txt")
    cmd = open(tmp_path / "conv" / "nx_400" / "command.txt").read().split()
    assert cmd[1] == "400" and cmd[cmd.index("--energy-model") + 1] == "eq15"
    assert cmd[cmd.index("--data") + 1] == harness.data_dir("water")


def test_convergence_rejects_bad_levels(tmp_path):
    with pytest.raises(ValueError):
        harness.convergence(_stub(tmp_path), tmp_path, nxs=(1000, 500, 2000))
    with pytest.raises(ValueError):
        harness.convergence(_stub(tmp_path), tmp_path, nxs=(500, 1000))


def test_table2_runner_reports_metrics_and_paper_differences(tmp_path):
    exe = _stub(tmp_path)
    rows = harness.table2(exe, tmp_path / "t2", nx=2000, time=40.0)
    assert [(r["material"], r["energy"]) for r in rows] == list(harness.TABLE2_CASES)
    for r in rows:
        x, y = harness.load_curve(tmp_path / "t2" / f"{r['material']}_{r['energy']}" / "idd_output.txt")
        assert r["metrics_cm"] == metrics.analyze(x, y)
        for k in ("BP", "P90", "D90", "D20"):
            assert r["diff_mm"][k] == pytest.approx(10 * (r["metrics_cm"][k] - r["paper_cm"][k]))
    # re-reading the written files gives the same table
    again = harness.table2(exe, tmp_path / "t2", skip_run=True)
    assert [r["metrics_cm"] for r in again] == [r["metrics_cm"] for r in rows]


def test_solver_failure_is_reported(tmp_path):
    bad = tmp_path / "bad_solver"
    bad.write_text(f"#!{sys.executable}\nimport sys\nprint('Error: nope')\nsys.exit(1)\n")
    bad.chmod(bad.stat().st_mode | stat.S_IEXEC)
    with pytest.raises(RuntimeError, match="exited with 1"):
        harness.run_solver(bad, 10, [], tmp_path / "case")


def test_grid_variants_must_change_exactly_one_family():
    base = (2000, 20, 20, 20, 20)
    harness.check_variant("depth", "d", base, (1000, 20, 20, 20, 20))
    harness.check_variant("transverse", "t", base, (2000, 40, 40, 20, 20))
    harness.check_variant("angular", "a", base, (2000, 20, 20, 16, 16))
    with pytest.raises(ValueError, match="did not change"):
        harness.check_variant("angular", "a", base, (1000, 20, 20, 20, 20))
    with pytest.raises(ValueError, match="together with transverse"):
        harness.check_variant("depth", "d", base, (1000, 40, 40, 20, 20))
    with pytest.raises(ValueError, match="unknown variant axis"):
        harness.check_variant("energy", "e", base, (2000, 20, 20, 20, 20))
    with pytest.raises(ValueError, match="positive"):
        harness.check_variant("depth", "d", base, (0, 20, 20, 20, 20))


def test_grid_sensitivity_runner(tmp_path):
    exe = _stub(tmp_path)
    out = tmp_path / "gs"
    rows = harness.grid_sensitivity(
        exe, out, ("base", (2000, 20, 20, 20, 20)),
        [("depth", "half_dx", (1000, 20, 20, 20, 20)),
         ("transverse", "wide", (2000, 40, 40, 20, 20)),
         ("angular", "fine_angle", (2000, 20, 20, 30, 30))], time=40.0, energy=150)
    assert [r["label"] for r in rows] == ["base", "half_dx", "wide", "fine_angle"]
    base, half, wide, fine = rows
    assert base["norm_l2_vs_baseline"] == 0.0 and base["norm_linf_vs_baseline"] == 0.0
    # a pure scale of the dose leaves the Dmax-normalised curve unchanged
    assert wide["dmax"] == pytest.approx(1.2 * base["dmax"], rel=1e-12)
    assert wide["entrance"] == pytest.approx(1.2 * base["entrance"], rel=1e-12)
    assert wide["norm0"] == pytest.approx(base["norm0"], rel=1e-12)
    assert wide["norm_l2_vs_baseline"] < 1e-12 and wide["bp"] == base["bp"]
    # the added wave is seen; the depth variant samples the same curve at every
    # other baseline depth, so interpolation is exact there
    assert 1e-4 < fine["norm_l2_vs_baseline"] < 0.1
    assert half["norm_l2_vs_baseline"] < 1e-12 and abs(half["bp"] - base["bp"]) <= 2 * 40.0 / 1000
    x, y = harness.metrics.load_idd(out / "fine_angle" / "idd_output.txt")
    assert fine["entrance"] == y[0] and fine["dmax"] == y.max()
    cmd = open(out / "wide" / "command.txt").read().split()
    assert cmd[1] == "2000" and cmd[cmd.index("--ny") + 1] == "40" and cmd[cmd.index("--nz") + 1] == "40"
    csv_rows = list(csv.DictReader(open(out / "grid_sensitivity.csv")))
    assert [r["label"] for r in csv_rows] == ["base", "half_dx", "wide", "fine_angle"]
    md = (out / "grid_sensitivity.md").read_text().splitlines()
    assert md[0].startswith("| axis | label | grid |") and len(md) == 6
    assert md[4].startswith("| transverse | wide | 2000 / 40x40 / 20x20 |")
    # re-reading the written files gives the same rows
    again = harness.grid_sensitivity(
        exe, out, ("base", (2000, 20, 20, 20, 20)),
        [("depth", "half_dx", (1000, 20, 20, 20, 20)),
         ("transverse", "wide", (2000, 40, 40, 20, 20)),
         ("angular", "fine_angle", (2000, 20, 20, 30, 30))], skip_run=True)
    assert again == rows


def test_grid_sensitivity_checks_before_running(tmp_path):
    exe = _stub(tmp_path)
    with pytest.raises(ValueError, match="together with"):
        harness.grid_sensitivity(exe, tmp_path / "gs", ("base", (2000, 20, 20, 20, 20)),
                                 [("depth", "bad", (1000, 20, 20, 16, 16))])
    assert not (tmp_path / "gs").exists()
    with pytest.raises(ValueError, match="five positive"):
        harness.grid_sensitivity(exe, tmp_path / "gs", ("base", (2000, 0, 20, 20, 20)),
                                 [("depth", "d", (1000, 0, 20, 20, 20))])
    with pytest.raises(ValueError, match="unique"):
        harness.grid_sensitivity(exe, tmp_path / "gs", ("base", (2000, 20, 20, 20, 20)),
                                 [("depth", "base", (1000, 20, 20, 20, 20))])


def test_grid_sensitivity_cli(tmp_path, capsys):
    exe = _stub(tmp_path)
    rc = harness.main(["grid-sensitivity", "--exe", str(exe), "--outdir", str(tmp_path / "cli"),
                       "--baseline", "base", "1000", "20", "20", "20", "20",
                       "--variant", "angular", "coarse", "1000", "20", "20", "16", "16"])
    assert rc == 0
    text = capsys.readouterr().out
    assert "| angular | coarse | 1000 / 20x20 / 16x16 |" in text
    with pytest.raises(SystemExit):
        harness.main(["grid-sensitivity", "--exe", str(exe), "--outdir", str(tmp_path / "cli2"),
                      "--baseline", "base", "1000", "20", "20", "20", "20",
                      "--variant", "depth", "same", "1000", "20", "20", "20", "20"])
