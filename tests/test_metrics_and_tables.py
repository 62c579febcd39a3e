"""Harness restatements pinned by reference-held data.

* metrics.analyze (scripts/analyze_idd.py:21-69) against the BP/P90/D90/D20
  values that results/stage6_final_water230/summary.md:10-13 reports for the
  two golden IDD files.
* tables.synthetic_tables (scripts/generate_test_data.py:20-100) against the
  committed generator output data/{data_cross,cross_total}.txt (Ng=500,
  Lg=259), byte for byte.
"""
import os

import numpy as np
import pytest

from bfpn_gpu_amd import metrics
from bfpn_gpu_amd.tables import synthetic_tables, write_synthetic_tables
from conftest import GOLDEN

SUMMARY = {  # results/stage6_final_water230/summary.md:10-13
    "8_7": {"BP": 32.5600, "P90": 32.3066, "D90": 32.7443, "D20": 33.2153},
    "12_9": {"BP": 32.5600, "P90": 32.3051, "D90": 32.7437, "D20": 33.2150},
}


@pytest.mark.parametrize("tag", ["8_7", "12_9"])
def test_metrics_match_reference_summary(tag):
    x, y = metrics.load_idd(os.path.join(GOLDEN, "stage6_water230", f"{tag}_nx1000",
                                         "idd_output.txt"))
    m = metrics.analyze(x, y)
    for k, v in SUMMARY[tag].items():
        assert abs(m[k] - v) <= 5e-5, (k, m[k], v)


def test_metrics_edge_cases():
    x = np.linspace(0.1, 1.0, 10)
    y = np.array([1, 2, 3, 4, 10, 4, 3, 2, 1, 0.5], dtype=float)
    m = metrics.analyze(x, y)
    assert m["BP"] == pytest.approx(0.5)
    # left crossing of 0.9 between samples 4 (0.4) and 5 (1.0)
    assert m["P90"] == pytest.approx(0.4 + (0.9 - 0.4) / (1.0 - 0.4) * 0.1)
    assert m["D90"] == pytest.approx(0.5 + (0.9 - 1.0) / (0.4 - 1.0) * 0.1)
    # no distal crossing of 0.2 -> value at the crossing 0.3 -> 0.2
    assert m["D20"] == pytest.approx(0.7 + (0.2 - 0.3) / (0.2 - 0.3) * 0.1)
    mono = metrics.analyze(x, x)  # peak at the last sample: no distal crossing
    assert np.isnan(mono["D90"]) and np.isnan(mono["D20"])
    with pytest.raises(ValueError):
        metrics.analyze(x, -np.ones_like(x))


def test_synthetic_generator_is_byte_identical(tmp_path):
    write_synthetic_tables(500, 259.0, str(tmp_path))
    for name in ("data_cross.txt", "cross_total.txt"):
        ref = open(os.path.join(GOLDEN, "synthetic_ng500_lg259", name)).read()
        assert open(os.path.join(tmp_path, name)).read() == ref


def test_synthetic_generator_first_row_closed_form():
    """Row 0 of mu0 is 1.5 + 2 ln(E0/10) with E0 = 1 + dg/2 = 1.259 MeV."""
    mu, sigma = synthetic_tables(500, 259.0)
    assert mu[0, 0] == pytest.approx(1.5 + 2 * np.log(1.259 / 10), rel=1e-14)
    assert np.all(sigma >= 1e-10) and np.all(mu[:, 2] >= 1e-10)
