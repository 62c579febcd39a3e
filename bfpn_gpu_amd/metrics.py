"""Depth-dose metrics of an IDD curve (Bragg peak, P90, D90, D20).

Definitions follow scripts/analyze_idd.py:21-69 of the reference:
  BP   x at the sample argmax of the IDD,
  P90  proximal (left) crossing of 0.9 of the peak-normalised curve,
  D90  distal (right) crossing of 0.9, D20 distal crossing of 0.2,
with linear interpolation between samples and NaN when no crossing exists.
Paper Table 2 targets (deterministic column): scripts/analyze_idd.py:11-18.
"""
from __future__ import annotations

import math

import numpy as np

TABLE2 = {
    ("water", 50): {"BP": 2.150, "P90": 2.075, "D90": 2.205, "D20": 2.345},
    ("water", 100): {"BP": 7.560, "P90": 7.418, "D90": 7.671, "D20": 7.945},
    ("water", 230): {"BP": 32.540, "P90": 32.172, "D90": 32.828, "D20": 33.513},
    ("bone", 50): {"BP": 1.360, "P90": 1.319, "D90": 1.402, "D20": 1.488},
    ("bone", 100): {"BP": 4.770, "P90": 4.675, "D90": 4.837, "D20": 5.005},
    ("bone", 230): {"BP": 20.390, "P90": 20.168, "D90": 20.585, "D20": 21.023},
}


def _interp(x0, y0, x1, y1, level):
    if y1 == y0:
        return x0
    return x0 + (level - y0) / (y1 - y0) * (x1 - x0)


def proximal_crossing(x, y, peak, level):
    """Walk left from the peak to the first sample pair bracketing `level` from below."""
    for k in range(peak, 0, -1):
        lo, hi = y[k - 1], y[k]
        if (lo - level) * (hi - level) <= 0 and lo <= level <= hi:
            return _interp(x[k - 1], lo, x[k], hi, level)
    return math.nan


def distal_crossing(x, y, peak, level):
    """Walk right from the peak to the first sample pair bracketing `level` from above."""
    for k in range(peak, len(x) - 1):
        hi, lo = y[k], y[k + 1]
        if (hi - level) * (lo - level) <= 0 and hi >= level >= lo:
            return _interp(x[k], hi, x[k + 1], lo, level)
    return math.nan


def analyze(x, dose) -> dict:
    x = np.asarray(x, dtype=np.float64)
    dose = np.asarray(dose, dtype=np.float64)
    keep = np.isfinite(x) & np.isfinite(dose)
    x, dose = x[keep], dose[keep]
    if x.size == 0:
        raise ValueError("no finite IDD samples")
    if dose.max() <= 0:
        raise ValueError("IDD has no positive dose values")
    peak = int(np.argmax(dose))
    dmax = float(dose[peak])
    y = dose / dmax
    return {
        "BP": float(x[peak]),
        "P90": proximal_crossing(x, y, peak, 0.9),
        "D90": distal_crossing(x, y, peak, 0.9),
        "D20": distal_crossing(x, y, peak, 0.2),
        "Dmax": dmax,
    }


def load_idd(path):
    data = np.loadtxt(path, comments="#")
    if data.ndim != 2 or data.shape[1] < 2:
        raise ValueError(f"{path} must have at least two columns: x dose")
    return data[:, 0], data[:, 1]


def compare(idd, ref, band=0.10) -> dict:
    """Pointwise error of `idd` against `ref` where ref >= band * max(ref)."""
    idd = np.asarray(idd, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    m = ref >= band * ref.max()
    rel = np.abs(idd[m] - ref[m]) / ref[m]
    return {"max_rel": float(rel.max()), "mean_rel": float(rel.mean()), "points": int(m.sum())}
