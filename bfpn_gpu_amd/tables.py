"""Synthetic cross-section tables for any (Ng, Lg).

Restates generate_cross_sections of scripts/generate_test_data.py:20-100 of
the reference (the generator behind its data/ directory): per group centre E
  mu0 = 1.5 + 2 ln(E/10)                (angular kernel shape)
  mu1 = 15 + 5 sqrt(E/100)              (angular kernel width)
  mu2 = dg * E / S_col(E)               (energy kernel length, S_col Bethe-like)
  sigma_c = 0.9 * 2 pi hbarc^2/alpha^2/4 * Zt^2/E^2 * ln(1+(200/E)^2)/1000
            minus 0.0002*{0.5,1.0,...} on groups 200..242,
both clipped below at 1e-10, written with '%.10e' like the reference.
The committed reference output (Ng=500, Lg=259) pins this restatement.
"""
from __future__ import annotations

import os

import numpy as np

_ME = 9.10956e-31
_MP = 1.673e-27
_ALPHA = 1 / 137.0
_HBARC = 197.3
_C = 2.998e8


def synthetic_tables(Ng: int, Lg: float):
    dg = Lg / Ng
    E = np.arange(Ng) * dg + dg / 2 + 1.0
    mu = np.zeros((Ng, 3))
    sigma = np.zeros(Ng)
    Zt = 3
    steps = np.arange(0.5, 22, 0.5)
    for i, e in enumerate(E):
        if e < 938.3:
            v = np.sqrt(2 * e * 1.6e-13 / _MP)
        else:
            gam = e / 938.3 + 1
            v = _C * np.sqrt(1 - 1 / gam**2)
        b2 = (v / _C) ** 2
        mu[i, 0] = 1.5 + 2.0 * np.log(e / 10.0)
        mu[i, 1] = 15.0 + 5.0 * np.sqrt(e / 100.0)
        s_col = 0.170 / b2 * (np.log(1.02e6 * b2 / (1 - b2)) - b2 - 4.31) + 0.02
        mu[i, 2] = (e / s_col if s_col > 0 else 1000.0) * dg
        sn = 2 * np.pi * _HBARC**2 / _ALPHA**2 / 4 * Zt**2 / e**2
        sn *= np.log(1 + (200 / e) ** 2) / 1000
        sigma[i] = sn * 0.9
        if 200 <= i < 250 and i - 200 < len(steps):
            sigma[i] -= 0.0002 * steps[i - 200]
    sigma = np.maximum(sigma, 1e-10)
    mu[:, 2] = np.maximum(mu[:, 2], 1e-10)
    return mu, sigma


def write_synthetic_tables(Ng: int, Lg: float, out_dir: str) -> str:
    os.makedirs(out_dir, exist_ok=True)
    mu, sigma = synthetic_tables(Ng, Lg)
    np.savetxt(os.path.join(out_dir, "data_cross.txt"), mu, fmt="%.10e")
    np.savetxt(os.path.join(out_dir, "cross_total.txt"), sigma, fmt="%.10e")
    return out_dir


TABLE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables")


def material_dir(material: str) -> str:
    """Directory of the shipped material tables (water, bone, air, synthetic)."""
    return os.path.join(TABLE_DIR, material)
