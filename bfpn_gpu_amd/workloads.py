"""Seeded beamlet and serving workloads (batched and served solves).

beamlet_batch: pencil beams with energy ~ U[70, 230] MeV, sigma_yz ~ U[0.2, 0.5] cm,
  Ny = Nz in {20, 40} over 8x8 cm, 20x20 directions, Ng = 300 on [1, 260] MeV,
  water (rho 1.0) or bone (rho 1.85) with equal probability, dt = 0.02 cm and a
  depth of 1.2x the CSDA range computed from the solver's own S(E).
serving_requests: Poisson arrivals at `rate` per second over `duration` s with
  energy ~ U[70, 230] MeV, Ny = Nz in {20, 40, 80}, otherwise as above.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .abi import Case
from .tables import material_dir

DENSITY = {"water": 1.0, "bone": 1.85}


def stopping_power(material: str, Ng: int, Lg: float) -> np.ndarray:
    S = np.zeros(Ng + 1)
    check(lib().bfpn_stopping_power(material.encode(), Ng, Lg, C.c_void_p(S.ctypes.data)))
    return S


def csda_range(energy: float, material: str, rho: float, Ng: int = 2000, Lg: float = 259.0) -> float:
    """Continuous-slowing-down range in cm: integral of dE / (rho S(E)) from 1 MeV."""
    S = stopping_power(material, Ng, Lg)
    dg = Lg / Ng
    E = np.concatenate([[1.0], (np.arange(1, Ng) - 0.5) * dg + 1.0, [Lg + 1.0]])
    m = E <= energy
    Em, Sm = E[m], S[m]
    return float(np.trapezoid(1.0 / (rho * Sm), Em) + (energy - Em[-1]) / (rho * Sm[-1]))


def beam_case(energy, sigma_yz, ny, material, depth, Ng=300, Nmu=20, Nom=20):
    return Case(Nx=2000, Ny=ny, Nz=ny, Ng=Ng, Nmu=Nmu, Nom=Nom, Ly=8.0, Lz=8.0, Lu=2.0, Lv=2.0,
                energy=energy, sigma_yz=sigma_yz, material=material, density=DENSITY[material],
                data=material_dir(material), time=min(depth, 40.0),
                flags=dict(eq15_straggling=1))


def beamlet_batch(n: int = 64, seed: int = 20261202, Ng: int = 300):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(n):
        E = float(rng.uniform(70.0, 230.0))
        sig = float(rng.uniform(0.2, 0.5))
        ny = int(rng.choice([20, 40]))
        mat = "water" if rng.random() < 0.5 else "bone"
        depth = 1.2 * csda_range(E, mat, DENSITY[mat])
        cases.append(beam_case(E, sig, ny, mat, depth, Ng=Ng))
    return cases


def serving_requests(rate: float, duration: float, seed: int = 20261202, Ng: int = 300,
                     sizes=(20, 40, 80)):
    """(arrival_s, Case) pairs of a Poisson stream."""
    rng = np.random.default_rng(seed + int(rate * 1000))
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out
        E = float(rng.uniform(70.0, 230.0))
        mat = "water" if rng.random() < 0.5 else "bone"
        ny = int(rng.choice(list(sizes)))
        depth = 1.2 * csda_range(E, mat, DENSITY[mat])
        out.append((t, beam_case(E, 0.3, ny, mat, depth, Ng=Ng)))
