"""Shared small cases for the GPU parity tests (sizes the oracle finishes in seconds)."""
import os

from bfpn_gpu_amd.abi import Case
from conftest import TABLES


def small_case(**over):
    """Water 230 MeV beam on a small but complete grid: every operator active,
    uneven y/z and mu/om sizes so that transposed indices show up."""
    kw = dict(Nx=1000, Ny=5, Nz=4, Ng=500, Nmu=5, Nom=4, Ly=8, Lz=6, Lu=2, Lv=2, energy=230,
              sigma_yz=0.6, material="water", data=os.path.join(TABLES, "water"), time=40,
              flags=dict(discrete_angle_delta=1, normalize_initial_mass=1, trapezoidal_yz=1))
    flags = over.pop("flags", {})
    kw.update(over)
    c = Case(**kw)
    c.flags.update(flags)
    return c


def golden_case(tag="8_7"):
    ny, nmu = {"8_7": (8, 7), "12_9": (12, 9)}[tag]
    return Case(Nx=1000, Ny=ny, Nz=ny, Ng=500, Nmu=nmu, Nom=nmu, Ly=8, Lz=8, Lu=2, Lv=2,
                energy=230, sigma_yz=0.3, material="water", data=os.path.join(TABLES, "water"),
                time=40, flags=dict(discrete_angle_delta=1, normalize_initial_mass=1,
                                    trapezoidal_yz=1))
