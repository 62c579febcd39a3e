"""ctypes mirrors of the POD structs declared in include/bfpn.h.

GridParams / PhysicsParams of the reference (src/bfpm_solver.cuh:21-62) map
onto ``Grid`` / ``Phys``; ``Deltas`` is the result of
GridParams::computeDeltas (src/bfpm_solver.cu:1116-1126).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field


class Grid(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("Nx", "Ny", "Nz", "Ng", "Nmu", "Nom")] + [
        (n, C.c_double) for n in ("Lx", "Ly", "Lz", "Lu", "Lv", "Lg")
    ]


class Phys(C.Structure):
    _fields_ = [
        ("sig_E", C.c_double),
        ("a_1", C.c_double),
        ("a_2", C.c_double),
        ("beam_energy", C.c_double),
        ("density", C.c_double),
        ("energy_density", C.c_double),
        ("material_name", C.c_char * 32),
    ] + [
        (n, C.c_int)
        for n in (
            "primary_only",
            "eq15_straggling",
            "energy_only",
            "no_transport",
            "no_angle",
            "no_spatial_clipping",
            "discrete_angle_delta",
            "normalize_initial_mass",
            "trapezoidal_yz",
            "dose_catastrophic_loss",
            "calibrate_cross_sections",
            "save_energy_moments",
            "legacy_energy",
            "lite_memory",
        )
    ]


class Layer(C.Structure):
    """bfpn_layer: material tables and density up to depth x_end (cm)."""
    _fields_ = [("x_end", C.c_double), ("material", C.c_char_p), ("density", C.c_double),
                ("data_path", C.c_char_p)]


class Deltas(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dt", "dy", "dz", "dg", "du", "dv")]


MATERIAL_DENSITY = {"water": 1.0, "bone": 1.757, "air": 0.001205}  # src/main.cu:126-134


@dataclass
class Case:
    """One solver configuration, with the reference CLI's defaults
    (src/main.cu:46-90, include/config.hpp:21-35)."""

    Nx: int = 4000
    Ny: int = 10
    Nz: int = 10
    Ng: int = 500
    Nmu: int = 11
    Nom: int = 11
    Ly: float = 1.0
    Lz: float = 1.0
    Lu: float = 1.0
    Lv: float = 1.0
    Lg: float = 259.0
    Lx: float = 40.0
    energy: float = 230.0
    sigma_e: float = 1.0
    sigma_yz: float = 0.1
    material: str = ""
    density: float | None = None
    data: str = "./data"
    time: float = 40.0
    flags: dict = field(default_factory=dict)

    def grid(self) -> Grid:
        return Grid(self.Nx, self.Ny, self.Nz, self.Ng, self.Nmu, self.Nom,
                    self.Lx, self.Ly, self.Lz, self.Lu, self.Lv, self.Lg)

    def phys(self) -> Phys:
        rho = self.density
        if rho is None:
            rho = MATERIAL_DENSITY.get(self.material, 1.0)
        p = Phys()
        p.sig_E = self.sigma_e
        p.a_1 = 1.0 / 2.0 / self.sigma_yz / self.sigma_yz      # src/main.cu:294
        p.a_2 = 1.0 / 2.0 / 1e-6 / 1e-6                        # src/main.cu:295
        p.beam_energy = self.energy
        p.density = rho
        p.energy_density = rho
        p.material_name = self.material.encode()
        p.dose_catastrophic_loss = 1                           # src/main.cu:82
        for k, v in self.flags.items():
            setattr(p, k, int(bool(v)))
        return p

    @property
    def nyz(self) -> int:
        return (self.Ny + 1) * (self.Nz + 1)

    @property
    def nang(self) -> int:
        return self.Nmu * self.Nom

    @property
    def cells(self) -> int:
        return self.Ng * self.nyz * self.nang

    @property
    def dt(self) -> float:
        return self.Lx / self.Nx

    def num_steps(self, t_final: float | None = None) -> int:
        t = self.time if t_final is None else t_final
        return int(math.ceil(t / self.dt - 1e-12))             # src/bfpm_solver.cu:2178
