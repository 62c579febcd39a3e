"""Python view of the CPU FP64 oracle (oracle/bfpn_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing that
is measured or shipped.

Pinning status of each restated function is recorded in the docstrings of
tests/test_oracle_*.py; every function here is pinned by at least one
reference fixture, closed form, or brute-force check.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from bfpn_gpu_amd.abi import Case, Deltas, Grid, Phys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbfpn_oracle.so")
REF_SERIAL = os.path.join(HERE, "_ref", "serial_bfpn")

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile libbfpn_oracle.so with the recipe in oracle/Makefile (when it is
    missing or older than its source, or when forced)."""
    src = os.path.join(HERE, "bfpn_oracle.cpp")
    stale = os.path.exists(LIB_PATH) and os.path.getmtime(src) > os.path.getmtime(LIB_PATH)
    if force or stale or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "libbfpn_oracle.so"], cwd=HERE, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_last_error.restype = C.c_char_p
        L.orc_threads.restype = C.c_int
        L.orc_deltas.argtypes = [C.POINTER(Grid), C.POINTER(Deltas)]
        L.orc_stopping_power.argtypes = [C.c_char_p, C.c_int, C.c_double, _dp]
        L.orc_straggling.argtypes = [C.c_int, _dp]
        L.orc_transport_xs.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_load_tables.argtypes = [C.c_char_p, C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_int)]
        L.orc_omend.argtypes = [C.POINTER(Grid), _dp]
        L.orc_build_kernels.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                        _ip, _ip, _dp]
        L.orc_init_beam.argtypes = [C.POINTER(Grid), C.POINTER(Phys), _dp, _dp]
        L.orc_energy_half.argtypes = [_dp, C.c_void_p, _dp, _dp, C.c_void_p, C.c_double,
                                      C.c_double, C.c_double, C.c_int, C.c_longlong, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_int]
        L.orc_energy_half_legacy.argtypes = [_dp, C.c_void_p, _dp, _dp, C.c_void_p, C.c_double,
                                             C.c_double, C.c_double, C.c_int, C.c_longlong,
                                             C.c_int, C.c_void_p, C.c_int]
        L.orc_secondary_source.argtypes = [_dp, C.c_void_p, _dp, _dp, _dp, _dp, _ip, _ip, C.c_int,
                                           C.c_longlong, C.c_int, C.c_double]
        L.orc_secondary_source_factorized.argtypes = L.orc_secondary_source.argtypes
        L.orc_set_direct_source.argtypes = [C.c_int]
        L.orc_transport_half.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                         C.c_double, C.c_double, C.c_int]
        L.orc_angle_full.argtypes = [_dp, _dp, C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_double, C.c_double]
        L.orc_idd.argtypes = [_dp, C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp, _dp] + [C.c_int] * 4 + \
            [C.c_double] * 6 + [C.c_int, C.c_int]
        L.orc_idd.restype = C.c_double
        L.orc_energy_moments.argtypes = [_dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                         C.c_longlong, C.c_int] + [C.c_double] * 4 + [_dp]
        L.orc_spot_plane.argtypes = [_dp, C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp, _dp,
                                     C.c_int, C.c_longlong, C.c_int] + [C.c_double] * 4 + \
            [C.c_int, _dp]
        L.orc_create.argtypes = [C.POINTER(Grid), C.POINTER(Phys), C.c_char_p]
        L.orc_create.restype = C.c_void_p
        L.orc_create_layered.argtypes = [C.POINTER(Grid), C.POINTER(Phys), C.c_int, _dp,
                                         C.POINTER(C.c_char_p), _dp, C.POINTER(C.c_char_p)]
        L.orc_create_layered.restype = C.c_void_p
        L.orc_create_mapped.argtypes = [C.POINTER(Grid), C.POINTER(Phys), C.c_int,
                                        C.POINTER(C.c_char_p), _dp, C.POINTER(C.c_char_p), C.c_int,
                                        _dp, np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")]
        L.orc_create_mapped.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_num_steps.argtypes = [C.c_void_p, C.c_double]
        L.orc_advance.argtypes = [C.c_void_p, C.c_int]
        L.orc_history.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_cells.argtypes = [C.c_void_p]
        L.orc_cells.restype = C.c_size_t
        L.orc_mu_rows.argtypes = [C.c_void_p]
        L.orc_masses.argtypes = [C.c_void_p, _dp]
        L.orc_state.argtypes = [C.c_void_p, C.c_int]
        L.orc_state.restype = C.POINTER(C.c_double)
        L.orc_table.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_size_t)]
        L.orc_table.restype = C.POINTER(C.c_double)
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def threads() -> int:
    return lib().orc_threads()


def deltas(case: Case) -> Deltas:
    d = Deltas()
    g = case.grid()
    lib().orc_deltas(C.byref(g), C.byref(d))
    return d


# ---------------------------------------------------------------- tables ---
def stopping_power(material: str, Ng: int, Lg: float) -> np.ndarray:
    S = np.zeros(Ng + 1)
    lib().orc_stopping_power(material.encode(), Ng, Lg, S)
    return S


def straggling(Ng: int) -> np.ndarray:
    T = np.zeros(Ng)
    lib().orc_straggling(Ng, T)
    return T


def transport_xs(Ng: int, dg: float) -> np.ndarray:
    s = np.zeros(Ng)
    lib().orc_transport_xs(Ng, dg, s)
    return s


def load_tables(path: str, Ng: int, calibrate: bool = False):
    mu = np.zeros(3 * Ng)
    sig = np.zeros(Ng)
    rows = C.c_int(0)
    if lib().orc_load_tables(path.encode(), Ng, int(calibrate), mu, sig, C.byref(rows)):
        raise OSError(lib().orc_last_error().decode())
    return mu.reshape(Ng, 3), sig, rows.value


def omend(case: Case) -> np.ndarray:
    om = np.zeros(2 * case.nang)
    g = case.grid()
    lib().orc_omend(C.byref(g), om)
    return om


def build_kernels(Ng, dg, nang, mu, sigma, om):
    ke1 = np.zeros(Ng * Ng)
    ke2 = np.zeros(Ng * Ng)
    beg = np.zeros(Ng, np.int32)
    end = np.zeros(Ng, np.int32)
    kv = np.zeros(Ng * nang * nang)
    lib().orc_build_kernels(Ng, dg, nang, np.ascontiguousarray(mu, dtype=np.float64).ravel(),
                            np.ascontiguousarray(sigma, dtype=np.float64), om, ke1, ke2, beg, end, kv)
    return ke1.reshape(Ng, Ng), ke2.reshape(Ng, Ng), beg, end, kv.reshape(Ng, nang, nang)


def init_beam(case: Case):
    F = np.zeros(case.cells)
    m = np.zeros(4)
    g, p = case.grid(), case.phys()
    if lib().orc_init_beam(C.byref(g), C.byref(p), F, m):
        raise ValueError(lib().orc_last_error().decode())
    return F, m


# ------------------------------------------------------------- operators ---
def energy_half(F, f, S, sigma, T, h, dg, density, Ng, nyz, nang, Q1=None, Q2=None,
                straggling=False):
    """In place on float64 arrays F, f (f may be None)."""
    lib().orc_energy_half(F, _ptr(f), S, sigma, _ptr(T), h, dg, density, Ng, nyz, nang,
                          _ptr(Q1), _ptr(Q2), int(straggling))


def energy_half_legacy(F, f, S, sigma, T, h, dg, density, Ng, nyz, nang, Q1=None,
                       straggling=False):
    """Legacy (first-order upwind) energy step, in place; f may be None (lite)."""
    lib().orc_energy_half_legacy(F, _ptr(f), S, sigma, _ptr(T), h, dg, density, Ng, nyz, nang,
                                 _ptr(Q1), int(straggling))


def secondary_source(Fp, fp, ke1, ke2, kv, beg, end, Ng, nyz, nang, dg):
    Q = np.zeros_like(Fp)
    lib().orc_secondary_source(Fp, _ptr(fp), Q, np.ascontiguousarray(ke1).ravel(),
                               np.ascontiguousarray(ke2).ravel(), np.ascontiguousarray(kv).ravel(),
                               np.ascontiguousarray(beg, dtype=np.int32),
                               np.ascontiguousarray(end, dtype=np.int32), Ng, nyz, nang, dg)
    return Q


def secondary_source_factorized(Fp, fp, ke1, ke2, kv, beg, end, Ng, nyz, nang, dg):
    Q = np.zeros_like(Fp)
    lib().orc_secondary_source_factorized(
        Fp, _ptr(fp), Q, np.ascontiguousarray(ke1).ravel(), np.ascontiguousarray(ke2).ravel(),
        np.ascontiguousarray(kv).ravel(), np.ascontiguousarray(beg, dtype=np.int32),
        np.ascontiguousarray(end, dtype=np.int32), Ng, nyz, nang, dg)
    return Q


def set_direct_source(on: bool):
    """Make the full solve use the reference-order source loop (slow)."""
    lib().orc_set_direct_source(int(on))


def transport_half(F, om, Ny, Nz, Ng, nang, dy, dz, h, preserve_sign):
    lib().orc_transport_half(F, om, Ny, Nz, Ng, nang, dy, dz, h, int(preserve_sign))


def angle_full(F, sig_tr, nyz, Ng, Nmu, Nom, du, dv, dt, density):
    lib().orc_angle_full(F, np.ascontiguousarray(sig_tr, dtype=np.float64), nyz, Ng, Nmu, Nom,
                         du, dv, dt, density)


def idd(F, f, F1, f1, S, sigma, en, Ng, Ny, Nz, nang, dg, du, dv, dy, dz, density,
        trapezoidal=False, catastrophic=True) -> float:
    return lib().orc_idd(F, _ptr(f), _ptr(F1), _ptr(f1), S, sigma, en, Ng, Ny, Nz, nang, dg, du,
                         dv, dy, dz, density, int(trapezoidal), int(catastrophic))


def energy_moments(F, f, F1, f1, Ng, nyz, nang, du, dv, dy, dz):
    """Eq.25 moments (energyMomentKernel): array [Ng, 2] of (int psi^1, int psi^2)."""
    out = np.zeros(2 * Ng)
    lib().orc_energy_moments(F, _ptr(f), _ptr(F1), _ptr(f1), Ng, nyz, nang, du, dv, dy, dz, out)
    return out.reshape(Ng, 2)


def spot_plane(F, f, F1, f1, S, sigma, en, Ng, nyz, nang, dg, du, dv, density, catastrophic=True):
    """Y-Z dose plane (spotDosePlaneKernel): nyz values, s = iz (Ny+1) + iy."""
    out = np.zeros(nyz)
    lib().orc_spot_plane(F, _ptr(f), _ptr(F1), _ptr(f1), S, sigma, en, Ng, nyz, nang, dg, du, dv,
                         density, int(catastrophic), out)
    return out


# ------------------------------------------------------------------ solver ---
TABLE_NAMES = ("S", "sigma", "T", "sig_tr", "en", "ker_e1", "ker_e2", "ker_v", "omend", "mu",
               "ker_beg", "ker_end")


class Solver:
    """Whole in-core eq15 solve on the CPU (BFPnSolver, FP64)."""

    def __init__(self, case: Case, layers=None, mapped=None):
        """layers: optional [(x_end_cm, material, density, table_dir), ...] in
        increasing depth; the last layer extends to any depth.
        mapped: optional material map (materials, slab_end, map):
        materials [(material, density, table_dir), ...], slab_end the depths
        (cm) ending each slab (the last: any depth), map an integer array
        (slabs, (Ny+1)(Nz+1)) of material indices, node s = iz (Ny+1) + iy."""
        self.case = case
        g, p = case.grid(), case.phys()
        if mapped is not None:
            materials, slab_end, mp = mapped
            n = len(materials)
            names = (C.c_char_p * n)(*[m[0].encode() for m in materials])
            dens = np.array([float(m[1]) for m in materials])
            paths = (C.c_char_p * n)(*[m[2].encode() for m in materials])
            ends = np.ascontiguousarray(slab_end, dtype=np.float64)
            mp = np.ascontiguousarray(mp, dtype=np.uint8)
            self._h = lib().orc_create_mapped(C.byref(g), C.byref(p), n, names, dens, paths,
                                              len(ends), ends, mp.ravel())
        elif layers:
            n = len(layers)
            x_end = np.array([float(l[0]) for l in layers])
            mats = (C.c_char_p * n)(*[l[1].encode() for l in layers])
            dens = np.array([float(l[2]) for l in layers])
            paths = (C.c_char_p * n)(*[l[3].encode() for l in layers])
            self._h = lib().orc_create_layered(C.byref(g), C.byref(p), n, x_end, mats, dens, paths)
        else:
            self._h = lib().orc_create(C.byref(g), C.byref(p), case.data.encode())
        if not self._h:
            raise OSError(lib().orc_last_error().decode())

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def cells(self) -> int:
        return lib().orc_cells(self._h)

    def num_steps(self, t_final=None) -> int:
        return lib().orc_num_steps(self._h, self.case.time if t_final is None else t_final)

    def advance(self, n: int = 1):
        lib().orc_advance(self._h, n)

    def history(self):
        n = lib().orc_history(self._h, None, None, 0)
        idd = np.zeros(n)
        px = np.zeros(n)
        lib().orc_history(self._h, _ptr(idd), _ptr(px), n)
        return idd, px

    def state(self, which: int) -> np.ndarray:
        """Writable view of state array `which` (0 F, 1 f, 2 F1, 3 f1)."""
        p = lib().orc_state(self._h, which)
        return np.ctypeslib.as_array(p, shape=(self.cells,))

    def table(self, which) -> np.ndarray:
        if isinstance(which, str):
            which = TABLE_NAMES.index(which)
        n = C.c_size_t(0)
        p = lib().orc_table(self._h, which, C.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def masses(self):
        m = np.zeros(4)
        lib().orc_masses(self._h, m)
        return m

    def mu_rows(self) -> int:
        return lib().orc_mu_rows(self._h)

    def run(self, t_final=None):
        self.advance(self.num_steps(t_final))
        return self.history()
