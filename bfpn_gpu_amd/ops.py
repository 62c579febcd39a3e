"""Operator-level calls (the reference's sub-solver seam) on DeviceArrays.

Each function mirrors one reference operator; state arrays are DeviceArrays
of float32 or float64 (the dtype selects the _f32 / _f64 entry point), table
arrays are float64 DeviceArrays, except `sig_tr` of angle_full which is a
host numpy array (the CN factors are built on the host).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceArray, Workspace, check, lib

_default_ws = None


def _ws(ws):
    global _default_ws
    if ws is not None:
        return ws
    if _default_ws is None:
        _default_ws = Workspace()
    return _default_ws


def _suffix(a: DeviceArray) -> str:
    if a.dtype == np.float32:
        return "f32"
    if a.dtype == np.float64:
        return "f64"
    raise TypeError(f"state arrays must be float32 or float64, got {a.dtype}")


def _p(a):
    return None if a is None else a.ptr


def energy_half(F, f, S, sigma, T, h, dg, density, Ng, nyz, nang, Q1=None, straggling=False,
                ws=None):
    """EnergyDepositionSolver::solve, eq15 (src/energy_solver.cu:444-506), in place."""
    fn = getattr(lib(), f"bfpn_energy_half_{_suffix(F)}")
    check(fn(_ws(ws).ptr, F.ptr, f.ptr, S.ptr, sigma.ptr, _p(T), h, dg, density, Ng, nyz, nang,
             _p(Q1), None, int(straggling), None))


def secondary_source(Fp, fp, Q, ke1, ke2, kv, beg, end, Ng, nyz, nang, dg, ws=None,
                     prepared=False):
    """secondarySourceFromPrimaryKernel (src/energy_solver.cu:287-333) into Q.

    The library caches the source operands by table pointer; unless
    `prepared` says that secondary_source_prepare ran on these very tables
    (which must not have changed since), they are rebuilt first, because a
    freed table's address can come back with other contents."""
    if not prepared:
        secondary_source_prepare(ke1, ke2, kv, beg, end, Ng, nang, dg, Fp.dtype, ws)
    fn = getattr(lib(), f"bfpn_secondary_source_{_suffix(Fp)}")
    check(fn(_ws(ws).ptr, Fp.ptr, fp.ptr, Q.ptr, ke1.ptr, ke2.ptr, kv.ptr, beg.ptr, end.ptr, Ng,
             nyz, nang, dg, None))


def secondary_source_prepare(ke1, ke2, kv, beg, end, Ng, nang, dg, dtype=np.float32, ws=None):
    """Builds the source operands once (the kernel upload of BFPnSolver::initialize,
    src/bfpm_solver.cu:1638-1813); later secondary_source calls with the same
    tables reuse them."""
    suffix = "f32" if np.dtype(dtype) == np.float32 else "f64"
    fn = getattr(lib(), f"bfpn_secondary_source_prepare_{suffix}")
    check(fn(_ws(ws).ptr, ke1.ptr, ke2.ptr, kv.ptr, beg.ptr, end.ptr, Ng, nang, dg, None))


def workspace_stats(ws=None):
    """(host-side operand builds so far, device bytes held) of a workspace."""
    builds, nbytes = C.c_int(0), C.c_size_t(0)
    check(lib().bfpn_workspace_stats(_ws(ws).ptr, C.byref(builds), C.byref(nbytes)))
    return builds.value, nbytes.value


def transport_half(F, omend, Ny, Nz, Ng, nang, dy, dz, h, preserve_sign, ws=None):
    """SpatialTransportSolver::solveEq18 (src/spatial_transport.cu:202-262), in place."""
    fn = getattr(lib(), f"bfpn_transport_half_{_suffix(F)}")
    check(fn(_ws(ws).ptr, F.ptr, omend.ptr, Ny, Nz, Ng, nang, dy, dz, h, int(preserve_sign), None))


def angle_full(F, sig_tr, nyz, Ng, Nmu, Nom, du, dv, dt, density, ws=None):
    """AngleDiffusionSolver::solveEq19 (src/angle_diffusion.cu:420-483), in place."""
    fn = getattr(lib(), f"bfpn_angle_full_{_suffix(F)}")
    st = np.ascontiguousarray(sig_tr, dtype=np.float64)
    check(fn(_ws(ws).ptr, F.ptr, C.c_void_p(st.ctypes.data), nyz, Ng, Nmu, Nom, du, dv, dt,
             density, None))


def idd(F, f, F1, f1, S, sigma, en, Ng, Ny, Nz, nang, dg, du, dv, dy, dz, density,
        trapezoidal=False, catastrophic=True, ws=None) -> float:
    """computeIntegratedDepthDose (src/bfpm_solver.cu:3473-3502)."""
    fn = getattr(lib(), f"bfpn_idd_{_suffix(F)}")
    out = C.c_double(0.0)
    check(fn(_ws(ws).ptr, F.ptr, _p(f), _p(F1), _p(f1), S.ptr, sigma.ptr, en.ptr, Ng, Ny, Nz, nang,
             dg, du, dv, dy, dz, density, int(trapezoidal), int(catastrophic), C.byref(out), None))
    return out.value
