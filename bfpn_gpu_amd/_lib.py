"""ctypes binding of libbfpn_hip.so (include/bfpn.h)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import Case, Deltas, Grid, Layer, Phys

PKG = os.path.dirname(os.path.abspath(__file__))
_LIB_FILE = os.path.join(PKG, "libbfpn_hip.so")

FP32 = 0
FP64 = 1

_lib = None


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_active", C.c_int), ("slice_steps", C.c_int),
                ("precision", C.c_int), ("fill_cells", C.c_size_t)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("active", C.c_int), ("queued", C.c_int), ("peak_active", C.c_int),
                ("cached_tables", C.c_int), ("bytes_in_use", C.c_size_t)]


class BfpnError(RuntimeError):
    """A bfpn_* call returned a non-zero status."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"bfpn status {status}: {msg}")
        self.status = status


def lib_path() -> str:
    return _LIB_FILE


def lib():
    """Load libbfpn_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_FILE):
        raise ImportError(
            f"{_LIB_FILE} is missing: build it with `python -m bfpn_gpu_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(_LIB_FILE)
    vp, i, d, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
    L.bfpn_last_error.restype = C.c_char_p
    L.bfpn_version.restype = C.c_char_p
    L.bfpn_device_count.argtypes = [C.POINTER(C.c_int)]
    L.bfpn_device_info.argtypes = [i, C.c_char_p, sz, C.POINTER(sz), C.POINTER(C.c_int)]
    L.bfpn_malloc.argtypes = [C.POINTER(vp), sz]
    L.bfpn_free.argtypes = [vp]
    L.bfpn_memcpy_h2d.argtypes = [vp, vp, sz]
    L.bfpn_memcpy_d2h.argtypes = [vp, vp, sz]
    L.bfpn_memset.argtypes = [vp, i, sz]
    L.bfpn_device_copy.argtypes = [vp, vp, sz]
    L.bfpn_mem_info.argtypes = [C.POINTER(sz), C.POINTER(sz)]
    L.bfpn_workspace_create.argtypes = [C.POINTER(vp)]
    L.bfpn_workspace_destroy.argtypes = [vp]
    L.bfpn_workspace_stats.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(sz)]
    for suffix in ("f32", "f64"):
        getattr(L, f"bfpn_energy_half_{suffix}").argtypes = [vp, vp, vp, vp, vp, vp, d, d, d, i, i, i,
                                                             vp, vp, i, vp]
        getattr(L, f"bfpn_secondary_source_{suffix}").argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                                                  i, i, i, d, vp]
        getattr(L, f"bfpn_secondary_source_prepare_{suffix}").argtypes = [vp, vp, vp, vp, vp, vp, i,
                                                                          i, d, vp]
        getattr(L, f"bfpn_transport_half_{suffix}").argtypes = [vp, vp, vp, i, i, i, i, d, d, d, i, vp]
        getattr(L, f"bfpn_angle_full_{suffix}").argtypes = [vp, vp, vp, i, i, i, i, d, d, d, d, vp]
        getattr(L, f"bfpn_idd_{suffix}").argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i, i, i, i, d, d,
                                                     d, d, d, d, i, i, C.POINTER(d), vp]
    L.bfpn_create.argtypes = [C.POINTER(Grid), C.POINTER(Phys), C.c_char_p, i, C.POINTER(vp)]
    L.bfpn_create_layered.argtypes = [C.POINTER(Grid), C.POINTER(Phys), C.POINTER(Layer), i, i,
                                      C.POINTER(vp)]
    L.bfpn_create_mapped.argtypes = [C.POINTER(Grid), C.POINTER(Phys), C.POINTER(Layer), i, vp,
                                     i, vp, i, C.POINTER(vp)]
    L.bfpn_destroy.argtypes = [vp]
    L.bfpn_get_deltas.argtypes = [vp, C.POINTER(Deltas)]
    L.bfpn_cells.argtypes = [vp]
    L.bfpn_cells.restype = sz
    L.bfpn_device_bytes.argtypes = [vp]
    L.bfpn_device_bytes.restype = sz
    L.bfpn_num_steps.argtypes = [vp, d]
    L.bfpn_get_table.argtypes = [vp, i, vp, sz]
    L.bfpn_advance.argtypes = [vp, i]
    L.bfpn_history.argtypes = [vp, vp, vp, i, C.POINTER(C.c_int)]
    L.bfpn_run.argtypes = [vp, d, vp, vp, i, C.POINTER(C.c_int)]
    L.bfpn_read_state.argtypes = [vp, i, vp, sz]
    L.bfpn_write_state.argtypes = [vp, i, vp, sz]
    L.bfpn_stream.argtypes = [vp]
    L.bfpn_stream.restype = vp
    L.bfpn_set_profiling.argtypes = [vp, i]
    L.bfpn_set_overlap.argtypes = [vp, i]
    L.bfpn_set_group_skip.argtypes = [vp, i]
    L.bfpn_active_groups.argtypes = [vp, C.POINTER(i), C.POINTER(i)]
    L.bfpn_energy_moments.argtypes = [vp, vp, i]
    L.bfpn_spot_plane.argtypes = [vp, vp, C.c_size_t]
    L.bfpn_phase_stats.argtypes = [vp, vp, vp, vp]
    L.bfpn_stopping_power.argtypes = [C.c_char_p, i, d, vp]
    L.bfpn_server_create.argtypes = [C.POINTER(ServerOpts), C.POINTER(vp)]
    L.bfpn_server_submit.argtypes = [vp, C.POINTER(Grid), C.POINTER(Phys), C.c_char_p, d,
                                     C.POINTER(C.c_uint64)]
    L.bfpn_server_submit_layered.argtypes = [vp, C.POINTER(Grid), C.POINTER(Phys),
                                            C.POINTER(Layer), i, d, C.POINTER(C.c_uint64)]
    L.bfpn_server_submit_mapped.argtypes = [vp, C.POINTER(Grid), C.POINTER(Phys),
                                            C.POINTER(Layer), i, vp, i, vp, d,
                                            C.POINTER(C.c_uint64)]
    L.bfpn_server_poll.argtypes = [vp, C.c_uint64]
    L.bfpn_server_wait.argtypes = [vp, C.c_uint64, d, vp, i, C.POINTER(C.c_int), vp]
    L.bfpn_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
    L.bfpn_server_destroy.argtypes = [vp]
    L.bfpn_server_estimate.argtypes = [C.POINTER(Grid), C.POINTER(Phys), i, C.POINTER(sz),
                                       C.POINTER(sz)]
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        raise BfpnError(status, lib().bfpn_last_error().decode())


def stopping_power(material: str, Ng: int, Lg: float) -> np.ndarray:
    """The solver's stopping-power table S (MeV cm^2/g, Ng + 1 points:
    E = 1, (i - 1/2) Lg/Ng + 1, 1 + Lg), computed on the host."""
    S = np.zeros(Ng + 1)
    check(lib().bfpn_stopping_power(material.encode(), Ng, Lg, _addr(S)))
    return S


def device_count() -> int:
    n = C.c_int(0)
    try:
        check(lib().bfpn_device_count(C.byref(n)))
    except BfpnError:
        return 0
    return n.value


def _addr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class DeviceArray:
    """A device allocation holding a copy of a numpy array (dtype kept)."""

    def __init__(self, host: np.ndarray | None = None, *, shape=None, dtype=np.float64):
        if host is not None:
            host = np.ascontiguousarray(host)
            shape, dtype = host.shape, host.dtype
        self.shape = tuple(np.atleast_1d(shape)) if shape is not None else (0,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p(0)
        check(lib().bfpn_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p
        if host is not None:
            self.upload(host)

    def upload(self, host: np.ndarray) -> None:
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes, "size mismatch"
        check(lib().bfpn_memcpy_h2d(self.ptr, _addr(host), self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(lib().bfpn_memcpy_d2h(_addr(out), self.ptr, self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr and self.ptr.value:
            lib().bfpn_free(self.ptr)
            self.ptr = C.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Workspace:
    """Scratch owner for operator-level calls (bfpn_workspace)."""

    def __init__(self):
        p = C.c_void_p(0)
        check(lib().bfpn_workspace_create(C.byref(p)))
        self.ptr = p

    def __del__(self):
        try:
            if self.ptr:
                lib().bfpn_workspace_destroy(self.ptr)
        except Exception:
            pass


TABLES = ("S", "sigma", "T", "sig_tr", "en", "ker_e1", "ker_e2", "ker_v", "omend", "mu",
          "ker_beg", "ker_end")


class Solver:
    """One beam: tables, state and depth march on the GPU (bfpn_ctx)."""

    def __init__(self, case: Case, precision: int = FP32, layers=None, mapped=None):
        """layers: optional depth layers [(x_end_cm, material, density, table_dir), ...]
        in increasing depth, the last extending to any depth (bfpn_create_layered).
        mapped: optional material map (materials, slab_end, map) (bfpn_create_mapped):
        materials [(material, density, table_dir), ...], slab_end the depths (cm)
        ending each slab (the last: any depth), map an integer array
        (slabs, (Ny+1)(Nz+1)) of material indices, node s = iz (Ny+1) + iy.
        The beam and flags come from `case`."""
        self.case = case
        self.precision = precision
        self.layers = list(layers) if layers else None
        g, p = case.grid(), case.phys()
        h = C.c_void_p(0)
        if mapped is not None:
            materials, slab_end, mp = mapped
            n = len(materials)
            spec = (Layer * n)(*[Layer(0.0, m.encode(), float(rho), path.encode())
                                 for m, rho, path in materials])
            ends = np.ascontiguousarray(slab_end, dtype=np.float64)
            mp = np.ascontiguousarray(mp, dtype=np.uint8)
            if mp.size != len(ends) * case.nyz:
                raise ValueError("the map needs (Ny+1)(Nz+1) material indices per slab")
            check(lib().bfpn_create_mapped(C.byref(g), C.byref(p), spec, n, _addr(ends), len(ends),
                                           _addr(mp), precision, C.byref(h)))
        elif self.layers:
            n = len(self.layers)
            spec = (Layer * n)(*[Layer(float(x), m.encode(), float(rho), path.encode())
                                 for x, m, rho, path in self.layers])
            check(lib().bfpn_create_layered(C.byref(g), C.byref(p), spec, n, precision,
                                            C.byref(h)))
        else:
            check(lib().bfpn_create(C.byref(g), C.byref(p), case.data.encode(), precision,
                                    C.byref(h)))
        self._h = h
        d = Deltas()
        check(lib().bfpn_get_deltas(self._h, C.byref(d)))
        self.deltas = d

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().bfpn_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def cells(self) -> int:
        return lib().bfpn_cells(self._h)

    @property
    def device_bytes(self) -> int:
        return lib().bfpn_device_bytes(self._h)

    def num_steps(self, t_final=None) -> int:
        return lib().bfpn_num_steps(self._h, self.case.time if t_final is None else t_final)

    def table(self, which) -> np.ndarray:
        if isinstance(which, str):
            which = TABLES.index(which)
        c = self.case
        sizes = [c.Ng + 1, c.Ng, c.Ng, c.Ng, c.Ng + 1, c.Ng * c.Ng, c.Ng * c.Ng,
                 c.Ng * c.nang * c.nang, 2 * c.nang, 3 * c.Ng, c.Ng, c.Ng]
        out = np.zeros(sizes[which])
        check(lib().bfpn_get_table(self._h, which, _addr(out), out.size))
        return out

    def advance(self, n: int = 1) -> None:
        check(lib().bfpn_advance(self._h, n))

    def history(self):
        n = C.c_int(0)
        check(lib().bfpn_history(self._h, None, None, 0, C.byref(n)))
        idd = np.zeros(max(n.value, 1))
        px = np.zeros(max(n.value, 1))
        check(lib().bfpn_history(self._h, _addr(idd), _addr(px), n.value, C.byref(n)))
        return idd[: n.value], px[: n.value]

    def run(self, t_final=None):
        self.advance(self.num_steps(t_final))
        return self.history()

    def state(self, which: int) -> np.ndarray:
        out = np.zeros(self.cells)
        check(lib().bfpn_read_state(self._h, which, _addr(out), out.size))
        return out

    def set_state(self, which: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().bfpn_write_state(self._h, which, _addr(v), v.size))

    def set_profiling(self, on: bool = True) -> None:
        check(lib().bfpn_set_profiling(self._h, int(on)))

    def energy_moments(self) -> np.ndarray:
        """Eq.25 moments of the current state: [Ng, 2] = (int psi^1, int psi^2) per group."""
        out = np.zeros(2 * self.case.Ng)
        check(lib().bfpn_energy_moments(self._h, _addr(out), out.size))
        return out.reshape(self.case.Ng, 2)

    def spot_plane(self) -> np.ndarray:
        """Y-Z dose plane of the current state, flat: s = iz (Ny+1) + iy."""
        out = np.zeros((self.case.Ny + 1) * (self.case.Nz + 1))
        check(lib().bfpn_spot_plane(self._h, _addr(out), out.size))
        return out

    def set_overlap(self, on: bool = True) -> None:
        """Secondary march on its own stream, overlapping the next primary step."""
        check(lib().bfpn_set_overlap(self._h, int(on)))

    def set_group_skip(self, on: bool = True) -> None:
        """Skip the energy groups above each species' top nonzero group
        (bit-identical results; on by default)."""
        check(lib().bfpn_set_group_skip(self._h, int(on)))

    def active_groups(self) -> tuple[int, int]:
        """(primary top group, secondary top group); -1 when all zero."""
        a, b = C.c_int(0), C.c_int(0)
        check(lib().bfpn_active_groups(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    PHASES = ("energy", "transport", "angle", "source", "dose")

    def phase_stats(self) -> dict:
        """{phase: (seconds, calls, algorithmic bytes)} since profiling was turned on."""
        sec = np.zeros(5)
        calls = np.zeros(5, np.int32)
        byt = np.zeros(5)
        check(lib().bfpn_phase_stats(self._h, _addr(sec), _addr(calls), _addr(byt)))
        return {p: (float(sec[k]), int(calls[k]), float(byt[k])) for k, p in enumerate(self.PHASES)}

    @property
    def stream(self):
        return lib().bfpn_stream(self._h)


class Server:
    """Continuous-batching request loop (bfpn_server): submit beams, wait for IDDs."""

    def __init__(self, max_active: int = 8, slice_steps: int = 8, precision: int = FP32,
                 hbm_budget_bytes: int = 0, fill_cells: int = 0):
        opts = ServerOpts(hbm_budget_bytes, max_active, slice_steps, precision, fill_cells)
        h = C.c_void_p(0)
        check(lib().bfpn_server_create(C.byref(opts), C.byref(h)))
        self._h = h

    def submit(self, case: Case, t_final: float | None = None, layers=None, mapped=None) -> int:
        """Queues one beam; layers and mapped as in Solver
        (bfpn_server_submit_layered, bfpn_server_submit_mapped)."""
        g, p = case.grid(), case.phys()
        rid = C.c_uint64(0)
        t = case.time if t_final is None else t_final
        if mapped is not None:
            materials, slab_end, mp = mapped
            n = len(materials)
            spec = (Layer * n)(*[Layer(0.0, m.encode(), float(rho), path.encode())
                                 for m, rho, path in materials])
            ends = np.ascontiguousarray(slab_end, dtype=np.float64)
            mp = np.ascontiguousarray(mp, dtype=np.uint8)
            if mp.size != len(ends) * case.nyz:
                raise ValueError("the map needs (Ny+1)(Nz+1) material indices per slab")
            check(lib().bfpn_server_submit_mapped(self._h, C.byref(g), C.byref(p), spec, n,
                                                  _addr(ends), len(ends), _addr(mp), t,
                                                  C.byref(rid)))
        elif layers:
            n = len(layers)
            spec = (Layer * n)(*[Layer(float(x), m.encode(), float(rho), path.encode())
                                 for x, m, rho, path in layers])
            check(lib().bfpn_server_submit_layered(self._h, C.byref(g), C.byref(p), spec, n, t,
                                                   C.byref(rid)))
        else:
            check(lib().bfpn_server_submit(self._h, C.byref(g), C.byref(p), case.data.encode(),
                                           t, C.byref(rid)))
        return rid.value

    def poll(self, rid: int) -> int:
        return lib().bfpn_server_poll(self._h, rid)

    def wait(self, rid: int, timeout: float = 0.0, capacity: int = 1 << 16):
        """Returns (idd, {"queued": s, "running": s, "total": s})."""
        buf = np.zeros(capacity)
        n = C.c_int(0)
        t = np.zeros(3)
        check(lib().bfpn_server_wait(self._h, rid, timeout, _addr(buf), capacity, C.byref(n),
                                     _addr(t)))
        return buf[: min(n.value, capacity)].copy(), dict(zip(("queued", "running", "total"),
                                                              t.tolist()))

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().bfpn_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    @staticmethod
    def estimate(case: Case, precision: int = FP32) -> tuple[int, int]:
        """(beam bytes, shared physics bytes) that admission charges for `case`."""
        g, p = case.grid(), case.phys()
        a, b = C.c_size_t(0), C.c_size_t(0)
        check(lib().bfpn_server_estimate(C.byref(g), C.byref(p), precision, C.byref(a),
                                         C.byref(b)))
        return a.value, b.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().bfpn_server_destroy(self._h)
            self._h = C.c_void_p(0)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
