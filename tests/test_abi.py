"""The C-ABI library loads and exports every entry point include/bfpn.h declares;
the CLI front end parses like the reference's main (src/main.cu:91-273).
These run without a GPU (no compute call is made)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "bfpn.h")
CLI = os.path.join(ROOT, "bfpn_gpu_amd", "bin", "bfp_solver")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"BFPN_API\s+[\w\s\*]+?\b(bfpn_\w+)\s*\(", text)))


def test_header_declares_the_operator_and_context_api():
    names = declared_symbols()
    for required in ("bfpn_create", "bfpn_advance", "bfpn_history", "bfpn_run", "bfpn_destroy",
                     "bfpn_energy_half_f32", "bfpn_transport_half_f64", "bfpn_angle_full_f32",
                     "bfpn_secondary_source_f32", "bfpn_idd_f64", "bfpn_last_error"):
        assert required in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol():
    import bfpn_gpu_amd as bg

    L = bg.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    assert L.bfpn_version().decode().startswith("bfpn-mi355x")
    out = subprocess.run(["nm", "-D", "--defined-only", bg.lib_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    stray = sorted(n for n in exported if not n.startswith("bfpn_") and not n.startswith("_"))
    assert not stray, f"unexpected exports: {stray[:10]}"


def test_errors_are_status_codes_not_exits():
    import ctypes as C

    import bfpn_gpu_amd as bg
    from bfpn_gpu_amd.abi import Case

    L = bg.lib()
    h = C.c_void_p(0)
    c = Case(Ny=0)  # invalid grid is rejected before any device call
    g, p = c.grid(), c.phys()
    st = L.bfpn_create(C.byref(g), C.byref(p), b"./data", 0, C.byref(h))
    assert st == 1 and not h.value
    assert b"positive" in L.bfpn_last_error()
    assert L.bfpn_advance(None, 1) == 1


def _cli(*args, cwd=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


def test_cli_help_and_test_mode():
    r = _cli("--help")
    assert r.returncode == 0 and "--eq15-straggling" in r.stdout and "<Nx>" in r.stdout
    r = _cli("--test", "--bogus-flag", "123")
    assert r.returncode == 0 and "Running in test mode" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (("--material", "steel"), "Unknown material: steel"),
    (("--energy-model", "foo"), "Unknown energy model: foo"),
    (("--lite-memory", "--streaming-full"), "mutually exclusive"),
])
def test_cli_rejects_like_reference(args, msg):
    r = _cli(*args)
    assert r.returncode == 1 and msg in r.stderr


@pytest.mark.parametrize("stack,msg", [
    ([], b"null argument"),
    ([(2.0, b"water", 0.0)], b"density must be positive"),
    ([(2.0, b"water", 1.0), (1.0, b"bone", 1.85), (9.0, b"water", 1.0)], b"must increase"),
    ([(-1.0, b"water", 1.0), (9.0, b"bone", 1.85)], b"must be positive"),
    ([(2.0, b"x" * 40, 1.0)], b"too long"),
])
def test_layered_create_rejects_bad_stacks_before_any_device_call(stack, msg):
    import ctypes as C

    import bfpn_gpu_amd as bg
    from bfpn_gpu_amd.abi import Case, Layer

    L = bg.lib()
    g, p = Case().grid(), Case().phys()
    arr = (Layer * max(1, len(stack)))(*[Layer(x, m, r, b"./data") for x, m, r in stack])
    h = C.c_void_p(0)
    st = L.bfpn_create_layered(C.byref(g), C.byref(p), arr, len(stack), 0, C.byref(h))
    assert st == 1 and not h.value
    assert msg in L.bfpn_last_error()


@pytest.mark.parametrize("arg,msg", [("2:steel", "Unknown material: steel"),
                                     ("2", "--layer expects"), ("x:water", "stod")])
def test_cli_rejects_bad_layers(arg, msg):
    r = _cli("--layer", arg)
    assert r.returncode == 1 and msg in r.stderr


def _map_file(tmp_path, text):
    p = tmp_path / "map.txt"
    p.write_text(text)
    return str(p)


def test_cli_rejects_bad_material_maps(tmp_path):
    """--material-map files are read before any device call (default grid:
    11 x 11 nodes, so a slab needs 121 indices)."""
    full = "material water\nslab inf\n" + ("0 " * 11 + "\n") * 11
    cases = [
        ("no such file", None, "Cannot open material map"),
        ("index beyond the materials", full.replace("0 ", "1 ", 1), "bad map entry 1"),
        ("short slab", "material water\nslab inf\n0 0 0\n", "complete slabs of 121"),
        ("values before a slab", "material water\n0 0\n", "before a slab line"),
        ("unknown material", "material steel\nslab inf\n0\n", "Unknown material: steel"),
    ]
    for what, text, msg in cases:
        path = str(tmp_path / "missing.txt") if text is None else _map_file(tmp_path, text)
        r = _cli("--material-map", path)
        assert r.returncode == 1 and msg in r.stderr, (what, r.stderr)
    r = _cli("--material-map", _map_file(tmp_path, full), "--layer", "2:water")
    assert r.returncode == 1 and "exclude each other" in r.stderr


@pytest.mark.parametrize("nmat,ends,entry,msg", [
    (0, [9.0], 0, b"null argument"),
    (1, [9.0], 1, b"names no material"),
    (2, [2.0, 1.0, 9.0], 0, b"must increase"),
    (2, [-1.0, 9.0], 0, b"must be positive"),
])
def test_mapped_create_rejects_bad_maps_before_any_device_call(nmat, ends, entry, msg):
    import ctypes as C

    import numpy as np

    import bfpn_gpu_amd as bg
    from bfpn_gpu_amd.abi import Case, Layer

    L = bg.lib()
    c = Case()
    g, p = c.grid(), c.phys()
    mats = (Layer * 2)(Layer(0.0, b"water", 1.0, b"./data"), Layer(0.0, b"bone", 1.85, b"./data"))
    e = np.array(ends, dtype=np.float64)
    mp = np.full(len(ends) * c.nyz, entry, dtype=np.uint8)
    h = C.c_void_p(0)
    st = L.bfpn_create_mapped(C.byref(g), C.byref(p), mats, nmat, e.ctypes.data, len(ends),
                              mp.ctypes.data, 0, C.byref(h))
    assert st == 1 and not h.value
    assert msg in L.bfpn_last_error()
