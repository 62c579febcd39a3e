"""bfpn_idd_f32_noalloc (no workspace, no device allocation) against
bfpn_idd_f32 through the C ABI.  Both launch the same kernels in the same
order on the same inputs, so the results are bit-identical.  Two calls in a
row, the second on a smaller grid (the shared scratch keeps the first call's
contents, which the second must not read), then a grid whose scratch exceeds
the fixed maximum is rejected on the host before any launch."""
import ctypes as C

import numpy as np
import pytest

from gpu_cases import small_case

pytestmark = pytest.mark.gpu


def _bind(L):
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.bfpn_idd_f32_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, vp, i, i, i, i, d, d, d, d, d, d,
                                       i, i, C.POINTER(d), vp]
    return L.bfpn_idd_f32_noalloc


def _inputs(gpu, oracle_mod, **over):
    c = small_case(**over)
    s = oracle_mod.Solver(c)
    s.advance(2)
    st = [gpu.DeviceArray(s.state(k).astype(np.float32)) for k in range(4)]
    tabs = [gpu.DeviceArray(np.ascontiguousarray(s.table(k), dtype=np.float64))
            for k in ("S", "sigma", "en")]
    s.close()
    return c, oracle_mod.deltas(c), st, tabs


@pytest.mark.parametrize("trap,cat", [(True, True), (False, False)])
def test_noalloc_idd_equals_workspace_idd(gpu, oracle_mod, trap, cat):
    fn = _bind(gpu.lib())
    for over in (dict(Ny=6, Nz=5, Nmu=6, Nom=5), dict(Ny=3, Nz=2, Nmu=3, Nom=4, Ng=200)):
        c, d, st, tabs = _inputs(gpu, oracle_mod, **over)
        args = (c.Ng, c.Ny, c.Nz, c.nang, d.dg, d.du, d.dv, d.dy, d.dz, 1.3, int(trap), int(cat))
        ref = gpu.ops.idd(*st, *tabs, *args[:10], trapezoidal=trap, catastrophic=cat)
        out = C.c_double(0.0)
        gpu.check(fn(*(a.ptr for a in st), *(t.ptr for t in tabs), *args, C.byref(out), None))
        assert out.value == ref, (over, out.value, ref)
        assert ref > 0


def test_noalloc_idd_rejects_grids_beyond_its_scratch(gpu, oracle_mod):
    fn = _bind(gpu.lib())
    c, d, st, tabs = _inputs(gpu, oracle_mod)
    out = C.c_double(-1.0)
    # 401x401 nodes x 400 directions: 250k partial sums alone exceed 2 MiB
    st_ = fn(*(a.ptr for a in st), *(t.ptr for t in tabs), c.Ng, 400, 400, 400, d.dg, d.du, d.dv,
             d.dy, d.dz, 1.0, 1, 1, C.byref(out), None)
    assert st_ == 1 and out.value == -1.0
    assert b"scratch" in gpu.lib().bfpn_last_error()
