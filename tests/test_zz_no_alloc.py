"""ell_device_solve_pcg_noalloc: the Jacobi-PCG of ell_device_solve_pcg with
b and x in caller-owned device memory and the work vectors in the library's
file-scope device array.  Its results must equal the allocating entry
point's bit for bit (same kernels, same reduction order), on two calls in a
row, the second on a smaller system (the array keeps the first call's
contents), and a matrix above ELL_NOALLOC_MAX_ROWS rows is refused."""
import ctypes as C

import numpy as np
import pytest

from madmax_amd import Ell, lib, rve_spec
from madmax_amd._lib import EllSolver
from tests.helpers import DevBuf, oracle_rve, same_bits

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def _rve_system(n, eps):
    r, _ = oracle_rve(rve_spec(n, [(0.45, 0.5, 0.3)]))
    cols, vals = r.assemble()
    x0 = r.set_disp_0(eps)
    res, _ = r.residual(x0)
    return cols, vals, -res


def _device(cols, vals):
    e = Ell.from_arrays(cols, vals)
    h = C.c_void_p()
    assert lib().ell_device_create(C.byref(e.m), 0, C.byref(h)) == 0
    return e, h


def test_noalloc_pcg_equals_allocating_pcg():
    systems = [_rve_system(9, np.array([1e-3, 0, 2e-4, 0, 1e-4, 0])),
               _rve_system(5, np.array([0, -1e-3, 0, 3e-4, 0, 0]))]
    for i, (cols, vals, b) in enumerate(systems):
        n = cols.shape[0]
        print("noalloc call", i, "rows", n, flush=True)
        e, h = _device(cols, vals)
        try:
            x0 = np.linspace(-1e-6, 1e-6, n) if i == 1 else np.zeros(n)  # second call: nonzero guess
            x_ref = x0.copy()
            s_ref = EllSolver(100000, 0, 1e-10, 0.0)
            rc_ref = lib().ell_device_solve_pcg(h, C.byref(s_ref), b.ctypes.data_as(C.POINTER(C.c_double)),
                                                x_ref.ctypes.data_as(C.POINTER(C.c_double)))
            db, dx = DevBuf(b), DevBuf(x0)
            try:
                s = EllSolver(100000, 0, 1e-10, 0.0)
                rc = lib().ell_device_solve_pcg_noalloc(h, C.byref(s), db.ptr, dx.ptr)
                x = dx.get()
            finally:
                db.free()
                dx.free()
        finally:
            lib().ell_device_destroy(h)
            e.free()
        assert rc_ref == 0 and rc == 0
        assert s.its == s_ref.its and s.its > 0
        assert s.err == s_ref.err
        assert same_bits(x, x_ref)


def test_noalloc_rejects_matrices_above_the_scratch():
    n = MAX_ROWS + 1  # a diagonal matrix one row too large
    cols = np.arange(n, dtype=np.int32).reshape(n, 1)
    vals = np.ones((n, 1))
    e, h = _device(cols, vals)
    db, dx = DevBuf(np.ones(n)), DevBuf(np.zeros(n))
    try:
        s = EllSolver(10, 0, 1e-10, 0.0)
        assert lib().ell_device_solve_pcg_noalloc(h, C.byref(s), db.ptr, dx.ptr) == 1
        assert float(np.max(np.abs(dx.get()))) == 0.0  # nothing ran
    finally:
        db.free()
        dx.free()
        lib().ell_device_destroy(h)
        e.free()
