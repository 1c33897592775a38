"""The fused persistent Jacobi-PCG (ell_device_set_pcg_fused): one cooperative
kernel that runs whole iterations, its workgroups meeting at two grid
barriers per iteration, must give the bits of the two-launch iteration
(same row arithmetic, same reduction order) for any number of iterations per
launch, on systems where every workgroup owns one slab and on one where
workgroups own several (more slabs than co-resident workgroups)."""
import ctypes as C

import numpy as np
import pytest

from madmax_amd import Ell, lib, rve_spec
from madmax_amd._lib import EllSolver
from oracle import pyoracle as po
from tests.helpers import DevBuf, oracle_rve, same_bits

pytestmark = pytest.mark.gpu


def _rve_system(n, eps):
    r, _ = oracle_rve(rve_spec(n, [(0.45, 0.5, 0.3)]))
    cols, vals = r.assemble()
    x0 = r.set_disp_0(eps)
    res, _ = r.residual(x0)
    return r, cols, vals, -res


def _laplace_3d(n, shift=0.05):
    """7-point Laplacian plus `shift` on the diagonal over an n^3 grid (SPD),
    rows packed left with -1 after the last neighbour."""
    N = n ** 3
    idx = np.arange(N).reshape(n, n, n)
    nb = [idx]
    for ax in range(3):
        for step in (-1, 1):
            s = np.full((n, n, n), -1, dtype=np.int64)
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if step < 0:
                dst[ax], src[ax] = slice(1, None), slice(0, -1)
            else:
                dst[ax], src[ax] = slice(0, -1), slice(1, None)
            s[tuple(dst)] = idx[tuple(src)]
            nb.append(s)
    cols = np.stack([a.reshape(N) for a in nb], axis=1)
    vals = np.where(cols >= 0, -1.0, 0.0)
    vals[:, 0] = 6.0 + shift
    order = np.argsort(cols < 0, axis=1, kind="stable")
    return (np.ascontiguousarray(np.take_along_axis(cols, order, axis=1), dtype=np.int32),
            np.ascontiguousarray(np.take_along_axis(vals, order, axis=1)))


class Mirror:
    def __init__(self, cols, vals, slab_rows=0):
        self.e = Ell.from_arrays(cols, vals)
        self.h = C.c_void_p()
        assert lib().ell_device_create(C.byref(self.e.m), slab_rows, C.byref(self.h)) == 0

    def solve(self, b, x0, rtol, max_its, fused):
        assert lib().ell_device_set_pcg_fused(self.h, fused) == 0
        x = np.array(x0, dtype=np.float64)
        s = EllSolver(max_its, 0, rtol, 0.0)
        rc = lib().ell_device_solve_pcg(self.h, C.byref(s), b.ctypes.data_as(C.POINTER(C.c_double)),
                                        x.ctypes.data_as(C.POINTER(C.c_double)))
        return x, s.its, s.err, rc

    def close(self):
        lib().ell_device_destroy(self.h)
        self.e.free()


@pytest.mark.parametrize("per_launch", [1, 5, 100000])
def test_fused_equals_two_launch_iteration_on_rves(per_launch):
    for n, eps in ((10, np.array([1e-3, 0, 2e-4, 0, 1e-4, 0])), (7, np.array([0, -1e-3, 0, 3e-4, 0, 0]))):
        r, cols, vals, b = _rve_system(n, eps)
        print("fused pcg: rve", n, "rows", cols.shape[0], "per launch", per_launch, flush=True)
        m = Mirror(cols, vals)
        try:
            x0, its0, err0, rc0 = m.solve(b, np.zeros_like(b), 1e-10, 100000, 0)
            x1, its1, err1, rc1 = m.solve(b, np.zeros_like(b), 1e-10, 100000, per_launch)
        finally:
            m.close()
        assert rc0 == 0 and rc1 == 0
        assert its1 == its0 and its0 > 0
        assert err1 == err0
        assert same_bits(x1, x0)
        # and the solve itself is the oracle's PCG
        x_ref, its_ref, _, ierr = po.pcg(r.ell(), b, rtol=1e-10)
        assert ierr == 0 and abs(its1 - its_ref) <= 1
        assert np.max(np.abs(x1 - x_ref)) <= 1e-9 * np.max(np.abs(x_ref))


def test_fused_equals_two_launch_iteration_with_several_slabs_per_workgroup():
    cols, vals = _laplace_3d(85)  # 614,125 rows = 2,399 slabs of 256
    rng = np.random.default_rng(5)
    b = rng.standard_normal(cols.shape[0])
    print("fused pcg: laplace 85^3, 40 iterations", flush=True)
    m = Mirror(cols, vals)
    try:
        out = [m.solve(b, np.zeros_like(b), 1e-30, 40, f) for f in (0, 16, 100000)]
    finally:
        m.close()
    x0, its0, err0, rc0 = out[0]
    assert rc0 == 5 and its0 == 40
    for x, its, err, rc in out[1:]:
        assert rc == 5 and its == 40
        assert err == err0
        assert same_bits(x, x0)


def test_fused_iteration_cap_and_zero_rhs():
    _, cols, vals, b = _rve_system(6, np.array([1e-3, 0, 0, 0, 0, 0]))
    print("fused pcg: cap and zero rhs", flush=True)
    m = Mirror(cols, vals)
    try:
        x, its, err, rc = m.solve(b, np.zeros_like(b), 1e-14, 3, 4)
        assert rc == 5 and its == 3
        x, its, err, rc = m.solve(np.zeros_like(b), np.ones_like(b), 1e-10, 100, 4)
        assert rc == 0 and its == 0 and np.all(x == 0.0)
    finally:
        m.close()


def test_fused_setting_is_checked():
    _, cols, vals, _ = _rve_system(4, np.array([1e-3, 0, 0, 0, 0, 0]))
    m = Mirror(cols, vals, slab_rows=128)
    try:
        assert lib().ell_device_set_pcg_fused(m.h, -1) == 1
        assert lib().ell_device_set_pcg_fused(m.h, 8) == 1  # the fused kernel takes 256-row slabs only
        assert lib().ell_device_set_pcg_fused(m.h, 0) == 0
    finally:
        m.close()


def test_fused_noalloc_equals_two_launch_allocating():
    """The no-allocation entry point on a fused mirror: its work vectors and
    the barrier counter come from the library's file-scope array."""
    _, cols, vals, b = _rve_system(8, np.array([0, 0, 1e-3, 0, 0, 2e-4]))
    print("fused pcg: noalloc", flush=True)
    m = Mirror(cols, vals)
    db, dx = DevBuf(b), DevBuf(np.zeros_like(b))
    try:
        x0, its0, err0, rc0 = m.solve(b, np.zeros_like(b), 1e-10, 100000, 0)
        assert lib().ell_device_set_pcg_fused(m.h, 32) == 0
        s = EllSolver(100000, 0, 1e-10, 0.0)
        rc = lib().ell_device_solve_pcg_noalloc(m.h, C.byref(s), db.ptr, dx.ptr)
        x = dx.get()
    finally:
        db.free()
        dx.free()
        m.close()
    assert rc0 == 0 and rc == 0
    assert s.its == its0 and s.err == err0
    assert same_bits(x, x0)
