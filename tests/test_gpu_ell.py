"""GPU parity of the ELL API (include/ell.h) against the oracle: the SpMV is
bit-exact with src/ell.c:127-141, the device layout round-trips, PCG and the
stationary Jacobi match the oracle's trajectories."""
import numpy as np
import pytest

from madmax_amd import Ell, configs, lib
from oracle import pyoracle as po
from tests.helpers import oracle_rve, same_bits

pytestmark = pytest.mark.gpu


def _c1_system():
    spec = configs.c1_spec()
    r, _ = oracle_rve(spec)
    cols, vals = r.assemble()
    x = r.set_disp_0(configs.c1_strain())
    res, _ = r.residual(x)
    return r, cols, vals, -res


def test_ell_mvp_bitexact_c1():
    r, cols, vals, _ = _c1_system()
    rng = np.random.default_rng(1)
    x = rng.standard_normal(cols.shape[0])
    y_ref = np.zeros(cols.shape[0])
    po.lib().or_ell_mvp(r.ell(), po._p(x), po._p(y_ref))
    y = Ell.from_arrays(cols, vals).mvp(x)
    assert same_bits(y, y_ref)


@pytest.mark.parametrize("nrow,ncol,nnz", [(1, 1, 1), (63, 70, 5), (300, 300, 9), (1000, 517, 13),
                                            (257, 300, 16), (129, 1000, 81)])
def test_ell_mvp_bitexact_random_with_gaps(nrow, ncol, nnz):
    """Random rows (including empty rows and a -1 hole before live slots,
    which the reference's break skips) and non-multiple-of-slab sizes."""
    rng = np.random.default_rng(nrow * 7 + nnz)
    cols = np.full((nrow, nnz), -1, dtype=np.int32)
    vals = np.zeros((nrow, nnz))
    for i in range(nrow):
        k = rng.integers(0, nnz + 1)
        cols[i, :k] = rng.choice(ncol, size=k, replace=(k > ncol))
        vals[i, :k] = rng.standard_normal(k)
        if k >= 3 and i % 5 == 0:
            cols[i, rng.integers(1, k)] = -1  # everything after the first -1 is ignored
    x = rng.standard_normal(ncol)
    ref = po.OrEll(nrow, ncol, nnz, po._p(cols, po.ip), po._p(vals))
    y_ref = np.zeros(nrow)
    po.lib().or_ell_mvp(po.C.byref(ref), po._p(x), po._p(y_ref))
    y = Ell.from_arrays(cols, vals, ncol=ncol).mvp(x)
    assert same_bits(y, y_ref)


def test_ell_device_time_mvp_runs():
    import ctypes as C
    _, cols, vals, _ = _c1_system()
    e = Ell.from_arrays(cols, vals)
    h = C.c_void_p()
    assert lib().ell_device_create(C.byref(e.m), 0, C.byref(h)) == 0
    ms = C.c_double()
    assert lib().ell_device_time_mvp(h, 3, C.byref(ms)) == 0
    assert ms.value > 0
    assert lib().ell_device_time_mvp(h, 0, C.byref(ms)) == 1
    assert lib().ell_device_destroy(h) == 0


@pytest.mark.parametrize("slab", [64, 128, 256, 512])
def test_ell_device_layout_roundtrip(slab):
    import ctypes as C
    _, cols, vals, _ = _c1_system()
    e = Ell.from_arrays(cols, vals)
    h = C.c_void_p()
    assert lib().ell_device_create(C.byref(e.m), slab, C.byref(h)) == 0
    c2 = np.zeros_like(cols)
    v2 = np.zeros_like(vals)
    assert lib().ell_device_download(h, c2.ctypes.data_as(po.ip), v2.ctypes.data_as(po.dp)) == 0
    assert lib().ell_device_destroy(h) == 0
    assert np.array_equal(c2, cols)
    assert same_bits(v2, vals)


def test_ell_solve_pcg_matches_oracle_c1():
    r, cols, vals, b = _c1_system()
    x_ref, its_ref, rr_ref, ierr = po.pcg(r.ell(), b, rtol=1e-10)
    assert ierr == 0
    x, its, rr, rc = Ell.from_arrays(cols, vals).solve_pcg(b, rtol=1e-10)
    assert rc == 0
    assert abs(its - its_ref) <= 1
    assert rr <= 1e-10
    assert np.max(np.abs(x - x_ref)) <= 1e-9 * np.max(np.abs(x_ref))


@pytest.mark.parametrize("slab_rows", [64, 128, 512, 1024])
def test_ell_device_pcg_any_slab_size(slab_rows):
    """The device mirror's PCG with slabs of 64 to 1024 rows (one lane per
    row, so up to 1024 lanes per workgroup): the oracle's solution and
    iteration count, as with the default 256-row slabs."""
    import ctypes as C

    from madmax_amd._lib import EllSolver
    r, cols, vals, b = _c1_system()
    x_ref, its_ref, _, ierr = po.pcg(r.ell(), b, rtol=1e-10)
    assert ierr == 0
    e = Ell.from_arrays(cols, vals)
    h = C.c_void_p()
    print("device pcg, slab rows", slab_rows, flush=True)
    assert lib().ell_device_create(C.byref(e.m), slab_rows, C.byref(h)) == 0
    try:
        x = np.zeros_like(b)
        s = EllSolver(100000, 0, 1e-10, 0.0)
        rc = lib().ell_device_solve_pcg(h, C.byref(s), b.ctypes.data_as(C.POINTER(C.c_double)),
                                        x.ctypes.data_as(C.POINTER(C.c_double)))
    finally:
        lib().ell_device_destroy(h)
        e.free()
    assert rc == 0 and abs(s.its - its_ref) <= 1
    assert np.max(np.abs(x - x_ref)) <= 1e-9 * np.max(np.abs(x_ref))


def test_ell_solve_pcg_zero_rhs_and_cap():
    _, cols, vals, b = _c1_system()
    e = Ell.from_arrays(cols, vals)
    x, its, rr, rc = e.solve_pcg(np.zeros_like(b), x0=np.ones_like(b))
    assert rc == 0 and its == 0 and np.all(x == 0.0)
    x, its, rr, rc = e.solve_pcg(b, rtol=1e-14, max_its=3)
    assert rc == 5 and its == 3  # not converged after the iteration cap


def _laplace_1d(n=10):
    """The system of unity/unity_ell.c:14-31 with b fully initialised."""
    e = Ell(n, n, 4)
    me = np.array([1.0, -1.0, -1.0, 1.0])
    for i in range(n - 1):
        e.add_vals([i, i + 1], [i, i + 1], me)
    L = lib()
    import ctypes as C
    L.ell_set_zero_row(C.byref(e.m), 0, 1.0)
    L.ell_set_zero_row(C.byref(e.m), n - 1, 1.0)
    L.ell_set_zero_col(C.byref(e.m), 0, 1.0)
    L.ell_set_zero_col(C.byref(e.m), n - 1, 1.0)
    b = np.ones(n)
    b[n - 1] = 0.0
    return e, b


def test_ell_solve_jacobi_matches_oracle_trajectory():
    e, b = _laplace_1d()
    cols, vals = e.arrays()
    ref = po.OrEll(cols.shape[0], cols.shape[0], cols.shape[1], po._p(cols, po.ip), po._p(vals))
    x_ref = np.zeros(len(b))
    its_ref = po.C.c_int()
    err_ref = po.C.c_double()
    po.lib().or_ell_solve_jacobi(po.C.byref(ref), po._p(b), po._p(x_ref), 100000, 1e-10,
                                 po.C.byref(its_ref), po.C.byref(err_ref))
    x, its, err = e.solve_jacobi(b)
    assert abs(its - its_ref.value) <= 1
    assert err < 1e-10
    assert np.max(np.abs(x - x_ref)) <= 1e-9
