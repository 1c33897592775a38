"""The oracle pinned against the reference's own code: the committed fixtures
written by tests/golden/make_golden.py from src/ell.c, src/fem.c and
micro/src/micro_struct.c (compiled unmodified into oracle/_ref), and, where
that library is present, direct randomized comparisons with it."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name))


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_hex8_tables_bitwise():
    g = gold("ref_hex8.npz")
    xp, wp, sh, ds = po.hex8_tables()
    assert bits_equal(xp, g["xp"])
    assert bits_equal(wp, g["wp"])
    assert bits_equal(sh, g["sh"])
    assert bits_equal(ds, g["ds"])


def test_struct2d_tables_bitwise():
    g = gold("ref_struct2d.npz")
    dsh = np.zeros((4, 2, 4))
    wp = np.zeros(4)
    h = np.ascontiguousarray(g["h"])
    po.lib().or_struct_dsh(2, po._p(h), po._p(dsh), po._p(wp))
    assert bits_equal(dsh, g["dsh"])
    assert bits_equal(wp, g["wp"])


def _or_ell(nrow, ncol, nnz):
    e = po.OrEll()
    assert po.lib().or_ell_init(C.byref(e), nrow, ncol, nnz) == 0
    return e


def test_ell_ops_codes_and_state():
    g = gold("ref_ell_ops.npz")
    e = _or_ell(4, 5, 3)
    L = po.lib()
    codes = []
    for op, r, c, v in g["ops"]:
        f = L.or_ell_add_val if op == 0 else L.or_ell_set_val
        codes.append(f(C.byref(e), int(r), int(c), float(v)))
    assert codes == list(g["codes"])
    for (rc_ref, v_ref), (r, c) in zip(g["getv"], [(0, 1), (0, 3), (0, 2), (1, 1), (3, 0), (2, 2),
                                                   (4, 0), (0, -1)]):
        v = C.c_double(-99.0)
        assert L.or_ell_get_val(C.byref(e), r, c, C.byref(v)) == rc_ref
        assert v.value == v_ref
    cols, vals = po.ell_to_numpy(e)
    assert np.array_equal(cols, g["cols"]) and bits_equal(vals, g["vals"])
    L.or_ell_set_zero_row(C.byref(e), 3, 1.0)
    L.or_ell_set_zero_col(C.byref(e), 4, 2.0)
    cols, vals = po.ell_to_numpy(e)
    assert np.array_equal(cols, g["cols2"]) and bits_equal(vals, g["vals2"])


def test_ell_1d_system_mvp_and_jacobi_trajectory():
    g = gold("ref_ell_1d.npz")
    n = 10
    e = _or_ell(n, n, 4)
    L = po.lib()
    me = np.array([1.0, -1.0, -1.0, 1.0])
    for i in range(n - 1):
        ix = np.array([i, i + 1], dtype=np.int32)
        L.or_ell_add_vals(C.byref(e), po._p(ix, po.ip), 2, po._p(ix, po.ip), 2, po._p(me))
    L.or_ell_set_zero_row(C.byref(e), 0, 1.0)
    L.or_ell_set_zero_row(C.byref(e), n - 1, 1.0)
    L.or_ell_set_zero_col(C.byref(e), 0, 1.0)
    L.or_ell_set_zero_col(C.byref(e), n - 1, 1.0)
    cols, vals = po.ell_to_numpy(e)
    assert np.array_equal(cols, g["cols"]) and bits_equal(vals, g["vals"])
    # ell_set_zero_col stops at the first empty slot of row 0 (src/ell.c:110):
    # column 0 keeps its -1 in row 1
    assert vals[1, list(cols[1]).index(0)] == -1.0
    y = np.zeros(n)
    L.or_ell_mvp(C.byref(e), po._p(np.ascontiguousarray(g["xm"])), po._p(y))
    assert bits_equal(y, g["y"])
    x = np.zeros(n)
    its, err = C.c_int(), C.c_double()
    L.or_ell_solve_jacobi(C.byref(e), po._p(np.ascontiguousarray(g["b"])), po._p(x), 100000, 1e-10,
                          C.byref(its), C.byref(err))
    assert its.value == int(g["its"])
    assert err.value == float(g["err"])
    assert bits_equal(x, g["x"])


def test_rve_assembly_bitwise_vs_reference_insertion():
    """3x3x3 RVE: the oracle's ELL equals the one the reference's own
    ell_add_vals / ell_set_zero_row / ell_set_zero_col build."""
    g = gold("ref_ell_rve.npz")
    n = int(g["n"])
    m = po.mesh(3, n, n, n)
    et = po.phase_map(m, [tuple(g["fiber"])])
    assert np.array_equal(et, g["elem_type"])
    r = po.Rve(3, n, et, list(g["young"]), list(g["poisson"]))
    cols, vals = r.assemble()
    assert np.array_equal(cols, g["cols"])
    assert bits_equal(vals, g["vals"])


def _grid_fibers(lx, ly, nfx, nfy, r, desv):
    """fiber centres of micro/src/micro_struct.c:109-117 (same evaluation order)"""
    out = []
    for i in range(nfx):
        for j in range(nfy):
            dx = desv[0] - lx / 2 + (lx / nfx) / 2 + i * (lx / nfx)
            dy = desv[1] - ly / 2 + (ly / nfy) / 2 + j * (ly / nfy)
            out.append((lx / 2 + dx, ly / 2 + dy, r))
    return out


def test_phase_map_vs_reference_micro_struct():
    g = gold("ref_phase.npz")
    m = po.mesh(3, 10, 10, 10)
    assert np.array_equal(po.phase_map(m, [(0.5, 0.5, 0.25)]), g["one"])
    assert int(g["one"].sum()) == 160  # the 160 fiber elements of configuration 1
    assert np.array_equal(po.phase_map(m, _grid_fibers(1.0, 1.0, 2, 2, 0.2, (0.03, -0.05))), g["grid"])


needs_ref = pytest.mark.skipif(po.ref_lib() is None, reason="reference not built (oracle/_ref)")


@needs_ref
@pytest.mark.parametrize("seed", range(5))
def test_random_ell_sequences_vs_reference(seed):
    """Random insert / set / zero sequences, then mvp: oracle == reference."""
    from tests.golden.make_golden import RefEll, ref

    R = ref()
    rng = np.random.default_rng(seed)
    nrow, ncol, nnz = int(rng.integers(1, 30)), int(rng.integers(1, 30)), int(rng.integers(1, 8))
    a, b = RefEll(), _or_ell(nrow, ncol, nnz)
    R.ell_init(C.byref(a), nrow, ncol, nnz)
    L = po.lib()
    for _ in range(200):
        op = rng.integers(0, 4)
        r, c = int(rng.integers(-1, nrow + 1)), int(rng.integers(-1, ncol + 1))
        v = float(rng.standard_normal())
        if op == 0:
            assert R.ell_add_val(C.byref(a), r, c, v) == L.or_ell_add_val(C.byref(b), r, c, v)
        elif op == 1:
            assert R.ell_set_val(C.byref(a), r, c, v) == L.or_ell_set_val(C.byref(b), r, c, v)
        elif op == 2 and 0 <= r < nrow:
            R.ell_set_zero_row(C.byref(a), r, v)
            L.or_ell_set_zero_row(C.byref(b), r, v)
        elif op == 3:
            R.ell_set_zero_col(C.byref(a), c, v)
            L.or_ell_set_zero_col(C.byref(b), c, v)
    from tests.golden.make_golden import ell_arrays
    ca, va = ell_arrays(a)
    cb, vb = po.ell_to_numpy(b)
    assert np.array_equal(ca, cb) and bits_equal(va, vb)
    x = rng.standard_normal(ncol)
    ya, yb = np.zeros(nrow), np.zeros(nrow)
    R.ell_mvp(C.byref(a), x.ctypes.data_as(po.dp), ya.ctypes.data_as(po.dp))
    L.or_ell_mvp(C.byref(b), po._p(x), po._p(yb))
    assert bits_equal(ya, yb)
