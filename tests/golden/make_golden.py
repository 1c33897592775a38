"""Writes the golden fixtures of tests/golden/ by running the reference's own
sources (src/ell.c, src/fem.c, micro/src/micro_struct.c), compiled unmodified
into oracle/_ref/libref.so by oracle/Makefile.  Run in a container that has
the reference tree:  make -C oracle && python tests/golden/make_golden.py

Fixtures (inputs and the reference's outputs only):
  ref_hex8.npz        fem_init hexa8 tables (src/fem.c:309-405)
  ref_struct2d.npz    fem_init_struct 2D tables (src/fem.c:3-59), h = (0.1, 0.2)
  ref_ell_ops.npz     return codes and states of ell_* edge cases (src/ell.c)
  ref_ell_1d.npz      unity/unity_ell.c system (b fully set), BC edits, mvp, and
                      the ell_solve_jacobi trajectory (its, err, x)
  ref_ell_rve.npz     3x3x3 hex8 RVE assembled with the reference's
                      ell_add_vals / ell_set_zero_row / ell_set_zero_col in the
                      element order of micro/src/assembly.c:93-135 (Ke from
                      the oracle), cols and vals
  ref_phase.npz       micro_struct_get_elem_id (fiber_cilin) on element
                      centroids of a 10^3 mesh, one and 2x2 fibers
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class RefEll(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("nnz", C.c_int), ("cols", ip), ("vals", dp)]


class RefSolver(C.Structure):
    _fields_ = [("max_its", C.c_int), ("its", C.c_int), ("min_err", C.c_double), ("err", C.c_double)]


def ref():
    R = po.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libref.so missing: run make -C oracle where the reference exists")
    for f in ("ell_init", "ell_set_val", "ell_add_val", "ell_add_vals", "ell_mvp", "ell_get_val",
              "ell_solve_jacobi", "ell_set_zero_row", "ell_set_zero_col", "ell_set_zero_mat"):
        getattr(R, f).restype = C.c_int
    R.ell_set_val.argtypes = [C.POINTER(RefEll), C.c_int, C.c_int, C.c_double]
    R.ell_add_val.argtypes = [C.POINTER(RefEll), C.c_int, C.c_int, C.c_double]
    R.ell_get_val.argtypes = [C.POINTER(RefEll), C.c_int, C.c_int, dp]
    R.ell_set_zero_row.argtypes = [C.POINTER(RefEll), C.c_int, C.c_double]
    R.ell_set_zero_col.argtypes = [C.POINTER(RefEll), C.c_int, C.c_double]
    return R


def ell_arrays(m):
    n = m.nrow * m.nnz
    return (np.ctypeslib.as_array(m.cols, shape=(n,)).reshape(m.nrow, m.nnz).copy(),
            np.ctypeslib.as_array(m.vals, shape=(n,)).reshape(m.nrow, m.nnz).copy())


def hex8(R):
    R.fem_init()
    xp = C.POINTER(dp).in_dll(R, "xp_hexa_8")
    wp = dp.in_dll(R, "wp_hexa_8")
    sh = C.POINTER(dp).in_dll(R, "sh_hexa_8")
    ds = C.POINTER(C.POINTER(dp)).in_dll(R, "ds_hexa_8")
    out = dict(xp=np.array([[xp[g][d] for d in range(3)] for g in range(8)]),
               wp=np.array([wp[g] for g in range(8)]),
               sh=np.array([[sh[a][g] for g in range(8)] for a in range(8)]),
               ds=np.array([[[ds[a][d][g] for g in range(8)] for d in range(3)] for a in range(8)]))
    np.savez(os.path.join(HERE, "ref_hex8.npz"), **out)


def struct2d(R):
    sh = C.POINTER(dp)()
    dsh = C.POINTER(C.POINTER(dp))()
    wp = dp()
    h = (C.c_double * 2)(0.1, 0.2)
    R.fem_init_struct(C.byref(sh), C.byref(dsh), C.byref(wp), h, 2)
    out = dict(h=np.array([0.1, 0.2]),
               sh=np.array([[sh[a][g] for g in range(4)] for a in range(4)]),
               dsh=np.array([[[dsh[a][d][g] for g in range(4)] for d in range(2)] for a in range(4)]),
               wp=np.array([wp[g] for g in range(4)]))
    np.savez(os.path.join(HERE, "ref_struct2d.npz"), **out)


def ell_ops(R):
    """Return codes of edge cases and the resulting matrix."""
    m = RefEll()
    R.ell_init(C.byref(m), 4, 5, 3)
    codes = []
    ops = [("add", 0, 1, 1.5), ("add", 0, 1, 2.0), ("add", 0, 3, -1.0), ("set", 0, 1, 7.0),
           ("add", 0, 4, 1.0), ("add", 0, 2, 9.0),  # row full -> 3
           ("add", 4, 0, 1.0), ("add", 1, 5, 1.0),  # beyond dims -> 1
           ("add", -1, 0, 1.0), ("set", 2, -3, 1.0),  # negative -> 2
           ("set", 3, 3, 4.0), ("add", 3, 0, 2.5), ("add", 2, 2, 0.5)]
    for op, r, c, v in ops:
        f = R.ell_add_val if op == "add" else R.ell_set_val
        codes.append(f(C.byref(m), r, c, v))
    getv = []
    for r, c in [(0, 1), (0, 3), (0, 2), (1, 1), (3, 0), (2, 2), (4, 0), (0, -1)]:
        v = C.c_double(-99.0)
        rc = R.ell_get_val(C.byref(m), r, c, C.byref(v))
        getv.append((rc, v.value))
    cols, vals = ell_arrays(m)
    R.ell_set_zero_row(C.byref(m), 3, 1.0)
    R.ell_set_zero_col(C.byref(m), 4, 2.0)
    cols2, vals2 = ell_arrays(m)
    np.savez(os.path.join(HERE, "ref_ell_ops.npz"),
             ops=np.array([[0 if o == "add" else 1, r, c, v] for o, r, c, v in ops], dtype=np.float64),
             codes=np.array(codes), getv=np.array(getv), cols=cols, vals=vals, cols2=cols2, vals2=vals2)


def ell_1d(R, n=10):
    m = RefEll()
    R.ell_init(C.byref(m), n, n, 4)
    me = (C.c_double * 4)(1, -1, -1, 1)
    ix = (C.c_int * 2)()
    for i in range(n - 1):
        ix[0], ix[1] = i, i + 1
        R.ell_add_vals(C.byref(m), ix, 2, ix, 2, me)
    R.ell_set_zero_row(C.byref(m), 0, 1.0)
    R.ell_set_zero_row(C.byref(m), n - 1, 1.0)
    R.ell_set_zero_col(C.byref(m), 0, 1.0)
    R.ell_set_zero_col(C.byref(m), n - 1, 1.0)
    cols, vals = ell_arrays(m)
    b = np.ones(n)
    b[n - 1] = 0.0
    xm = np.linspace(-1.0, 2.0, n)
    y = np.zeros(n)
    R.ell_mvp(C.byref(m), xm.ctypes.data_as(dp), y.ctypes.data_as(dp))
    x = np.zeros(n)
    s = RefSolver()
    R.ell_solve_jacobi(C.byref(s), C.byref(m), b.ctypes.data_as(dp), x.ctypes.data_as(dp))
    np.savez(os.path.join(HERE, "ref_ell_1d.npz"), cols=cols, vals=vals, b=b, xm=xm, y=y, x=x,
             its=s.its, err=s.err)


def ell_rve(R, n=3):
    mesh = po.mesh(3, n, n, n)
    et = po.phase_map(mesh, [(0.45, 0.55, 0.3)])
    E, nu = [1.0e7, 8.0e7], [0.25, 0.32]
    kes = [po.ke(3, (mesh.hx, mesh.hy, mesh.hz), E[k], nu[k]) for k in range(2)]
    m = RefEll()
    nrow = 3 * mesh.nn
    R.ell_init(C.byref(m), nrow, nrow, 81)
    idx = np.zeros(24, dtype=np.int32)
    for e in range(mesh.nelm):
        po.lib().or_elem_index(C.byref(mesh), e, idx.ctypes.data_as(ip))
        kk = np.ascontiguousarray(kes[et[e]])
        R.ell_add_vals(C.byref(m), idx.ctypes.data_as(ip), 24, idx.ctypes.data_as(ip), 24,
                       kk.ctypes.data_as(dp))
    bnd = np.zeros(mesh.nn, dtype=np.int32)
    nb = po.lib().or_boundary_nodes(C.byref(mesh), bnd.ctypes.data_as(ip))
    for d in range(3):
        for i in range(nb):
            R.ell_set_zero_row(C.byref(m), int(bnd[i]) * 3 + d, 1.0)
            R.ell_set_zero_col(C.byref(m), int(bnd[i]) * 3 + d, 1.0)
    cols, vals = ell_arrays(m)
    np.savez(os.path.join(HERE, "ref_ell_rve.npz"), n=n, fiber=np.array([0.45, 0.55, 0.3]),
             young=np.array(E), poisson=np.array(nu), elem_type=et, cols=cols, vals=vals)


class MicroStruct(C.Structure):
    _fields_ = [("type", C.c_int), ("size", dp), ("data", C.c_void_p)]


def phase(R):
    R.micro_struct_init.argtypes = [C.c_int, C.c_char_p, C.POINTER(MicroStruct)]
    R.micro_struct_get_elem_id.argtypes = [C.c_int, C.POINTER(MicroStruct), dp, ip]
    out = {}
    mesh = po.mesh(3, 10, 10, 10)
    cent = np.zeros((mesh.nelm, 3))
    for e in range(mesh.nelm):
        po.lib().or_elem_centroid(C.byref(mesh), e, cent[e].ctypes.data_as(dp))
    for name, s in [("one", b"fiber_cilin 1.0 1.0 1 1 0.25 0.0 0.0"),
                    ("grid", b"fiber_cilin 1.0 1.0 2 2 0.2 0.03 -0.05")]:
        ms = MicroStruct()
        assert R.micro_struct_init(2, s, C.byref(ms)) == 0
        ids = np.zeros(mesh.nelm, dtype=np.int32)
        v = C.c_int()
        for e in range(mesh.nelm):
            c2 = np.ascontiguousarray(cent[e, :2])
            R.micro_struct_get_elem_id(2, C.byref(ms), c2.ctypes.data_as(dp), C.byref(v))
            ids[e] = v.value
        out[name] = ids
    np.savez(os.path.join(HERE, "ref_phase.npz"), **out)


if __name__ == "__main__":
    R = ref()
    hex8(R)
    struct2d(R)
    ell_ops(R)
    ell_1d(R)
    ell_rve(R)
    phase(R)
    print("fixtures written to", HERE)
