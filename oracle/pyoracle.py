"""ctypes view of the FP64 CPU oracle (oracle/liboracle.so) and of the
reference sources compiled unmodified into oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.  The product package never imports this file.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class OrEll(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("nnz", C.c_int),
                ("cols", ip), ("vals", dp)]


class OrMesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("nex", C.c_int), ("ney", C.c_int), ("nez", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nn", C.c_int),
                ("nelm", C.c_int), ("npe", C.c_int), ("nvoi", C.c_int), ("ngp", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
                ("vol", C.c_double), ("vol_elm", C.c_double)]


def _p(a, t=dp):
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.or_rve_create.restype = C.c_void_p
        L.or_rve_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, ip, C.c_int, dp, dp]
        L.or_rve_free.argtypes = [C.c_void_p]
        L.or_rve_ell.restype = C.POINTER(OrEll)
        L.or_rve_ell.argtypes = [C.c_void_p]
        L.or_rve_mesh.restype = C.POINTER(OrMesh)
        L.or_rve_mesh.argtypes = [C.c_void_p]
        L.or_rve_assemble_jac.argtypes = [C.c_void_p]
        L.or_rve_residual.argtypes = [C.c_void_p, dp, dp, dp]
        L.or_rve_averages.argtypes = [C.c_void_p, dp, dp, dp]
        L.or_rve_set_disp_0.argtypes = [C.c_void_p, dp, dp]
        L.or_rve_homog_fe2.argtypes = [C.c_void_p, dp, C.c_int, C.c_double, C.c_int, C.c_double,
                                       dp, dp, dp, ip, ip]
        L.or_rve_c_tangent_fd.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, dp]
        L.or_rve_c_tangent_unit.argtypes = [C.c_void_p, C.c_double, C.c_double, dp, ip]
        L.or_rve_taylor.argtypes = [C.c_void_p, C.c_int, dp]
        L.or_ell_init.argtypes = [C.POINTER(OrEll), C.c_int, C.c_int, C.c_int]
        L.or_ell_free.argtypes = [C.POINTER(OrEll)]
        L.or_ell_add_vals.argtypes = [C.POINTER(OrEll), ip, C.c_int, ip, C.c_int, dp]
        L.or_ell_add_val.argtypes = [C.POINTER(OrEll), C.c_int, C.c_int, C.c_double]
        L.or_ell_set_val.argtypes = [C.POINTER(OrEll), C.c_int, C.c_int, C.c_double]
        L.or_ell_set_zero_row.argtypes = [C.POINTER(OrEll), C.c_int, C.c_double]
        L.or_ell_set_zero_col.argtypes = [C.POINTER(OrEll), C.c_int, C.c_double]
        L.or_ell_set_zero_mat.argtypes = [C.POINTER(OrEll)]
        L.or_ell_mvp.argtypes = [C.POINTER(OrEll), dp, dp]
        L.or_ell_get_val.argtypes = [C.POINTER(OrEll), C.c_int, C.c_int, dp]
        L.or_ell_solve_jacobi.argtypes = [C.POINTER(OrEll), dp, dp, C.c_int, C.c_double, ip, dp]
        L.or_pcg.argtypes = [C.POINTER(OrEll), dp, dp, C.c_double, C.c_int, ip, dp]
        L.or_hex8_tables.argtypes = [dp, dp, dp, dp]
        L.or_struct_dsh.argtypes = [C.c_int, dp, dp, dp]
        L.or_struct_bmat.argtypes = [C.c_int, dp, dp, dp]
        L.or_material_c.argtypes = [C.c_int, C.c_double, C.c_double, dp]
        L.or_ke.argtypes = [C.c_int, dp, dp, dp, dp]
        L.or_mesh_init.argtypes = [C.POINTER(OrMesh), C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_double, C.c_double]
        L.or_elem_index.argtypes = [C.POINTER(OrMesh), C.c_int, ip]
        L.or_elem_centroid.argtypes = [C.POINTER(OrMesh), C.c_int, dp]
        L.or_boundary_nodes.argtypes = [C.POINTER(OrMesh), ip]
        L.or_phase_map.argtypes = [C.POINTER(OrMesh), C.c_int, dp, dp, dp, ip]
        L.or_rve_homog_per_lm.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_macro_bmat.argtypes = [dp, dp, dp]
        L.or_macro_ke.argtypes = [dp, dp, dp]
        L.or_macro_strain.argtypes = [dp, dp, dp]
        L.or_macro_res.argtypes = [dp, dp, dp]
        L.or_ell_zero_rows_cols.argtypes = [C.POINTER(OrEll), C.c_int, ip, C.c_double]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own ell.c / fem.c / micro_struct.c, or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
    return _REF


# ------------------------------------------------------------------ helpers

def hex8_tables():
    xp = np.zeros((8, 3)); wp = np.zeros(8); sh = np.zeros((8, 8)); ds = np.zeros((8, 3, 8))
    lib().or_hex8_tables(_p(xp), _p(wp), _p(sh), _p(ds))
    return xp, wp, sh, ds


def struct_bmat(dim, h):
    npe = 4 if dim == 2 else 8
    nvoi = 3 if dim == 2 else 6
    b = np.zeros((nvoi, npe * dim, npe)); wp = np.zeros(npe)
    lib().or_struct_bmat(dim, _p(np.ascontiguousarray(h, dtype=np.float64)), _p(b), _p(wp))
    return b, wp


def material_c(dim, E, nu):
    nvoi = 3 if dim == 2 else 6
    c = np.zeros((nvoi, nvoi))
    lib().or_material_c(dim, E, nu, _p(c))
    return c


def ke(dim, h, E, nu):
    b, wp = struct_bmat(dim, h)
    c = material_c(dim, E, nu)
    nd = b.shape[1]
    k = np.zeros((nd, nd))
    lib().or_ke(dim, _p(b), _p(wp), _p(c), _p(k))
    return k


def mesh(dim, nex, ney, nez, lx=1.0, ly=1.0, lz=1.0):
    m = OrMesh()
    if lib().or_mesh_init(C.byref(m), dim, nex, ney, nez, lx, ly, lz):
        raise ValueError("bad mesh")
    return m


def phase_map(m, fibers):
    f = np.asarray(fibers, dtype=np.float64).reshape(-1, 3)
    fx = np.ascontiguousarray(f[:, 0]); fy = np.ascontiguousarray(f[:, 1])
    fr = np.ascontiguousarray(f[:, 2])
    et = np.zeros(m.nelm, dtype=np.int32)
    lib().or_phase_map(C.byref(m), len(f), _p(fx), _p(fy), _p(fr), _p(et, ip))
    return et


def ell_to_numpy(e):
    n = e.nrow * e.nnz
    cols = np.ctypeslib.as_array(e.cols, shape=(n,)).reshape(e.nrow, e.nnz).copy()
    vals = np.ctypeslib.as_array(e.vals, shape=(n,)).reshape(e.nrow, e.nnz).copy()
    return cols, vals


class Rve:
    """An oracle RVE: structured mesh, phase map, materials, ELL Jacobian."""

    def __init__(self, dim, n, elem_type, young, poisson, lengths=(1.0, 1.0, 1.0)):
        L = lib()
        nex, ney, nez = (n, n, n) if isinstance(n, int) else tuple(n) + ((1,) if dim == 2 else ())
        self.dim = dim
        et = np.ascontiguousarray(elem_type, dtype=np.int32)
        yo = np.ascontiguousarray(young, dtype=np.float64)
        po = np.ascontiguousarray(poisson, dtype=np.float64)
        self.h = L.or_rve_create(dim, nex, ney, nez, lengths[0], lengths[1], lengths[2],
                                 _p(et, ip), len(yo), _p(yo), _p(po))
        if not self.h:
            raise ValueError("or_rve_create failed")
        self.mesh = L.or_rve_mesh(self.h).contents
        self.nvoi = self.mesh.nvoi
        self.nrow = self.mesh.nn * dim

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_rve_free(self.h)
            self.h = None

    def assemble(self):
        lib().or_rve_assemble_jac(self.h)
        return ell_to_numpy(lib().or_rve_ell(self.h).contents)

    def ell(self):
        return lib().or_rve_ell(self.h)

    def residual(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(self.nrow); nrm = C.c_double()
        lib().or_rve_residual(self.h, _p(x), _p(r), C.byref(nrm))
        return r, nrm.value

    def averages(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        se = np.zeros(self.nvoi); ss = np.zeros(self.nvoi)
        lib().or_rve_averages(self.h, _p(x), _p(se), _p(ss))
        return se, ss

    def set_disp_0(self, strain, x=None):
        x = np.zeros(self.nrow) if x is None else np.ascontiguousarray(x, dtype=np.float64)
        s = np.ascontiguousarray(strain, dtype=np.float64)
        lib().or_rve_set_disp_0(self.h, _p(s), _p(x))
        return x

    def homog(self, strain, rtol=1e-10, solver=0, nl_max_its=2, nl_min_norm=1e-4):
        s = np.ascontiguousarray(strain, dtype=np.float64)
        se = np.zeros(self.nvoi); ss = np.zeros(self.nvoi); x = np.zeros(self.nrow)
        its = C.c_int(); nl = C.c_int()
        ierr = lib().or_rve_homog_fe2(self.h, _p(s), solver, rtol, nl_max_its, nl_min_norm,
                                      _p(se), _p(ss), _p(x), C.byref(its), C.byref(nl))
        return dict(strain_ave=se, stress_ave=ss, x=x, its=its.value, nl_its=nl.value, ierr=ierr)

    def c_tangent_fd(self, strain, delta=0.005, rtol=1e-10):
        s = np.ascontiguousarray(strain, dtype=np.float64)
        c = np.zeros((self.nvoi, self.nvoi))
        lib().or_rve_c_tangent_fd(self.h, _p(s), delta, rtol, _p(c))
        return c

    def c_tangent_unit(self, amp=1e-3, rtol=1e-10):
        c = np.zeros((self.nvoi, self.nvoi)); its = C.c_int()
        ierr = lib().or_rve_c_tangent_unit(self.h, amp, rtol, _p(c), C.byref(its))
        if ierr:
            raise RuntimeError("oracle PCG did not converge")
        return c, its.value

    def homog_per_lm(self, strain):
        """BC_PER_LM (periodic, Lagrange multipliers) solved densely; small RVEs."""
        s = np.ascontiguousarray(strain, dtype=np.float64)
        se = np.zeros(self.nvoi); ss = np.zeros(self.nvoi); x = np.zeros(self.nrow)
        if lib().or_rve_homog_per_lm(self.h, _p(s), _p(se), _p(ss), _p(x)):
            raise RuntimeError("singular periodic system")
        return dict(strain_ave=se, stress_ave=ss, x=x)

    def c_tangent_per_lm(self, amp=1e-3):
        c = np.zeros((6, 6))
        for i in range(6):
            c[:, i] = self.homog_per_lm(amp * np.eye(6)[i])["stress_ave"] / amp
        return c

    def taylor(self, mode):
        c = np.zeros((self.nvoi, self.nvoi))
        if lib().or_rve_taylor(self.h, mode, _p(c)):
            raise RuntimeError("singular mixture")
        return c


def pcg(ell_ptr, b, x0=None, rtol=1e-10, max_its=100000):
    n = ell_ptr.contents.nrow
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    its = C.c_int(); rr = C.c_double()
    ierr = lib().or_pcg(ell_ptr, _p(b), _p(x), rtol, max_its, C.byref(its), C.byref(rr))
    return x, its.value, rr.value, ierr


# ------------------------------------------------------------- macro hex8

def macro_bmat(coor):
    coor = np.ascontiguousarray(coor, dtype=np.float64).reshape(8, 3)
    b = np.zeros((6, 24, 8)); dj = np.zeros(8)
    lib().or_macro_bmat(_p(coor), _p(b), _p(dj))
    return b, dj


def macro_ke(coor, c_gp):
    coor = np.ascontiguousarray(coor, dtype=np.float64).reshape(8, 3)
    c = np.ascontiguousarray(np.broadcast_to(c_gp, (8, 6, 6)), dtype=np.float64)
    ke = np.zeros((24, 24))
    lib().or_macro_ke(_p(coor), _p(c), _p(ke))
    return ke


def macro_strain(coor, ue):
    coor = np.ascontiguousarray(coor, dtype=np.float64).reshape(8, 3)
    ue = np.ascontiguousarray(ue, dtype=np.float64)
    st = np.zeros((8, 6))
    lib().or_macro_strain(_p(coor), _p(ue), _p(st))
    return st


def macro_res(coor, stress_gp):
    coor = np.ascontiguousarray(coor, dtype=np.float64).reshape(8, 3)
    s = np.ascontiguousarray(np.broadcast_to(stress_gp, (8, 6)), dtype=np.float64)
    r = np.zeros(24)
    lib().or_macro_res(_p(coor), _p(s), _p(r))
    return r


def ell_zero_rows_cols(ell_ptr, rows, diag=1.0):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    return lib().or_ell_zero_rows_cols(ell_ptr, len(rows), _p(rows, ip), diag)
