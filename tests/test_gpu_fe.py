"""GPU parity of the element-assembly path (fe.hip) against the oracle:
  nonlinear request path  micro with linear_materials = 0: every request runs
      homog_fe2's Newton loop with the Jacobian re-assembled element by
      element on the GPU; the assembled ELL equals the oracle's bit for bit,
      stresses and FD tangents meet 1e-8, Newton iteration counts match
  periodic BC             micro with BC_PER_LM: stresses and tangent equal
      the oracle's Lagrange-multiplier solve within 1e-8, a homogeneous cell
      gives C, <eps> = eps_mac, and the tangent lies below the USTRAIN one
  macro assembly          isoparametric hex8 on a distorted mesh: ELL,
      gauss-point strains and residual bit for bit, the solve within 1e-9
  FE^2                    one macro Newton step whose gauss points are solved
      by the batched engine, against the same step assembled by the oracle
  VTU                     micro_write_vtu writes the reference's fields."""
import ctypes as C
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from madmax_amd import (MICRO_SOLVER_PCG, MM_PREC_FP64, MM_TANGENT_UNIT, Engine, configs, lib, rve_spec,
                        unit_loads)
from madmax_amd import _lib
from madmax_amd.engine import Micro
from oracle import pyoracle as po
from tests.helpers import normwise_rel, oracle_rve, same_bits

pytestmark = pytest.mark.gpu

TOL = 1e-8


def _micro(spec, **kw):
    return Micro(spec, solver=MICRO_SOLVER_PCG, rtol=1e-12, **kw)


@pytest.mark.parametrize("n", [4, 10])
def test_nonlinear_path_reassembles_bitexact(n):
    spec = configs.c1_spec() if n == 10 else rve_spec(4, [(0.45, 0.55, 0.3)])
    ref, _ = oracle_rve(spec)
    eps = np.array([1e-3, -3e-4, 2e-4, 1e-4, 0.0, -2e-4])
    print("nonlinear path n", n, flush=True)
    with _micro(spec, linear_materials=0) as mi:
        se, ss = mi.strain_stress(eps)  # no linear shortcut: homog_fe2 runs
        st = mi.stats()
        cols, vals = mi.ell()
        c = mi.c_tangent(eps)  # homog_calculate_c_tangent: 7 more homog_fe2
        st2 = mi.stats()
    out = ref.homog(eps, rtol=1e-12)
    cols_ref, vals_ref = ref.assemble()
    assert np.array_equal(cols, cols_ref)
    assert same_bits(vals, vals_ref)
    assert normwise_rel(ss, out["stress_ave"]) <= TOL
    assert normwise_rel(se, out["strain_ave"]) <= 1e-12
    assert st["sweeps"] == out["nl_its"] and st["assemblies"] == out["nl_its"]
    assert st2["assemblies"] == st["assemblies"] + 7 * out["nl_its"]  # re-assembled every iteration
    c_ref = ref.c_tangent_fd(eps, 0.005, 1e-12)
    assert normwise_rel(c, c_ref) <= TOL


@pytest.mark.parametrize("n", [4, 6])
def test_periodic_bc_vs_lagrange_multiplier_oracle(n):
    spec = rve_spec(n, [(0.5, 0.45, 0.3)])
    ref, _ = oracle_rve(spec)
    eps = np.array([1e-3, 2e-4, -3e-4, 4e-4, -1e-4, 2e-4])
    print("periodic n", n, flush=True)
    with _micro(spec, fe2_bc=_lib.MICRO_BC_PER_LM) as mi:
        c = mi.c_tangent(eps)  # c_tangent_linear from the periodic solves at init
        se, ss = mi.fe2(eps)
        x = mi.displacement()
        nrow, _, _ = mi.sizes()
    assert nrow == 3 * n ** 3  # unknowns on the torus
    assert x[0] == 0.0 and x[1] == 0.0 and x[2] == 0.0  # pinned corner fluctuation
    out = ref.homog_per_lm(eps)
    assert normwise_rel(ss, out["stress_ave"]) <= TOL
    assert normwise_rel(se, eps) <= 1e-12
    c_ref = ref.c_tangent_per_lm()
    assert normwise_rel(c, c_ref) <= TOL
    cu, _ = ref.c_tangent_unit(1e-3, 1e-12)
    assert np.min(np.linalg.eigvalsh(0.5 * ((cu - c) + (cu - c).T))) >= -1e-8 * np.max(np.abs(cu))


def test_periodic_homogeneous_cell():
    spec = rve_spec(4, [], young=(2e7, 2e7), poisson=(0.3, 0.3))
    with _micro(spec, fe2_bc=_lib.MICRO_BC_PER_LM) as mi:
        c = mi.c_tangent(np.zeros(6))
    assert normwise_rel(c, po.material_c(3, 2e7, 0.3)) <= 1e-10


def test_ustress_is_refused():
    cfg = _lib.MicroConfig()
    lib().micro_config_default(C.byref(cfg))
    cfg.fe2_bc = _lib.MICRO_BC_USTRESS
    assert lib().micro_init(C.byref(cfg)) == _lib.MM_ERR_ARG


def _macro_mesh(n, seed=5):
    rng = np.random.default_rng(seed)
    nx = n + 1
    X = np.array([[ix, iy, iz] for iz in range(nx) for iy in range(nx) for ix in range(nx)], dtype=float) / n
    inner = [(ix, iy, iz) for iz in range(1, n) for iy in range(1, n) for ix in range(1, n)]
    for p in inner:
        X[p[0] + p[1] * nx + p[2] * nx * nx] += rng.uniform(-0.2, 0.2, 3) / n
    return X


def _elem_nodes(e, n):
    nx = n + 1
    ex, ey, ez = e % n, (e // n) % n, e // (n * n)
    n0 = ex + ey * nx + ez * nx * nx
    return [n0, n0 + 1, n0 + nx + 1, n0 + nx, n0 + nx * nx, n0 + nx * nx + 1, n0 + nx * nx + nx + 1, n0 + nx * nx + nx]


class Macro:
    def __init__(self, n, coords):
        self.h = C.c_void_p()
        _lib.check(lib().mm_macro_create(n, n, n, 1.0, 1.0, 1.0, coords.ctypes.data_as(_lib.dp), 0,
                                         C.byref(self.h)), "mm_macro_create")
        nr, ne = C.c_int(), C.c_int()
        lib().mm_macro_sizes(self.h, C.byref(nr), C.byref(ne))
        self.nrow, self.nelm = nr.value, ne.value

    def close(self):
        lib().mm_macro_destroy(self.h)


def _oracle_macro(n, X, c_gp, dirs):
    nelm = n ** 3
    nrow = 3 * (n + 1) ** 3
    K = po.OrEll()
    po.lib().or_ell_init(C.byref(K), nrow, nrow, 81)
    for e in range(nelm):
        nods = _elem_nodes(e, n)
        idx = np.array([3 * a + d for a in nods for d in range(3)], dtype=np.int32)
        ke = po.macro_ke(X[nods], c_gp[e])
        po.lib().or_ell_add_vals(C.byref(K), po._p(idx, po.ip), 24, po._p(idx, po.ip), 24, po._p(ke))
    po.ell_zero_rows_cols(C.byref(K), dirs, 1.0)
    cols, vals = po.ell_to_numpy(K)
    po.lib().or_ell_free(C.byref(K))
    return cols, vals


def _dense(cols, vals):
    A = np.zeros((cols.shape[0], cols.shape[0]))
    for i in range(cols.shape[0]):
        for j in range(cols.shape[1]):
            if cols[i, j] < 0:
                break
            A[i, cols[i, j]] += vals[i, j]
    return A


def test_macro_assembly_bitexact_on_distorted_mesh():
    n = 3
    X = _macro_mesh(n)
    rng = np.random.default_rng(7)
    nelm = n ** 3
    c0 = po.material_c(3, 2e7, 0.28)
    c_gp = np.array([[c0 * (1 + 0.3 * rng.uniform()) for _ in range(8)] for _ in range(nelm)])
    sig = rng.standard_normal((nelm, 8, 6)) * 1e4
    u = rng.standard_normal(3 * (n + 1) ** 3) * 1e-3
    dirs = np.array([3 * k + d for k in range((n + 1) ** 3) if k % (n + 1) == 0 for d in range(3)], dtype=np.int32)
    m = Macro(n, X)
    try:
        _lib.check(lib().mm_macro_set_dirichlet(m.h, len(dirs), dirs.ctypes.data_as(_lib.ip)), "dirichlet")
        _lib.check(lib().mm_macro_assemble_a(m.h, np.ascontiguousarray(c_gp).ctypes.data_as(_lib.dp)), "A")
        cols = np.zeros((m.nrow, 81), dtype=np.int32)
        vals = np.zeros((m.nrow, 81))
        _lib.check(lib().mm_macro_get_ell(m.h, cols.ctypes.data_as(_lib.ip), vals.ctypes.data_as(_lib.dp)), "ell")
        _lib.check(lib().mm_macro_set_u(m.h, u.ctypes.data_as(_lib.dp)), "u")
        eps = np.zeros((nelm, 8, 6))
        _lib.check(lib().mm_macro_strain(m.h, eps.ctypes.data_as(_lib.dp)), "strain")
        b = np.zeros(m.nrow)
        nrm = C.c_double()
        _lib.check(lib().mm_macro_assemble_b(m.h, np.ascontiguousarray(sig).ctypes.data_as(_lib.dp),
                                             b.ctypes.data_as(_lib.dp), C.byref(nrm)), "b")
        du = np.zeros(m.nrow)
        its = C.c_int()
        rc = lib().mm_macro_solve(m.h, b.ctypes.data_as(_lib.dp), du.ctypes.data_as(_lib.dp), 1e-12, C.byref(its))
    finally:
        m.close()
    cols_ref, vals_ref = _oracle_macro(n, X, c_gp, dirs)
    assert np.array_equal(cols, cols_ref)
    assert same_bits(vals, vals_ref)
    res_ref = np.zeros(m.nrow)
    for e in range(nelm):
        nods = _elem_nodes(e, n)
        ue = np.array([u[3 * a + d] for a in nods for d in range(3)])
        assert same_bits(eps[e], po.macro_strain(X[nods], ue))
        r = po.macro_res(X[nods], sig[e])
        for a, nd in enumerate(nods):
            for d in range(3):
                res_ref[3 * nd + d] += r[3 * a + d]
    res_ref[dirs] = 0.0
    assert same_bits(b, -res_ref)
    assert abs(nrm.value - np.linalg.norm(res_ref)) <= 1e-12 * np.linalg.norm(res_ref)
    assert rc == 0
    A = _dense(cols_ref, vals_ref)
    du_ref = np.linalg.solve(A, -res_ref)
    assert normwise_rel(du, du_ref) <= 1e-9


def test_fe2_macro_step_with_engine_gauss_points():
    """One FE^2 Newton step: a 2x2x2 macro cube, x-face clamped and the
    opposite face pulled; every gauss point's strain is sent to the batched
    engine (6 unit loads + the strain), whose tangents and stresses are
    assembled on the GPU into the macro system and solved.  The same step
    with the oracle's tangent of the RVE gives the same update."""
    n = 2
    nx = n + 1
    X = np.array([[ix, iy, iz] for iz in range(nx) for iy in range(nx) for ix in range(nx)], dtype=float) / n
    spec = rve_spec(6, [(0.5, 0.5, 0.3)])
    ref, _ = oracle_rve(spec)
    c_hom, _ = ref.c_tangent_unit(1e-3, 1e-12)
    nrow = 3 * nx ** 3
    left = [k for k in range(nx ** 3) if k % nx == 0]
    right = [k for k in range(nx ** 3) if k % nx == n]
    dirs = np.array(sorted([3 * k + d for k in left + right for d in range(3)]), dtype=np.int32)
    u = np.zeros(nrow)
    for k in right:
        u[3 * k] = 1e-3  # imposed stretch
    m = Macro(n, X)
    try:
        _lib.check(lib().mm_macro_set_dirichlet(m.h, len(dirs), dirs.ctypes.data_as(_lib.ip)), "dirichlet")
        _lib.check(lib().mm_macro_set_u(m.h, u.ctypes.data_as(_lib.dp)), "u")
        nelm = m.nelm
        eps = np.zeros((nelm, 8, 6))
        _lib.check(lib().mm_macro_strain(m.h, eps.ctypes.data_as(_lib.dp)), "strain")
        c_gp = np.zeros((nelm, 8, 6, 6))
        sig = np.zeros((nelm, 8, 6))
        with Engine(nrhs=7, precision=MM_PREC_FP64, hbm_budget=1 << 30, max_rves=80, inner_rtol=1e-12,
                    rtol=1e-12, max_sweeps=3) as eng:
            for e in range(nelm):
                for g in range(8):
                    assert eng.submit(e * 8 + g + 1, spec, unit_loads() + [eps[e, g]], MM_TANGENT_UNIT) == 0
            eng.drain()
            for r in eng.poll(128):
                e, g = divmod(r["id"] - 1, 8)
                assert r["status"] == 0
                c_gp[e, g] = r["c_tangent"]
                sig[e, g] = r["stress"][6]
        _lib.check(lib().mm_macro_assemble_a(m.h, np.ascontiguousarray(c_gp).ctypes.data_as(_lib.dp)), "A")
        b = np.zeros(nrow)
        nrm = C.c_double()
        _lib.check(lib().mm_macro_assemble_b(m.h, np.ascontiguousarray(sig).ctypes.data_as(_lib.dp),
                                             b.ctypes.data_as(_lib.dp), C.byref(nrm)), "b")
        du = np.zeros(nrow)
        its = C.c_int()
        assert lib().mm_macro_solve(m.h, b.ctypes.data_as(_lib.dp), du.ctypes.data_as(_lib.dp), 1e-12,
                                    C.byref(its)) == 0
    finally:
        m.close()
    # the oracle's step: same mesh, the oracle tangent at every gauss point
    c_ref = np.broadcast_to(c_hom, (n ** 3, 8, 6, 6))
    cols_ref, vals_ref = _oracle_macro(n, X, c_ref, dirs)
    res_ref = np.zeros(nrow)
    for e in range(n ** 3):
        nods = _elem_nodes(e, n)
        ue = np.array([u[3 * a + d] for a in nods for d in range(3)])
        st = po.macro_strain(X[nods], ue)
        r = po.macro_res(X[nods], st @ c_hom.T)
        for a, nd in enumerate(nods):
            res_ref[3 * nd:3 * nd + 3] += r[3 * a:3 * a + 3]
    res_ref[dirs] = 0.0
    du_ref = np.linalg.solve(_dense(cols_ref, vals_ref), -res_ref)
    assert normwise_rel(b, -res_ref) <= TOL
    assert normwise_rel(du, du_ref) <= TOL


def test_micro_write_vtu(tmp_path):
    spec = rve_spec(4, [(0.5, 0.5, 0.3)])
    eps = np.array([1e-3, 0, 0, 0, 2e-4, 0])
    name = str(tmp_path / "micro")
    with _micro(spec) as mi:
        mi.fe2(eps)
        x = mi.displacement()
        et = mi.elem_type()
        assert lib().micro_write_vtu(name.encode()) == 0
    assert os.path.exists(name + ".pvtu")
    root = ET.parse(name + "_0.vtu").getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == 125 and int(piece.get("NumberOfCells")) == 64
    arrays = {a.get("Name"): np.array(a.text.split(), dtype=float) for a in root.iter("DataArray")}
    assert normwise_rel(arrays["displ"], x) <= 1e-5  # printed with %e
    assert np.array_equal(arrays["elem_type"].astype(int), et)
    assert arrays["stress"].shape == (64 * 6,) and arrays["residual"].shape == (125 * 3,)


@pytest.mark.parametrize("bc", [_lib.MICRO_BC_USTRAIN, _lib.MICRO_BC_PER_LM])
def test_micro_residual_bitexact(bc):
    """assembly_res_ell on the GPU (micro_residual) equals the oracle's
    restatement bit for bit for USTRAIN; for the periodic condition the
    residual of the fluctuation is compared with the oracle residual of the
    full displacement (images summed onto their masters) within rounding."""
    spec = rve_spec(5, [(0.5, 0.45, 0.3)])
    ref, _ = oracle_rve(spec)
    rng = np.random.default_rng(11)
    eps = np.array([1e-3, -2e-4, 3e-4, 1e-4, 2e-4, -1e-4])
    with _micro(spec, fe2_bc=bc) as mi:
        nrow, _, _ = mi.sizes()
        x = rng.standard_normal(nrow) * 1e-4
        if bc == _lib.MICRO_BC_PER_LM:
            x[:3] = 0.0
        r, nrm = mi.residual(x, eps)
    if bc == _lib.MICRO_BC_USTRAIN:
        r_ref, n_ref = ref.residual(x)
        assert same_bits(r, r_ref)
        assert abs(nrm - n_ref) <= 1e-13 * n_ref
        return
    # periodic: full displacement u = E X + x(wrapped); oracle residual of u
    n = 5
    nx = n + 1
    u = np.zeros(3 * nx ** 3)
    for iz in range(nx):
        for iy in range(nx):
            for ix in range(nx):
                X = np.array([ix, iy, iz]) / n
                aff = np.array([eps[0] * X[0] + eps[3] / 2 * X[1] + eps[5] / 2 * X[2],
                                eps[3] / 2 * X[0] + eps[1] * X[1] + eps[4] / 2 * X[2],
                                eps[5] / 2 * X[0] + eps[4] / 2 * X[1] + eps[2] * X[2]])
                m = (ix % n) + (iy % n) * n + (iz % n) * n * n
                k = ix + iy * nx + iz * nx * nx
                u[3 * k:3 * k + 3] = x[3 * m:3 * m + 3] + aff
    # unconstrained element sums: zero no dof, then fold images onto masters
    m_ = po.mesh(3, n, n, n)
    full = np.zeros(3 * nx ** 3)
    b, wp = po.struct_bmat(3, np.array([1.0 / n] * 3))
    et = po.phase_map(m_, [(0.5, 0.45, 0.3)])
    cm = [po.material_c(3, 1e7, 0.25), po.material_c(3, 1e8, 0.3)]
    for e in range(n ** 3):
        idx = np.zeros(24, dtype=np.int32)
        po.lib().or_elem_index(C.byref(m_), e, po._p(idx, po.ip))
        ue = u[idx]
        re = np.zeros(24)
        for g in range(8):
            st = cm[et[e]] @ (b[:, :, g] @ ue)
            re += b[:, :, g].T @ st * wp[g]
        full[idx] += re
    folded = np.zeros(3 * n ** 3)
    for iz in range(nx):
        for iy in range(nx):
            for ix in range(nx):
                m = (ix % n) + (iy % n) * n + (iz % n) * n * n
                k = ix + iy * nx + iz * nx * nx
                folded[3 * m:3 * m + 3] += full[3 * k:3 * k + 3]
    folded[:3] = 0.0
    assert normwise_rel(r, folded) <= 1e-12
