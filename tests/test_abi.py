"""CPU checks of the drop-in boundary: the shared library loads, exports every
function include/*.h declares, and its host-side ELL construction (the part
of include/ell.h that edits the caller's row-major matrix, src/ell.c) behaves
exactly like the reference on the reference's own fixtures.  No kernel runs."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from madmax_amd import _lib
from madmax_amd.engine import Ell
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    names = []
    for h in ("ell.h", "micro_api.h", "macro_api.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|double)\s+\**\s*(\w+)\s*\(", txt, flags=re.M):
            names.append(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers them all with a signature
    assert sorted(set(names)) == sorted(set(_lib.SIGNATURES))


def test_ctypes_struct_sizes_match_c_layout():
    # natural alignment on x86-64 Linux: same as the C compiler's
    assert C.sizeof(_lib.EllMatrix) == 4 * 3 + 4 + 8 + 8
    assert C.sizeof(_lib.EllSolver) == 24
    assert C.sizeof(_lib.RveSpec) == 12 + 4 + 24 + 4 + 4 + 3 * 64 + 32
    assert C.sizeof(_lib.Request) % 8 == 0


def test_host_ell_ops_match_reference_fixture():
    g = np.load(os.path.join(GOLD, "ref_ell_ops.npz"))
    e = Ell(4, 5, 3)
    L = _lib.lib()
    codes = []
    for op, r, c, v in g["ops"]:
        f = L.ell_add_val if op == 0 else L.ell_set_val
        codes.append(f(C.byref(e.m), int(r), int(c), float(v)))
    assert codes == list(g["codes"])
    for (rc_ref, v_ref), (r, c) in zip(g["getv"], [(0, 1), (0, 3), (0, 2), (1, 1), (3, 0), (2, 2),
                                                   (4, 0), (0, -1)]):
        v = C.c_double(-99.0)
        assert L.ell_get_val(C.byref(e.m), r, c, C.byref(v)) == rc_ref
        assert v.value == v_ref
    cols, vals = e.arrays()
    assert np.array_equal(cols, g["cols"]) and np.array_equal(vals, g["vals"])
    L.ell_set_zero_row(C.byref(e.m), 3, 1.0)
    L.ell_set_zero_col(C.byref(e.m), 4, 2.0)
    cols, vals = e.arrays()
    assert np.array_equal(cols, g["cols2"]) and np.array_equal(vals, g["vals2"])


def test_host_ell_rve_insertion_matches_reference_fixture():
    """The 3x3x3 RVE built with the library's ell_add_vals / zero_row /
    zero_col in the reference's element order equals the reference's ELL."""
    g = np.load(os.path.join(GOLD, "ref_ell_rve.npz"))
    n = int(g["n"])
    mesh = po.mesh(3, n, n, n)
    kes = [po.ke(3, (mesh.hx, mesh.hy, mesh.hz), g["young"][k], g["poisson"][k]) for k in range(2)]
    nrow = 3 * mesh.nn
    e = Ell(nrow, nrow, 81)
    idx = np.zeros(24, dtype=np.int32)
    for el in range(mesh.nelm):
        po.lib().or_elem_index(C.byref(mesh), el, po._p(idx, po.ip))
        assert e.add_vals(idx, idx, np.ascontiguousarray(kes[g["elem_type"][el]])) == 0
    bnd = np.zeros(mesh.nn, dtype=np.int32)
    nb = po.lib().or_boundary_nodes(C.byref(mesh), po._p(bnd, po.ip))
    L = _lib.lib()
    for d in range(3):
        for i in range(nb):
            L.ell_set_zero_row(C.byref(e.m), int(bnd[i]) * 3 + d, 1.0)
            L.ell_set_zero_col(C.byref(e.m), int(bnd[i]) * 3 + d, 1.0)
    cols, vals = e.arrays()
    assert np.array_equal(cols, g["cols"])
    assert np.array_equal(vals.view(np.uint64), g["vals"].view(np.uint64))


def test_argument_errors_without_gpu():
    L = _lib.lib()
    assert L.ell_add_vals(None, None, 0, None, 0, None) == 1
    assert L.ell_mvp(None, None, None) == 1
    assert L.homog_fe2(None, None, None) == _lib.MM_ERR_STATE  # micro not initialised
    cfg = _lib.EngineConfig()
    L.mm_engine_config_default(C.byref(cfg))
    cfg.nrhs = 5  # unsupported block width
    h = C.c_void_p()
    assert L.mm_engine_create(C.byref(cfg), C.byref(h)) == _lib.MM_ERR_ARG
    L.mm_engine_config_default(C.byref(cfg))
    assert cfg.solver_streams == 2
    cfg.solver_streams = 3  # one or two stream groups only
    assert L.mm_engine_create(C.byref(cfg), C.byref(h)) == _lib.MM_ERR_ARG
    L.mm_engine_config_default(C.byref(cfg))
    assert cfg.big_slabs == 1024 and cfg.march_planes == 8
    cfg.big_slabs = 0  # per-RVE totals from a slab count of at least 1
    assert L.mm_engine_create(C.byref(cfg), C.byref(h)) == _lib.MM_ERR_ARG
    assert L.ell_device_set_pcg_fused(None, 4) == 1
    # a column id outside [-1, ncol) is refused before anything reaches the GPU
    e = Ell.from_arrays(np.array([[0, 3]], dtype=np.int32), np.array([[1.0, 2.0]]), ncol=2)
    assert L.ell_device_create(C.byref(e.m), 0, C.byref(h)) == 1
    e = Ell.from_arrays(np.array([[0, -2]], dtype=np.int32), np.array([[1.0, 2.0]]), ncol=2)
    assert L.ell_device_create(C.byref(e.m), 0, C.byref(h)) == 1


def test_request_validation_refuses_loads_the_tangent_cannot_divide_by():
    """mm_request_validate (run by mm_engine_submit before anything reaches the
    GPU): a UNIT tangent divides the stress of load i by its component i, an
    FD tangent divides by the difference of loads i+1 and 0 in component i
    (micro/src/homogenize.c:59-94), so loads of any other shape are refused
    instead of producing a wrong tangent with status MM_OK."""
    from madmax_amd.engine import fd_loads, make_request, rve_spec, unit_loads
    L = _lib.lib()
    spec = rve_spec(4, [(0.5, 0.5, 0.3)])

    def rc(loads, mode, nrhs=None):
        rq = make_request(1, spec, loads, mode)
        return L.mm_request_validate(C.byref(rq), nrhs or len(loads))

    unit = unit_loads()
    assert rc(unit, _lib.MM_TANGENT_UNIT) == 0
    assert rc(unit + [np.array([1e-3, 2e-4, 0, 0, 0, 0])], _lib.MM_TANGENT_UNIT) == 0
    assert rc([2.5 * u for u in unit], _lib.MM_TANGENT_UNIT) == 0  # any a_i != 0
    bad = [u.copy() for u in unit]
    bad[3][0] = 1e-4  # load 3 has an off-diagonal component
    assert rc(bad, _lib.MM_TANGENT_UNIT) == _lib.MM_ERR_ARG
    bad = [u.copy() for u in unit]
    bad[5][5] = 0.0  # zero diagonal
    assert rc(bad, _lib.MM_TANGENT_UNIT) == _lib.MM_ERR_ARG
    eps = np.array([1e-3, -2e-4, 0.0, 3e-4, 0.0, 1e-4])
    fd = fd_loads(eps)
    assert rc(fd, _lib.MM_TANGENT_FD) == 0
    bad = [l.copy() for l in fd]
    bad[2][4] += 1e-6  # load 2 = eps + d e_1 must not move component 4
    assert rc(bad, _lib.MM_TANGENT_FD) == _lib.MM_ERR_ARG
    bad = [l.copy() for l in fd]
    bad[4][3] = bad[0][3]  # no step in its own component
    assert rc(bad, _lib.MM_TANGENT_FD) == _lib.MM_ERR_ARG
    assert rc(unit[:5], _lib.MM_TANGENT_UNIT) == _lib.MM_ERR_ARG  # too few loads
    assert rc(unit, _lib.MM_TANGENT_UNIT, nrhs=5) == _lib.MM_ERR_ARG  # more loads than the block
    assert rc(unit, 9) == _lib.MM_ERR_ARG  # unknown tangent mode
    assert rc([eps], _lib.MM_TANGENT_NONE) == 0
    nan = eps.copy()
    nan[2] = float("nan")
    assert rc([nan], _lib.MM_TANGENT_NONE) == _lib.MM_ERR_ARG
