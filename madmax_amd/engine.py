"""Python mirror of the C-ABI: RVE specs, the batched / serving engine, the
process-global micro entry points and the device ELL.

Every numerical call is a call into libmadmax_amd.so (HIP on the GPU).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (MM_MAX_FIBERS, MM_MAX_RHS, MM_NVOI, EllMatrix, EllSolver, EngineConfig,
                   EngineInfo, MicroConfig, MicroStats, Request, Result, RveSpec, check, lib)


def _dp(a):
    return a.ctypes.data_as(_lib.dp)


def _ip(a):
    return a.ctypes.data_as(_lib.ip)


def rve_spec(n, fibers=(), young=(1.0e7, 1.0e8), poisson=(0.25, 0.3), lengths=(1.0, 1.0, 1.0)):
    """An RVE: n (int or (nex, ney, nez)) hex8 elements per side, z-fibers
    given as (x, y, r) triples, MATRIX / FIBER moduli."""
    s = RveSpec()
    nex, ney, nez = (n, n, n) if isinstance(n, (int, np.integer)) else tuple(int(v) for v in n)
    s.nex, s.ney, s.nez = nex, ney, nez
    s.lx, s.ly, s.lz = (float(v) for v in lengths)
    fibers = list(fibers)
    if len(fibers) > MM_MAX_FIBERS:
        raise ValueError("at most %d fibers" % MM_MAX_FIBERS)
    s.nfib = len(fibers)
    for i, (x, y, r) in enumerate(fibers):
        s.fib_x[i], s.fib_y[i], s.fib_r[i] = float(x), float(y), float(r)
    s.young[0], s.young[1] = float(young[0]), float(young[1])
    s.poisson[0], s.poisson[1] = float(poisson[0]), float(poisson[1])
    return s


def spec_fibers(s):
    return [(s.fib_x[i], s.fib_y[i], s.fib_r[i]) for i in range(s.nfib)]


def unit_loads(amp=1.0e-3):
    """The six unit-strain load cases amp * e_i of one tangent."""
    return [amp * np.eye(MM_NVOI)[i] for i in range(MM_NVOI)]


def fd_loads(strain, delta=_lib.HOMOGENIZE_DELTA_STRAIN):
    """sigma(eps) and sigma(eps + delta e_i) (micro/src/homogenize.c:59-94)."""
    e = np.asarray(strain, dtype=np.float64)
    return [e.copy()] + [e + delta * np.eye(MM_NVOI)[i] for i in range(MM_NVOI)]


def make_request(rid, spec, loads, tangent_mode=_lib.MM_TANGENT_NONE, warm_key=0, warm_start=0):
    rq = Request()
    rq.id = int(rid)
    rq.rve = spec
    loads = [np.asarray(l, dtype=np.float64) for l in loads]
    if not 1 <= len(loads) <= MM_MAX_RHS:
        raise ValueError("1..%d load cases" % MM_MAX_RHS)
    rq.nload = len(loads)
    for k, l in enumerate(loads):
        for i in range(MM_NVOI):
            rq.strain[k][i] = float(l[i])
    rq.tangent_mode = tangent_mode
    rq.warm_key = int(warm_key)
    rq.warm_start = int(warm_start)
    return rq


def result_dict(r, nload):
    return dict(id=int(r.id), status=int(r.status), pcg_its=int(r.pcg_its), sweeps=int(r.sweeps),
                residual_rel=float(r.residual_rel),
                stress=np.array([[r.stress[k][i] for i in range(MM_NVOI)] for k in range(nload)]),
                strain=np.array([[r.strain[k][i] for i in range(MM_NVOI)] for k in range(nload)]),
                c_tangent=np.array([r.c_tangent[i] for i in range(36)]).reshape(6, 6),
                t_submit=r.t_submit, t_admit=r.t_admit, t_done=r.t_done)


class Engine:
    """Batched / serving RVE engine (mm_engine_*)."""

    def __init__(self, nrhs=6, precision=_lib.MM_PREC_MIXED, hbm_budget=0, max_rves=0,
                 slab_nodes=0, rtol=1e-10, abs_tol=0.0, inner_rtol=1e-5, max_inner_its=20000,
                 max_sweeps=8, steps_per_launch=16, use_graph=True, sync_every_step=False,
                 device=0, ell_storage=_lib.MM_ELL_DICT, initial_guess=_lib.MM_GUESS_AFFINE,
                 solver_streams=2, profile_step_kernel=False, march_planes=None,
                 big_slabs=None):
        L = lib()
        c = EngineConfig()
        check(L.mm_engine_config_default(C.byref(c)), "mm_engine_config_default")
        c.device = device
        c.hbm_budget = int(hbm_budget)
        c.max_rves = int(max_rves)
        c.slab_nodes = int(slab_nodes)
        c.nrhs = int(nrhs)
        c.precision = int(precision)
        c.rtol = rtol
        c.abs_tol = abs_tol
        c.inner_rtol = inner_rtol
        c.max_inner_its = int(max_inner_its)
        c.max_sweeps = int(max_sweeps)
        c.steps_per_launch = int(steps_per_launch)
        c.use_graph = int(bool(use_graph))
        c.sync_every_step = int(bool(sync_every_step))
        c.ell_storage = int(ell_storage)
        c.initial_guess = int(initial_guess)
        c.solver_streams = int(solver_streams)
        c.profile_step_kernel = int(bool(profile_step_kernel))
        if march_planes is not None:
            c.march_planes = int(march_planes)
        if big_slabs is not None:
            c.big_slabs = int(big_slabs)
        self.cfg = c
        self.nrhs = nrhs
        h = C.c_void_p()
        check(L.mm_engine_create(C.byref(c), C.byref(h)), "mm_engine_create")
        self.h = h
        self._nload = {}

    def close(self):
        if getattr(self, "h", None):
            lib().mm_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, rid, spec, loads, tangent_mode=_lib.MM_TANGENT_NONE, warm_key=0,
               warm_start=0):
        rq = make_request(rid, spec, loads, tangent_mode, warm_key, warm_start)
        rc = lib().mm_engine_submit(self.h, C.byref(rq))
        if rc == 0:
            self._nload[int(rid)] = rq.nload
        return rc

    def step(self):
        return check(lib().mm_engine_step(self.h), "mm_engine_step")

    def drain(self):
        return check(lib().mm_engine_drain(self.h), "mm_engine_drain")

    def start(self):
        return check(lib().mm_engine_start(self.h), "mm_engine_start")

    def stop(self):
        return lib().mm_engine_stop(self.h)

    def now(self):
        return lib().mm_engine_now(self.h)

    def poll(self, max_results=1024):
        buf = getattr(self, "_buf", None)
        if buf is None or len(buf) < max_results:
            buf = self._buf = (Result * max_results)()
        n = C.c_int()
        check(lib().mm_engine_poll(self.h, buf, max_results, C.byref(n)), "mm_engine_poll")
        out = []
        for i in range(n.value):
            r = buf[i]
            out.append(result_dict(r, self._nload.pop(int(r.id), MM_MAX_RHS)))
        return out

    def info(self):
        i = EngineInfo()
        check(lib().mm_engine_info_get(self.h, C.byref(i)), "mm_engine_info_get")
        return {f: getattr(i, f) for f, _ in EngineInfo._fields_}

    def step_kernel_times(self, max_n=1 << 16):
        """Durations (ms) of the step-kernel launches recorded since the last
        call (engine created with profile_step_kernel=True)."""
        buf = np.zeros(max_n)
        n = C.c_int()
        check(lib().mm_engine_step_kernel_times(self.h, _dp(buf), max_n, C.byref(n)), "step_kernel_times")
        return buf[: n.value].copy()

    def rve_sizes(self, key):
        nr, ne = C.c_int(), C.c_int()
        check(lib().mm_engine_rve_sizes(self.h, key, C.byref(nr), C.byref(ne)), "mm_engine_rve_sizes")
        return nr.value, ne.value

    def download_ell(self, key):
        nrow, _ = self.rve_sizes(key)
        cols = np.zeros((nrow, 81), dtype=np.int32)
        vals = np.zeros((nrow, 81))
        check(lib().mm_engine_download_ell(self.h, key, _ip(cols), _dp(vals)), "download_ell")
        return cols, vals

    def download_phase(self, key):
        _, nelm = self.rve_sizes(key)
        et = np.zeros(nelm, dtype=np.int32)
        check(lib().mm_engine_download_phase(self.h, key, _ip(et)), "download_phase")
        return et

    def download_x(self, key, load=0):
        nrow, _ = self.rve_sizes(key)
        x = np.zeros(nrow)
        check(lib().mm_engine_download_x(self.h, key, load, _dp(x)), "download_x")
        return x

    def time_spmv(self, reps=20):
        ms, by = C.c_double(), C.c_double()
        check(lib().mm_engine_time_spmv(self.h, reps, C.byref(ms), C.byref(by)), "time_spmv")
        return ms.value, by.value

    def time_step(self, reps=20):
        """Average duration (ms) and algorithmic bytes of the PCG step's SpMV
        kernel (SpMV + p/q update) over the resident RVEs."""
        ms, by = C.c_double(), C.c_double()
        check(lib().mm_engine_time_step(self.h, reps, C.byref(ms), C.byref(by)), "time_step")
        return ms.value, by.value


class Micro:
    """The reference's process-global micro (micro_init + homog_* entry points)."""

    def __init__(self, spec, solver=_lib.MICRO_SOLVER_PCG, rtol=1e-10, nl_max_its=2,
                 nl_min_norm=1e-4, device=0, linear_materials=1, fe2_bc=_lib.MICRO_BC_USTRAIN):
        L = lib()
        c = MicroConfig()
        check(L.micro_config_default(C.byref(c)), "micro_config_default")
        c.rve = spec
        c.solver = solver
        c.rtol = rtol
        c.nl_max_its = nl_max_its
        c.nl_min_norm = nl_min_norm
        c.device = device
        c.linear_materials = int(linear_materials)
        c.fe2_bc = int(fe2_bc)
        check(L.micro_init(C.byref(c)), "micro_init")
        self.spec = spec

    def close(self):
        lib().micro_finalize()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def fe2(self, strain):
        e = np.ascontiguousarray(strain, dtype=np.float64)
        se, ss = np.zeros(6), np.zeros(6)
        check(lib().homog_fe2(_dp(e), _dp(se), _dp(ss)), "homog_fe2")
        return se, ss

    def strain_stress(self, strain):
        e = np.ascontiguousarray(strain, dtype=np.float64)
        se, ss = np.zeros(6), np.zeros(6)
        check(lib().homog_get_strain_stress(_dp(e), _dp(se), _dp(ss)), "homog_get_strain_stress")
        return se, ss

    def c_tangent(self, strain):
        e = np.ascontiguousarray(strain, dtype=np.float64)
        p = _lib.dp()
        check(lib().homog_get_c_tangent(_dp(e), C.byref(p)), "homog_get_c_tangent")
        return np.array([p[i] for i in range(36)]).reshape(6, 6)

    def calculate_c_tangent(self, strain):
        e = np.ascontiguousarray(strain, dtype=np.float64)
        c = np.zeros(36)
        check(lib().homog_calculate_c_tangent(_dp(e), _dp(c)), "homog_calculate_c_tangent")
        return c.reshape(6, 6)

    def stats(self):
        s = MicroStats()
        check(lib().micro_get_stats(C.byref(s)), "micro_get_stats")
        return dict(pcg_its=s.pcg_its, sweeps=s.sweeps, residual=s.residual, seconds=s.seconds,
                    assemblies=s.assemblies)

    def sizes(self):
        nr, nz, ne = C.c_int(), C.c_int(), C.c_int()
        check(lib().micro_get_sizes(C.byref(nr), C.byref(nz), C.byref(ne)), "micro_get_sizes")
        return nr.value, nz.value, ne.value

    def ell(self):
        nr, nz, _ = self.sizes()
        cols = np.zeros((nr, nz), dtype=np.int32)
        vals = np.zeros((nr, nz))
        check(lib().micro_get_ell(_ip(cols), _dp(vals)), "micro_get_ell")
        return cols, vals

    def elem_type(self):
        _, _, ne = self.sizes()
        et = np.zeros(ne, dtype=np.int32)
        check(lib().micro_get_elem_type(_ip(et)), "micro_get_elem_type")
        return et

    def residual(self, x, strain=None):
        """assembly_res_ell on the GPU for displacement x: (r, |r|)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(len(x))
        nrm = C.c_double()
        sp = _dp(np.ascontiguousarray(strain, dtype=np.float64)) if strain is not None else None
        check(lib().micro_residual(sp, _dp(x), _dp(r), C.byref(nrm)), "micro_residual")
        return r, nrm.value

    def displacement(self):
        nr, _, _ = self.sizes()
        x = np.zeros(nr)
        check(lib().micro_get_displacement(_dp(x)), "micro_get_displacement")
        return x


class Ell:
    """A host ELL matrix driven through the reference API (include/ell.h)."""

    def __init__(self, nrow, ncol, nnz):
        self.m = EllMatrix()
        check(lib().ell_init(C.byref(self.m), nrow, ncol, nnz), "ell_init")

    @classmethod
    def from_arrays(cls, cols, vals, ncol=None):
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        nrow, nnz = cols.shape
        e = cls(nrow, nrow if ncol is None else ncol, nnz)
        n = nrow * nnz
        C.memmove(e.m.cols, cols.ctypes.data, n * 4)
        C.memmove(e.m.vals, vals.ctypes.data, n * 8)
        return e

    def arrays(self):
        n = self.m.nrow * self.m.nnz
        cols = np.ctypeslib.as_array(self.m.cols, shape=(n,)).reshape(self.m.nrow, self.m.nnz).copy()
        vals = np.ctypeslib.as_array(self.m.vals, shape=(n,)).reshape(self.m.nrow, self.m.nnz).copy()
        return cols, vals

    def add_vals(self, ix, iy, vals):
        ix = np.ascontiguousarray(ix, dtype=np.int32)
        iy = np.ascontiguousarray(iy, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        return lib().ell_add_vals(C.byref(self.m), _ip(ix), len(ix), _ip(iy), len(iy), _dp(v))

    def mvp(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.m.nrow)
        check(lib().ell_mvp(C.byref(self.m), _dp(x), _dp(y)), "ell_mvp")
        return y

    def solve_pcg(self, b, x0=None, rtol=1e-10, max_its=100000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.m.nrow) if x0 is None else np.array(x0, dtype=np.float64)
        s = EllSolver(max_its, 0, rtol, 0.0)
        rc = lib().ell_solve_pcg(C.byref(s), C.byref(self.m), _dp(b), _dp(x))
        return x, s.its, s.err, rc

    def solve_jacobi(self, b, x0=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.m.nrow) if x0 is None else np.array(x0, dtype=np.float64)
        s = EllSolver(0, 0, 0.0, 0.0)
        check(lib().ell_solve_jacobi(C.byref(s), C.byref(self.m), _dp(b), _dp(x)), "ell_solve_jacobi")
        return x, s.its, s.err

    def free(self):
        if self.m.cols:
            lib().ell_free(C.byref(self.m))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Macro:
    """Macro hex8 mesh assembled on the GPU from gauss-point tangents and
    stresses (include/macro_api.h)."""

    def __init__(self, n, lengths=(1.0, 1.0, 1.0), coords=None, device=0):
        nex, ney, nez = (n, n, n) if isinstance(n, (int, np.integer)) else tuple(int(v) for v in n)
        self.h = C.c_void_p()
        cp = None
        if coords is not None:
            self._coords = np.ascontiguousarray(coords, dtype=np.float64)
            cp = _dp(self._coords)
        check(lib().mm_macro_create(nex, ney, nez, float(lengths[0]), float(lengths[1]), float(lengths[2]), cp,
                                    device, C.byref(self.h)), "mm_macro_create")
        nr, ne = C.c_int(), C.c_int()
        lib().mm_macro_sizes(self.h, C.byref(nr), C.byref(ne))
        self.nrow, self.nelm = nr.value, ne.value

    def close(self):
        if getattr(self, "h", None):
            lib().mm_macro_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_dirichlet(self, dofs):
        d = np.ascontiguousarray(dofs, dtype=np.int32)
        check(lib().mm_macro_set_dirichlet(self.h, len(d), _ip(d)), "mm_macro_set_dirichlet")

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        check(lib().mm_macro_set_u(self.h, _dp(u)), "mm_macro_set_u")

    def strain(self):
        out = np.zeros((self.nelm, 8, 6))
        check(lib().mm_macro_strain(self.h, _dp(out)), "mm_macro_strain")
        return out

    def assemble_b(self, stress_gp):
        s = np.ascontiguousarray(stress_gp, dtype=np.float64)
        b = np.zeros(self.nrow)
        nrm = C.c_double()
        check(lib().mm_macro_assemble_b(self.h, _dp(s), _dp(b), C.byref(nrm)), "mm_macro_assemble_b")
        return b, nrm.value

    def assemble_a(self, c_gp):
        c = np.ascontiguousarray(c_gp, dtype=np.float64)
        check(lib().mm_macro_assemble_a(self.h, _dp(c)), "mm_macro_assemble_a")

    def ell(self):
        cols = np.zeros((self.nrow, 81), dtype=np.int32)
        vals = np.zeros((self.nrow, 81))
        check(lib().mm_macro_get_ell(self.h, _ip(cols), _dp(vals)), "mm_macro_get_ell")
        return cols, vals

    def solve(self, b, rtol=1e-10):
        b = np.ascontiguousarray(b, dtype=np.float64)
        du = np.zeros(self.nrow)
        its = C.c_int()
        rc = lib().mm_macro_solve(self.h, _dp(b), _dp(du), rtol, C.byref(its))
        if rc not in (0, _lib.MM_NOT_CONVERGED):
            check(rc, "mm_macro_solve")
        return du, its.value, rc
