"""ctypes binding of libmadmax_amd.so (include/ell.h, include/micro_api.h).

The shared library is the product: every compute call goes through its HIP
kernels.  Importing this module fails loudly when the library is missing;
there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmadmax_amd.so")

MM_NVOI = 6
MM_MAX_FIBERS = 8
MM_MAX_RHS = 8
HOMOGENIZE_DELTA_STRAIN = 0.005

MM_OK, MM_ERR_ARG, MM_ERR_ALLOC, MM_ERR_FULL = 0, 1, 2, 3
MM_NOT_CONVERGED, MM_ERR_HIP, MM_ERR_STATE, MM_REJECTED = 5, 6, 7, 8
MICRO_SOLVER_PCG, MICRO_SOLVER_PCG_MIXED = 0, 1
MICRO_BC_USTRAIN, MICRO_BC_USTRESS, MICRO_BC_PER_LM = 1, 2, 3
MM_PREC_FP64, MM_PREC_MIXED = 0, 1
MM_ELL_FULL, MM_ELL_DICT = 0, 1
MM_GUESS_ZERO, MM_GUESS_AFFINE = 0, 1
MM_TANGENT_NONE, MM_TANGENT_UNIT, MM_TANGENT_FD = 0, 1, 2

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class EllMatrix(C.Structure):
    """inc/ell.h:15-21"""
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("nnz", C.c_int),
                ("cols", ip), ("vals", dp)]


class EllSolver(C.Structure):
    """inc/ell.h:23-28"""
    _fields_ = [("max_its", C.c_int), ("its", C.c_int), ("min_err", C.c_double),
                ("err", C.c_double)]


class RveSpec(C.Structure):
    _fields_ = [("nex", C.c_int), ("ney", C.c_int), ("nez", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("nfib", C.c_int),
                ("fib_x", C.c_double * MM_MAX_FIBERS), ("fib_y", C.c_double * MM_MAX_FIBERS),
                ("fib_r", C.c_double * MM_MAX_FIBERS),
                ("young", C.c_double * 2), ("poisson", C.c_double * 2)]


class MicroConfig(C.Structure):
    _fields_ = [("rve", RveSpec), ("solver", C.c_int), ("rtol", C.c_double),
                ("nl_max_its", C.c_int), ("nl_min_norm", C.c_double), ("rho", C.c_double),
                ("device", C.c_int), ("linear_materials", C.c_int), ("fe2_bc", C.c_int)]


class MicroMessage(C.Structure):
    """message_t of inc/comm.h:18-27 (400 bytes)"""
    _fields_ = [("action", C.c_int), ("num_voigt", C.c_int), ("strain_mac", C.c_double * 6),
                ("stress_ave", C.c_double * 6), ("c_tangent_ave", C.c_double * 36),
                ("rho", C.c_double)]


ACTION_NULL, ACTION_MICRO_CALC_STRESS, ACTION_MICRO_CALC_C_TANGENT = 0, 1, 2
ACTION_MICRO_CALC_RHO, ACTION_MICRO_END = 3, 4


class MicroStats(C.Structure):
    _fields_ = [("pcg_its", C.c_int), ("sweeps", C.c_int), ("residual", C.c_double),
                ("seconds", C.c_double), ("assemblies", C.c_int), ("pad_", C.c_int)]


class EngineConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_size_t), ("max_rves", C.c_int),
                ("slab_nodes", C.c_int), ("nrhs", C.c_int), ("precision", C.c_int),
                ("rtol", C.c_double), ("abs_tol", C.c_double), ("inner_rtol", C.c_double),
                ("max_inner_its", C.c_int), ("max_sweeps", C.c_int),
                ("steps_per_launch", C.c_int), ("use_graph", C.c_int),
                ("sync_every_step", C.c_int), ("ell_storage", C.c_int),
                ("initial_guess", C.c_int), ("solver_streams", C.c_int),
                ("profile_step_kernel", C.c_int),
                ("march_planes", C.c_int), ("big_slabs", C.c_int)]


class Request(C.Structure):
    _fields_ = [("id", C.c_uint64), ("rve", RveSpec), ("nload", C.c_int),
                ("strain", (C.c_double * MM_NVOI) * MM_MAX_RHS), ("tangent_mode", C.c_int),
                ("warm_key", C.c_uint64), ("warm_start", C.c_int)]


class Result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("pcg_its", C.c_int),
                ("sweeps", C.c_int), ("residual_rel", C.c_double),
                ("stress", (C.c_double * MM_NVOI) * MM_MAX_RHS),
                ("strain", (C.c_double * MM_NVOI) * MM_MAX_RHS),
                ("c_tangent", C.c_double * (MM_NVOI * MM_NVOI)),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


class EngineInfo(C.Structure):
    _fields_ = [("resident", C.c_int), ("queued", C.c_int), ("completed", C.c_longlong),
                ("steps", C.c_longlong), ("arena_bytes", C.c_size_t),
                ("arena_used", C.c_size_t), ("slab_nodes", C.c_int), ("spmv_ms", C.c_double)]


# name -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    # include/ell.h
    "ell_init": (C.c_int, [C.POINTER(EllMatrix), C.c_int, C.c_int, C.c_int]),
    "ell_set_val": (C.c_int, [C.POINTER(EllMatrix), C.c_int, C.c_int, C.c_double]),
    "ell_add_val": (C.c_int, [C.POINTER(EllMatrix), C.c_int, C.c_int, C.c_double]),
    "ell_add_vals": (C.c_int, [C.POINTER(EllMatrix), ip, C.c_int, ip, C.c_int, dp]),
    "ell_mvp": (C.c_int, [C.POINTER(EllMatrix), dp, dp]),
    "ell_get_val": (C.c_int, [C.POINTER(EllMatrix), C.c_int, C.c_int, dp]),
    "ell_solve_jacobi": (C.c_int, [C.POINTER(EllSolver), C.POINTER(EllMatrix), dp, dp]),
    "ell_set_zero_row": (C.c_int, [C.POINTER(EllMatrix), C.c_int, C.c_double]),
    "ell_set_zero_col": (C.c_int, [C.POINTER(EllMatrix), C.c_int, C.c_double]),
    "ell_set_zero_mat": (C.c_int, [C.POINTER(EllMatrix)]),
    "ell_print_full": (C.c_int, [C.POINTER(EllMatrix)]),
    "ell_free": (C.c_int, [C.POINTER(EllMatrix)]),
    "ell_solve_pcg": (C.c_int, [C.POINTER(EllSolver), C.POINTER(EllMatrix), dp, dp]),
    "ell_device_create": (C.c_int, [C.POINTER(EllMatrix), C.c_int, C.POINTER(C.c_void_p)]),
    "ell_device_update": (C.c_int, [C.c_void_p, C.POINTER(EllMatrix)]),
    "ell_device_download": (C.c_int, [C.c_void_p, ip, dp]),
    "ell_device_mvp": (C.c_int, [C.c_void_p, dp, dp]),
    "ell_device_time_mvp": (C.c_int, [C.c_void_p, C.c_int, dp]),
    "ell_device_solve_pcg": (C.c_int, [C.c_void_p, C.POINTER(EllSolver), dp, dp]),
    "ell_device_solve_jacobi": (C.c_int, [C.c_void_p, C.POINTER(EllSolver), dp, dp]),
    "ell_device_destroy": (C.c_int, [C.c_void_p]),
    "ell_device_set_pcg_fused": (C.c_int, [C.c_void_p, C.c_int]),
    "ell_device_solve_pcg_noalloc": (C.c_int, [C.c_void_p, C.POINTER(EllSolver), C.c_void_p, C.c_void_p]),
    # include/micro_api.h
    "micro_config_default": (C.c_int, [C.POINTER(MicroConfig)]),
    "micro_init": (C.c_int, [C.POINTER(MicroConfig)]),
    "micro_finalize": (C.c_int, []),
    "homog_get_strain_stress": (C.c_int, [dp, dp, dp]),
    "homog_get_strain_stress_non_linear": (C.c_int, [dp, dp, dp]),
    "homog_fe2": (C.c_int, [dp, dp, dp]),
    "homog_get_c_tangent": (C.c_int, [dp, C.POINTER(dp)]),
    "homog_calculate_c_tangent": (C.c_int, [dp, dp]),
    "homog_calculate_c_tangent_around_zero": (C.c_int, [dp]),
    "micro_serve_message": (C.c_int, [C.POINTER(MicroMessage)]),
    "micro_get_stats": (C.c_int, [C.POINTER(MicroStats)]),
    "micro_get_sizes": (C.c_int, [ip, ip, ip]),
    "micro_get_ell": (C.c_int, [ip, dp]),
    "micro_get_elem_type": (C.c_int, [ip]),
    "micro_get_displacement": (C.c_int, [dp]),
    "micro_write_vtu": (C.c_int, [C.c_char_p]),
    "micro_residual": (C.c_int, [dp, dp, dp, dp]),
    "mm_engine_config_default": (C.c_int, [C.POINTER(EngineConfig)]),
    "mm_request_validate": (C.c_int, [C.POINTER(Request), C.c_int]),
    "mm_engine_create": (C.c_int, [C.POINTER(EngineConfig), C.POINTER(C.c_void_p)]),
    "mm_engine_destroy": (C.c_int, [C.c_void_p]),
    "mm_engine_submit": (C.c_int, [C.c_void_p, C.POINTER(Request)]),
    "mm_engine_step": (C.c_int, [C.c_void_p]),
    "mm_engine_poll": (C.c_int, [C.c_void_p, C.POINTER(Result), C.c_int, ip]),
    "mm_engine_drain": (C.c_int, [C.c_void_p]),
    "mm_engine_start": (C.c_int, [C.c_void_p]),
    "mm_engine_stop": (C.c_int, [C.c_void_p]),
    "mm_engine_now": (C.c_double, [C.c_void_p]),
    "mm_engine_info_get": (C.c_int, [C.c_void_p, C.POINTER(EngineInfo)]),
    "mm_engine_step_kernel_times": (C.c_int, [C.c_void_p, dp, C.c_int, ip]),
    "mm_engine_rve_sizes": (C.c_int, [C.c_void_p, C.c_uint64, ip, ip]),
    "mm_engine_download_ell": (C.c_int, [C.c_void_p, C.c_uint64, ip, dp]),
    "mm_engine_download_phase": (C.c_int, [C.c_void_p, C.c_uint64, ip]),
    "mm_engine_download_x": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, dp]),
    "mm_engine_time_spmv": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
    "mm_engine_time_step": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
    "mm_hbm_copy_probe": (C.c_int, [C.c_int, C.c_size_t, C.c_int, dp]),
    # include/macro_api.h
    "mm_macro_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, dp, C.c_int,
                                  C.POINTER(C.c_void_p)]),
    "mm_macro_destroy": (C.c_int, [C.c_void_p]),
    "mm_macro_sizes": (C.c_int, [C.c_void_p, ip, ip]),
    "mm_macro_set_dirichlet": (C.c_int, [C.c_void_p, C.c_int, ip]),
    "mm_macro_set_u": (C.c_int, [C.c_void_p, dp]),
    "mm_macro_strain": (C.c_int, [C.c_void_p, dp]),
    "mm_macro_assemble_b": (C.c_int, [C.c_void_p, dp, dp, dp]),
    "mm_macro_assemble_a": (C.c_int, [C.c_void_p, dp]),
    "mm_macro_get_ell": (C.c_int, [C.c_void_p, ip, dp]),
    "mm_macro_solve": (C.c_int, [C.c_void_p, dp, dp, C.c_double, ip]),
    "mm_hbm_stream_probe": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.c_int, dp]),
}

_LIB = None


def lib():
    """The loaded product library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "madmax_amd: %s is missing; build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` or `make -C madmax_amd/csrc`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with code %d" % (what, rc))
    return rc
