"""madmax_amd -- MI355X-native micro (RVE) homogenization engine, a drop-in
for the sputnik micro solve: the reference's ELL matrix API and homogenization
entry points over HIP kernels for gfx950 (libmadmax_amd.so)."""
from . import _lib
from ._lib import (MICRO_BC_PER_LM, MICRO_BC_USTRAIN, MICRO_SOLVER_PCG, MICRO_SOLVER_PCG_MIXED, MM_ELL_DICT, MM_ELL_FULL, MM_GUESS_AFFINE,
                   MM_GUESS_ZERO, MM_PREC_FP64, MM_PREC_MIXED, MM_TANGENT_FD, MM_TANGENT_NONE,
                   MM_TANGENT_UNIT, lib)
from .engine import (Ell, Engine, Macro, Micro, fd_loads, make_request, rve_spec, spec_fibers,
                     unit_loads)

__all__ = ["Ell", "Engine", "Macro", "Micro", "rve_spec", "spec_fibers", "unit_loads", "fd_loads",
           "make_request", "lib", "MICRO_SOLVER_PCG", "MICRO_SOLVER_PCG_MIXED", "MM_PREC_FP64",
           "MM_PREC_MIXED", "MM_TANGENT_FD", "MM_TANGENT_NONE", "MM_TANGENT_UNIT", "MM_ELL_FULL",
           "MM_ELL_DICT", "MM_GUESS_ZERO", "MM_GUESS_AFFINE", "MICRO_BC_USTRAIN", "MICRO_BC_PER_LM"]
