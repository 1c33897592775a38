"""GPU parity of the micro homogenization entry points (include/micro_api.h,
micro/inc/micro.h:150-157) and of the hex8 assembly kernel against the
oracle: ELL columns and FP64 values bit-exact, phases identical, averaged
stress and tangent within 1e-8 (normwise relative)."""
import numpy as np
import pytest

from madmax_amd import MICRO_SOLVER_PCG, MICRO_SOLVER_PCG_MIXED, Micro, configs, rve_spec
from tests.helpers import normwise_rel, oracle_rve, same_bits

pytestmark = pytest.mark.gpu

TOL = 1e-8  # stress / tangent bar vs the FP64 oracle

SPECS = {
    "1x1x1": rve_spec(1, [(0.5, 0.5, 0.25)]),
    "3x2x1": rve_spec((3, 2, 1), [(0.3, 0.4, 0.3)], lengths=(1.5, 1.0, 0.5)),
    "4^3 two fibers": rve_spec(4, [(0.25, 0.25, 0.2), (0.7, 0.7, 0.3)], (2.0e7, 9.0e7), (0.2, 0.33)),
    "c1": configs.c1_spec(),
}


@pytest.mark.parametrize("name", list(SPECS))
def test_assembly_bitexact(name):
    spec = SPECS[name]
    ref, et = oracle_rve(spec)
    cols_ref, vals_ref = ref.assemble()
    with Micro(spec) as m:
        cols, vals = m.ell()
        assert np.array_equal(m.elem_type(), et)
    assert np.array_equal(cols, cols_ref)
    assert same_bits(vals, vals_ref)


def test_c1_fiber_count():
    with Micro(configs.c1_spec()) as m:
        assert int(m.elem_type().sum()) == 160


@pytest.mark.parametrize("solver", [MICRO_SOLVER_PCG, MICRO_SOLVER_PCG_MIXED])
def test_fe2_stress_c1(solver):
    spec = configs.c1_spec()
    ref, _ = oracle_rve(spec)
    eps = configs.c1_strain()
    out = ref.homog(eps, rtol=1e-12)
    with Micro(spec, solver=solver, nl_min_norm=0.0 if solver == MICRO_SOLVER_PCG_MIXED else 1e-4) as m:
        se, ss = m.fe2(eps)
        st = m.stats()
        x = m.displacement()
    assert normwise_rel(ss, out["stress_ave"]) <= TOL
    assert np.max(np.abs(se - eps)) <= 1e-12  # <eps> = eps_mac under affine boundary values
    assert np.max(np.abs(x - out["x"])) <= 1e-6 * np.max(np.abs(out["x"]))
    assert st["pcg_its"] > 0


@pytest.mark.parametrize("case", [0, 1, 2])
def test_fp64_pcg_iterations_equal_oracle(case):
    """The FP64 path (Jacobi-PCG to 1e-10 of the first residual, Newton to
    nl_min_norm) takes exactly the oracle's PCG iterations on the same
    system: its step, direction and stopping logic are the reference's, one
    step per iteration (a step that updates nothing would add one per
    Newton sweep)."""
    spec, eps = [(configs.c1_spec(), configs.c1_strain()),
                 (rve_spec(6, [(0.5, 0.5, 0.3)]), np.array([1e-3, 2e-4, -1e-4, 0.0, 3e-4, 0.0])),
                 (rve_spec((7, 5, 4), [(0.4, 0.6, 0.25)]), np.array([0.0, 1e-3, 0.0, 5e-4, 0.0, -2e-4]))][case]
    ref, _ = oracle_rve(spec)
    out = ref.homog(eps, rtol=1e-10)
    with Micro(spec, solver=MICRO_SOLVER_PCG) as m:
        m.fe2(eps)
        st = m.stats()
    assert st["pcg_its"] == out["its"]


def test_tangent_linear_c1_and_shortcut():
    spec = configs.c1_spec()
    ref, _ = oracle_rve(spec)
    c_ref = ref.c_tangent_fd(np.zeros(6), 0.005, 1e-12)
    with Micro(spec) as m:
        c = m.c_tangent(np.zeros(6))  # c_tangent_linear, computed in micro_init
        assert normwise_rel(c, c_ref) <= TOL
        eps = np.array([1e-3, -2e-4, 3e-4, 1e-4, 0.0, -5e-5])
        se, ss = m.strain_stress(eps)  # linear shortcut (micro/src/homogenize.c:7-14)
        assert np.array_equal(se, eps)
        assert np.allclose(ss, c @ eps, rtol=0, atol=1e-9 * np.max(np.abs(ss)))
        # homog_calculate_c_tangent after init uses the shortcut like the reference
        c2 = m.calculate_c_tangent(eps)
        assert normwise_rel(c2, c) <= 1e-9


def test_message_dispatch_like_the_micro_loop():
    """micro/src/main.c:116-151 through micro_serve_message with the 400-byte
    message_t layout of inc/comm.h:18-27."""
    import ctypes as C

    from madmax_amd import _lib

    spec = configs.c1_spec()
    with Micro(spec) as m:
        c = m.c_tangent(np.zeros(6))
        msg = _lib.MicroMessage()
        msg.action = _lib.ACTION_MICRO_CALC_STRESS
        msg.num_voigt = 6
        eps = [1e-3, 0, 2e-4, 0, 0, 1e-4]
        for i in range(6):
            msg.strain_mac[i] = eps[i]
        assert _lib.lib().micro_serve_message(C.byref(msg)) == 0
        assert np.allclose(np.array(msg.stress_ave[:]), c @ np.array(eps), rtol=0, atol=1e-9 * 1e5)
        msg.action = _lib.ACTION_MICRO_CALC_C_TANGENT
        assert _lib.lib().micro_serve_message(C.byref(msg)) == 0
        assert np.array_equal(np.array(msg.c_tangent_ave[:]).reshape(6, 6), c)
        msg.action = _lib.ACTION_MICRO_CALC_RHO
        assert _lib.lib().micro_serve_message(C.byref(msg)) == 0 and msg.rho == 1.0e7
        msg.action = 99
        assert _lib.lib().micro_serve_message(C.byref(msg)) == 1
