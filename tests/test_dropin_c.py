"""Compiled-C drop-in check: tests/dropin/dropin.c is a C program written only
against include/ell.h and include/micro_api.h that follows the reference
micro's call order (ell_init, ell_set_zero_mat, ell_add_vals per element,
ell_set_zero_row / ell_set_zero_col on the boundary, then homog_fe2's loop
with ell_solve_jacobi or ell_solve_pcg, then micro_init and the message loop
through micro_serve_message; micro/src/alloc.c:54, micro/src/assembly.c:84-137,
micro/src/homogenize.c:175-333, micro/src/main.c:116-151).  It is built with
gcc and linked to libmadmax_amd.so; its results must equal the FP64 oracle's:
the assembled ELL bit for bit, displacements, stresses and tangent within
1e-8."""
import os
import subprocess

import numpy as np
import pytest

from madmax_amd import rve_spec
from tests.helpers import normwise_rel, oracle_rve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DIR = os.path.join(ROOT, "tests", "dropin")
EXE = os.path.join(DIR, "dropin")


def fnv(buf):
    h = 1469598103934665603
    for b in buf:
        h = ((h ^ b) * 1099511628211) & ((1 << 64) - 1)
    return h


def build():
    subprocess.run(["make", "-s", "-C", DIR], check=True, capture_output=True)
    return EXE


def test_dropin_builds_and_links_against_the_library():
    exe = build()
    out = subprocess.run(["ldd", exe], check=True, capture_output=True, text=True).stdout
    line = next(ln for ln in out.splitlines() if "libmadmax_amd.so" in ln)
    assert os.path.realpath(line.split("=>")[1].split("(")[0].strip()) == \
        os.path.realpath(os.path.join(ROOT, "madmax_amd", "libmadmax_amd.so"))


@pytest.mark.gpu
def test_dropin_matches_oracle(tmp_path):
    """The RVE is solved with ell_solve_pcg: the reference's stationary Jacobi
    diverges on 3D hex8 elasticity (the oracle's restatement of src/ell.c:
    143-184 reaches NaN on this system too), so Jacobi is checked on the
    reference's own 1D test system against the fixture of its trajectory."""
    solver = "pcg"
    n, rad = 5, 0.3
    eps = np.array([1e-3, -2e-4, 3e-4, 1e-4, 0.0, -5e-5])
    exe = build()
    xf = str(tmp_path / "x.bin")
    env = dict(os.environ, DROPIN_X=xf)
    print("dropin", solver, n, rad, flush=True)
    p = subprocess.run([exe, str(n), repr(rad), solver] + [repr(float(v)) for v in eps], env=env,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, (p.returncode, p.stdout[-2000:], p.stderr[-2000:])
    lines = {ln.split()[0]: ln.split()[1:] for ln in p.stdout.splitlines() if ln and not ln.startswith("\x1b")}
    assert "ok" in lines
    g = np.load(os.path.join(ROOT, "tests", "golden", "ref_ell_1d.npz"))
    assert abs(int(lines["unity"][0]) - int(g["its"])) <= 1
    assert float(lines["unity"][1]) < 1e-10
    assert np.max(np.abs(np.array([float(v) for v in lines["unityx"]]) - g["x"])) <= 1e-9
    spec = rve_spec(n, [(0.5, 0.5, rad)])
    ref, _ = oracle_rve(spec)
    cols, vals = ref.assemble()
    assert int(lines["ell"][0], 16) == fnv(np.ascontiguousarray(cols, dtype=np.int32).tobytes())
    assert int(lines["ell"][1], 16) == fnv(np.ascontiguousarray(vals, dtype=np.float64).tobytes())
    out = ref.homog(eps, rtol=1e-12)
    x = np.fromfile(xf, dtype=np.float64)
    assert normwise_rel(x, out["x"]) <= 1e-8
    _, s_x = ref.averages(x)
    assert normwise_rel(s_x, out["stress_ave"]) <= 1e-8
    c_ref = ref.c_tangent_fd(np.zeros(6), 0.005, 1e-12)
    stress = np.array([float(v) for v in lines["stress"]])
    tangent = np.array([float(v) for v in lines["tangent"]]).reshape(6, 6)
    fe2 = np.array([float(v) for v in lines["fe2"]])
    assert normwise_rel(tangent, c_ref) <= 1e-8
    assert normwise_rel(stress, c_ref @ eps) <= 1e-8  # the linear shortcut sigma = C_lin eps
    assert normwise_rel(fe2, out["stress_ave"]) <= 1e-8
