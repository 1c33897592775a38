"""GPU parity of the batched / serving engine (mm_engine_*): FP32 ELL slabs
equal the rounded oracle ELL, batched mixed-precision tangents of uneven RVEs
match the FP64 oracle within 1e-8, results are the same alone and in a mix,
resident RVEs are reused without re-assembly, the serving thread admits work
while other work is on the device, and oversize / invalid requests fail
cleanly."""
import time

import numpy as np
import pytest

from madmax_amd import (MM_ELL_DICT, MM_ELL_FULL, MM_GUESS_AFFINE, MM_GUESS_ZERO, MM_PREC_FP64,
                        MM_PREC_MIXED, MM_TANGENT_FD, MM_TANGENT_NONE, MM_TANGENT_UNIT, Engine,
                        configs, fd_loads, rve_spec, unit_loads)
from madmax_amd import _lib
from tests.helpers import normwise_rel, oracle_rve

pytestmark = pytest.mark.gpu

TOL = 1e-8
GB = 1 << 30


def _oracle_tangent(spec, rtol=1e-12):
    r, _ = oracle_rve(spec)
    c, _ = r.c_tangent_unit(1e-3, rtol)
    return c


def _solve_batch(eng, specs, key0=0):
    for i, s in enumerate(specs):
        assert eng.submit(i + 1, s, unit_loads(), MM_TANGENT_UNIT,
                          warm_key=(key0 + i + 1) if key0 else 0) == 0
    eng.drain()
    res = {r["id"]: r for r in eng.poll()}
    assert len(res) == len(specs)
    return [res[i + 1] for i in range(len(specs))]


@pytest.mark.parametrize("storage", [MM_ELL_FULL, MM_ELL_DICT])
def test_fp32_slabs_equal_rounded_oracle(storage):
    spec = rve_spec(5, [(0.4, 0.6, 0.3)])
    ref, et = oracle_rve(spec)
    cols_ref, vals_ref = ref.assemble()
    with Engine(nrhs=6, precision=MM_PREC_MIXED, hbm_budget=GB // 4, max_rves=8,
                ell_storage=storage) as eng:
        _solve_batch(eng, [spec], key0=100)
        cols, vals = eng.download_ell(101)
        assert np.array_equal(eng.download_phase(101), et)
    assert np.array_equal(cols, cols_ref)
    assert np.array_equal(vals, vals_ref.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("precision", [MM_PREC_FP64, MM_PREC_MIXED])
def test_batched_uneven_tangents_vs_oracle(precision):
    specs = [configs.c4_spec(4, i) for i in range(6)]
    specs = [s for s in specs if s.nex <= 20][:4] + [rve_spec(7, [(0.45, 0.5, 0.33)])]
    with Engine(nrhs=6, precision=precision, hbm_budget=2 * GB, max_rves=64,
                inner_rtol=1e-10 if precision == MM_PREC_FP64 else 1e-5) as eng:
        out = _solve_batch(eng, specs)
    for s, r in zip(specs, out):
        assert r["status"] == 0
        err = normwise_rel(r["c_tangent"], _oracle_tangent(s))
        assert err <= TOL, (s.nex, err)
        c = r["c_tangent"]
        assert normwise_rel(c, c.T) <= 1e-7  # major symmetry


def test_alone_vs_in_mix_identical():
    target = configs.c2_spec(2, 7, n=12)
    others = [configs.c4_spec(9, i) for i in range(5)]
    with Engine(nrhs=6, hbm_budget=2 * GB, max_rves=64) as eng:
        alone = _solve_batch(eng, [target])[0]
    with Engine(nrhs=6, hbm_budget=2 * GB, max_rves=64) as eng:
        mixed = _solve_batch(eng, others[:2] + [target] + others[2:])[2]
    # per-RVE reductions run over that RVE's slabs only: bitwise identical
    assert np.array_equal(alone["c_tangent"], mixed["c_tangent"])
    assert alone["pcg_its"] == mixed["pcg_its"]


def test_slab_shapes_agree():
    """The same uneven batch with 64, 128 and 256 nodes per slab: every slab
    shape meets the oracle bar, and the shapes agree closely with each other
    (only the grouping of the partial sums differs)."""
    specs = [configs.c4_spec(6, i) for i in range(6)]
    specs = [s for s in specs if s.nex <= 16][:3] + [rve_spec((9, 7, 5), [(0.4, 0.55, 0.3)])]
    refs = [_oracle_tangent(s) for s in specs]
    out = {}
    for sn in (64, 128, 256):
        with Engine(nrhs=6, hbm_budget=GB, max_rves=16, slab_nodes=sn) as eng:
            out[sn] = _solve_batch(eng, specs)
        for s, r, c_ref in zip(specs, out[sn], refs):
            assert r["status"] == 0
            assert normwise_rel(r["c_tangent"], c_ref) <= TOL, (sn, s.nex)
    for a, b, c in zip(out[64], out[128], out[256]):
        assert normwise_rel(a["c_tangent"], c["c_tangent"]) <= 1e-12
        assert normwise_rel(b["c_tangent"], c["c_tangent"]) <= 1e-12


@pytest.mark.parametrize("precision", [MM_PREC_MIXED, MM_PREC_FP64])
def test_plane_march_matches_slab_kernel(precision):
    """Kernel A slab by slab (march_planes=0) and as a z-plane march with two
    run lengths, in FP32 inside FP64 refinement (two load cases per tile) and
    in FP64 (one): every variant meets the oracle bar and they agree closely
    (the SpMV bits are the same; only the grouping of the p.q partial sums
    differs)."""
    specs = [rve_spec(8, [(0.45, 0.5, 0.3)]), rve_spec((12, 9, 7), [(0.4, 0.6, 0.25)]),
             configs.c2_spec(2, 5, n=14)]
    refs = [_oracle_tangent(s) for s in specs]
    out = {}
    for planes in (0, 8, 2):
        print("march planes", planes, "precision", precision, flush=True)
        with Engine(nrhs=6, precision=precision, hbm_budget=GB, max_rves=16, march_planes=planes,
                    inner_rtol=1e-10 if precision == MM_PREC_FP64 else 1e-5) as eng:
            out[planes] = _solve_batch(eng, specs)
        for s, r, c_ref in zip(specs, out[planes], refs):
            assert r["status"] == 0
            assert normwise_rel(r["c_tangent"], c_ref) <= TOL, (planes, s.nex)
    for a, b, c in zip(out[0], out[8], out[2]):
        assert normwise_rel(b["c_tangent"], a["c_tangent"]) <= 1e-12
        assert normwise_rel(c["c_tangent"], a["c_tangent"]) <= 1e-12


@pytest.mark.parametrize("storage", [MM_ELL_FULL, MM_ELL_DICT])
def test_march_and_slab_rves_in_one_batch(storage):
    """With 64 nodes per slab an RVE 34 elements wide cannot march (its window
    exceeds two positions per lane) and goes slab by slab beside RVEs that
    march: both kernels A run in the same step, and every RVE gets the bits it
    gets alone."""
    wide = rve_spec((34, 6, 5), [(0.5, 0.5, 0.3)])
    batch = [rve_spec(9, [(0.45, 0.55, 0.3)]), wide, rve_spec((12, 10, 8), [(0.5, 0.4, 0.28)])]
    with Engine(nrhs=6, hbm_budget=GB, max_rves=16, slab_nodes=64, ell_storage=storage) as eng:
        mixed = _solve_batch(eng, batch)
    for i, s in enumerate(batch):
        with Engine(nrhs=6, hbm_budget=GB, max_rves=16, slab_nodes=64, ell_storage=storage) as eng:
            alone = _solve_batch(eng, [s])[0]
        assert mixed[i]["status"] == 0
        assert np.array_equal(alone["c_tangent"], mixed[i]["c_tangent"]), i
        assert alone["pcg_its"] == mixed[i]["pcg_its"]
    assert normwise_rel(mixed[1]["c_tangent"], _oracle_tangent(wide)) <= TOL


@pytest.mark.parametrize("use_graph", [True, False])
def test_one_and_two_solver_streams_identical(use_graph):
    """Two stream groups (RVEs split in slab order, group 1 forked after group
    0's first kernel) give the same bits as one stream, RVE by RVE."""
    specs = [configs.c4_spec(5, i) for i in range(7)] + [configs.c2_spec(2, 3, n=10)]
    res = {}
    for streams in (1, 2):
        with Engine(nrhs=6, hbm_budget=2 * GB, max_rves=16, solver_streams=streams,
                    use_graph=use_graph) as eng:
            res[streams] = _solve_batch(eng, specs)
    for a, b in zip(res[1], res[2]):
        assert a["status"] == 0 and b["status"] == 0
        assert np.array_equal(a["c_tangent"], b["c_tangent"])
        assert a["pcg_its"] == b["pcg_its"]


@pytest.mark.parametrize("precision", [MM_PREC_MIXED, MM_PREC_FP64])
def test_per_rve_totals_same_bits(precision):
    """RVEs of at least big_slabs slabs take their per-iteration dot-product
    totals from kernels R and S (summed once per RVE) instead of summing the
    slab partials in every workgroup; both sum in the same order, so putting
    every RVE on that path (big_slabs=1), or only the larger ones (a
    threshold inside the batch's range of slab counts, so that one launch
    holds RVEs of both kinds), changes no bit.  Uneven RVEs (the march and
    the slab kernel A), two stream groups, graphs."""
    specs = [configs.c4_spec(5, i) for i in range(5)] + [configs.c2_spec(2, 3, n=20)]
    nslab = sorted(-(-(s.nex + 1) * (s.ney + 1) * (s.nez + 1) // 256) for s in specs)
    mid = nslab[len(nslab) // 2]
    assert nslab[0] < mid <= nslab[-1]
    res = {}
    for big in (None, 1, mid):
        print("per-RVE totals: precision", precision, "big_slabs", big, flush=True)
        with Engine(nrhs=6, precision=precision, hbm_budget=2 * GB, max_rves=16, big_slabs=big) as eng:
            res[big] = _solve_batch(eng, specs)
    for big in (1, mid):
        for a, b in zip(res[None], res[big]):
            assert a["status"] == 0 and b["status"] == 0
            assert np.array_equal(a["c_tangent"], b["c_tangent"])
            assert a["pcg_its"] == b["pcg_its"]


def test_fd_tangent_seven_loads():
    spec = rve_spec(6, [(0.5, 0.5, 0.3)])
    eps = np.array([1e-3, 2e-4, -1e-4, 0.0, 3e-4, 0.0])
    ref, _ = oracle_rve(spec)
    c_ref = ref.c_tangent_fd(eps, 0.005, 1e-12)
    with Engine(nrhs=7, precision=MM_PREC_MIXED, hbm_budget=GB // 2, max_rves=8) as eng:
        assert eng.submit(1, spec, fd_loads(eps), MM_TANGENT_FD) == 0
        eng.drain()
        r = eng.poll()[0]
    assert r["status"] == 0
    assert normwise_rel(r["c_tangent"], c_ref) <= TOL
    out = ref.homog(eps, rtol=1e-12)
    assert normwise_rel(r["stress"][0], out["stress_ave"]) <= TOL


def test_resident_reuse_and_warm_start():
    spec = rve_spec(10, [(0.5, 0.5, 0.3)])
    with Engine(nrhs=6, hbm_budget=GB, max_rves=8) as eng:
        first = _solve_batch(eng, [spec], key0=10)[0]
        x1 = eng.download_x(11, 0)
        # same key, cold start: no assembly, same answer
        assert eng.submit(2, spec, unit_loads(), MM_TANGENT_UNIT, warm_key=11, warm_start=0) == 0
        eng.drain()
        cold = eng.poll()[0]
        # same key, warm start from the kept displacements: converged at once
        assert eng.submit(3, spec, unit_loads(), MM_TANGENT_UNIT, warm_key=11, warm_start=1) == 0
        eng.drain()
        warm = eng.poll()[0]
        x3 = eng.download_x(11, 0)
    assert np.array_equal(cold["c_tangent"], first["c_tangent"])
    assert warm["pcg_its"] < cold["pcg_its"]
    assert normwise_rel(warm["c_tangent"], first["c_tangent"]) <= 1e-9
    assert np.max(np.abs(x3 - x1)) <= 1e-9 * np.max(np.abs(x1))


def test_serving_thread_admits_while_busy():
    specs = [configs.c2_spec(5, i, n=8 + 2 * (i % 4)) for i in range(12)]
    with Engine(nrhs=6, hbm_budget=GB, max_rves=6) as eng:
        ref = {i + 1: r for i, r in enumerate(_solve_batch(eng, specs))}
    with Engine(nrhs=6, hbm_budget=GB, max_rves=6, steps_per_launch=8) as eng:
        eng.start()
        for i, s in enumerate(specs):
            assert eng.submit(i + 1, s, unit_loads(), MM_TANGENT_UNIT) == 0
            time.sleep(0.002)
        got = {}
        t0 = time.time()
        while len(got) < len(specs) and time.time() - t0 < 60:
            for r in eng.poll():
                got[r["id"]] = r
            time.sleep(0.001)
        assert eng.stop() == 0
    assert len(got) == len(specs)
    overlap = any(got[i]["t_admit"] < max(got[j]["t_done"] for j in got if j != i)
                  and got[i]["t_done"] > min(got[j]["t_admit"] for j in got if j != i) for i in got)
    assert overlap
    for i in got:
        assert normwise_rel(got[i]["c_tangent"], ref[i]["c_tangent"]) <= 1e-12


def test_admission_limits():
    with Engine(nrhs=6, hbm_budget=64 << 20, max_rves=4) as eng:
        big = rve_spec(64)  # ~0.8 M rows of vectors alone exceed 64 MB
        assert eng.submit(1, big, unit_loads(), MM_TANGENT_UNIT) == _lib.MM_REJECTED
        bad = rve_spec(4, young=(float("nan"), 1e8))
        assert eng.submit(2, bad, unit_loads(), MM_TANGENT_UNIT) == _lib.MM_ERR_ARG
        loads = unit_loads()
        loads[2] = np.array([0, 0, float("inf"), 0, 0, 0])
        assert eng.submit(3, rve_spec(4), loads, MM_TANGENT_UNIT) == _lib.MM_ERR_ARG
        # more RVEs than slots: admitted as slots free up
        specs = [rve_spec(6, [(0.5, 0.5, 0.1 + 0.02 * i)]) for i in range(10)]
        out = _solve_batch(eng, specs)
        assert all(r["status"] == 0 for r in out)


def test_vector_run_limit_rejected():
    """An RVE whose solver-vector run (nrhs x 3 x nodes x 8 bytes, 7 FP64
    loads) reaches 2^31 bytes is refused even though the arena could hold its
    slabs: the SpMV addresses a run with 32-bit offsets. The 12 GB arena holds
    about 57,900 slabs (222 KB each at 7 FP64 loads); 235^3 elements need
    51,345 slabs and 2.21 GB per run (rejected), 200^3 need 31,722 slabs and
    1.36 GB per run (queued; nothing is solved here)."""
    with Engine(nrhs=7, precision=MM_PREC_FP64, hbm_budget=12 * GB, max_rves=2,
                slab_nodes=256) as eng:
        loads = [np.eye(6)[i % 6] * 1e-3 for i in range(7)]
        assert eng.submit(1, rve_spec(235), loads, MM_TANGENT_NONE) == _lib.MM_REJECTED
        assert eng.submit(2, rve_spec(200), loads, MM_TANGENT_NONE) == 0


def test_not_converged_status():
    with Engine(nrhs=6, hbm_budget=GB // 4, max_rves=4, max_inner_its=2, max_sweeps=1) as eng:
        assert eng.submit(1, rve_spec(8, [(0.5, 0.5, 0.3)]), unit_loads(), MM_TANGENT_UNIT) == 0
        eng.drain()
        r = eng.poll()[0]
    assert r["status"] == _lib.MM_NOT_CONVERGED


def test_c2_tangent_30cube_vs_oracle():
    spec = configs.c2_spec(2, 0)
    c_ref = _oracle_tangent(spec, rtol=1e-11)
    with Engine(nrhs=6, precision=MM_PREC_MIXED, hbm_budget=4 * GB, max_rves=8) as eng:
        r = _solve_batch(eng, [spec])[0]
    assert r["status"] == 0
    assert normwise_rel(r["c_tangent"], c_ref) <= TOL


def test_refinement_sweep_targets():
    """Looser FP32 sweep targets need more FP64 refinement sweeps but reach the
    same FP64 tolerance; the last sweep's target is capped by what the outer
    tolerance still needs, so a tight target does not over-solve."""
    specs = [configs.c2_spec(2, i, n=16) for i in range(3)]
    out = {}
    for tol in (1e-2, 1e-5):
        with Engine(nrhs=6, hbm_budget=GB, max_rves=8, inner_rtol=tol, rtol=1e-10,
                    max_sweeps=16) as eng:
            out[tol] = _solve_batch(eng, specs)
    for s, a, b in zip(specs, out[1e-2], out[1e-5]):
        c_ref = _oracle_tangent(s)
        assert a["status"] == 0 and b["status"] == 0
        assert normwise_rel(a["c_tangent"], c_ref) <= TOL
        assert normwise_rel(b["c_tangent"], c_ref) <= TOL
        assert a["sweeps"] > b["sweeps"]
        assert a["residual_rel"] <= 1e-10 and b["residual_rel"] <= 1e-10


@pytest.mark.parametrize("precision", [MM_PREC_FP64, MM_PREC_MIXED])
def test_dictionary_storage_equals_full_ell(precision):
    """The dictionary rows are the assembled rows bit for bit and the product
    keeps the slot order, so the whole solve is the same to the last bit."""
    specs = [configs.c4_spec(11, i) for i in range(8)]
    specs = [s for s in specs if s.nex <= 20][:3] + [rve_spec((5, 4, 6), [(0.3, 0.6, 0.25)])]
    out = {}
    for storage in (MM_ELL_FULL, MM_ELL_DICT):
        with Engine(nrhs=6, precision=precision, hbm_budget=GB, max_rves=16, ell_storage=storage,
                    inner_rtol=1e-10 if precision == MM_PREC_FP64 else 1e-5) as eng:
            out[storage] = _solve_batch(eng, specs, key0=40)
            cols, vals = eng.download_ell(40 + len(specs))
            out[storage, "ell"] = (cols, vals)
    for a, b in zip(out[MM_ELL_FULL], out[MM_ELL_DICT]):
        assert a["status"] == 0 and b["status"] == 0
        assert np.array_equal(a["c_tangent"], b["c_tangent"])
        assert a["pcg_its"] == b["pcg_its"] and a["sweeps"] == b["sweeps"]
    assert np.array_equal(out[MM_ELL_FULL, "ell"][0], out[MM_ELL_DICT, "ell"][0])
    assert np.array_equal(out[MM_ELL_FULL, "ell"][1], out[MM_ELL_DICT, "ell"][1])
    # and both equal the oracle's assembly of that RVE directly: FP64 bit for
    # bit, FP32 as its rounding
    ref, _ = oracle_rve(specs[-1])
    cols_ref, vals_ref = ref.assemble()
    if precision == MM_PREC_MIXED:
        vals_ref = vals_ref.astype(np.float32).astype(np.float64)
    for storage in (MM_ELL_FULL, MM_ELL_DICT):
        cols, vals = out[storage, "ell"]
        assert np.array_equal(cols, cols_ref)
        assert np.array_equal(vals.view(np.uint64), vals_ref.view(np.uint64))


@pytest.mark.parametrize("storage", [MM_ELL_FULL, MM_ELL_DICT])
def test_spmv_timing_runs(storage):
    with Engine(nrhs=6, hbm_budget=GB, max_rves=8, ell_storage=storage) as eng:
        _solve_batch(eng, [configs.c2_spec(3, i, n=12) for i in range(2)], key0=60)
        ms, nbytes = eng.time_spmv(reps=3)
        ms2, nbytes2 = eng.time_step(reps=3)
    assert ms > 0 and nbytes > 0
    assert ms2 > 0 and nbytes2 > nbytes


def test_autotuner_picks_a_converging_configuration():
    """The autotuner tries every slab shape, the march run lengths and the
    sweep tolerances on a small sample; it returns candidates it measured,
    all of which converged, and the chosen engine reproduces the tangents."""
    from madmax_amd import autotune
    specs = [configs.c2_spec(4, i, n=10) for i in range(3)]
    out = autotune.tune(specs, budget=GB, use_cache=False)
    assert out["slab_nodes"] in autotune.SLAB_NODES
    assert out["march_planes"] in autotune.MARCH_PLANES
    assert out["inner_rtol"] in autotune.INNER_RTOLS
    tried = {(t["slab_nodes"], t["march_planes"], t["inner_rtol"]) for t in out["trials"]}
    assert (out["slab_nodes"], out["march_planes"], out["inner_rtol"]) in tried
    assert {t["slab_nodes"] for t in out["trials"]} == set(autotune.SLAB_NODES)
    assert all(t["ok"] and t["step_ms"] > 0 for t in out["trials"])
    with Engine(nrhs=6, hbm_budget=GB, max_rves=8, slab_nodes=out["slab_nodes"],
                march_planes=out["march_planes"], inner_rtol=out["inner_rtol"], rtol=1e-9) as eng:
        res = _solve_batch(eng, specs)
    for s, r in zip(specs, res):
        assert r["status"] == 0
        assert normwise_rel(r["c_tangent"], _oracle_tangent(s)) <= TOL


@pytest.mark.parametrize("precision", [MM_PREC_FP64, MM_PREC_MIXED])
def test_single_load_sequential_reuse(precision):
    """One load case at a time on a resident RVE (nrhs 1, host check after
    every step): the plain single-RVE path."""
    spec = rve_spec(6, [(0.5, 0.5, 0.3)])
    ref, _ = oracle_rve(spec)
    with Engine(nrhs=1, precision=precision, hbm_budget=GB // 4, max_rves=2, slab_nodes=128,
                inner_rtol=1e-10 if precision == MM_PREC_FP64 else 1e-5, rtol=1e-11,
                max_sweeps=4 if precision == MM_PREC_FP64 else 8, sync_every_step=True,
                use_graph=False) as eng:
        for i, load in enumerate(unit_loads()):
            assert eng.submit(i + 1, spec, [load], MM_TANGENT_NONE, warm_key=1) == 0
            eng.drain()
            r = eng.poll()[0]
            out = ref.homog(load, rtol=1e-12)
            print("load", i, r["status"], r["pcg_its"], r["stress"][0], out["stress_ave"], flush=True)
            assert r["status"] == 0
            assert normwise_rel(r["stress"][0], out["stress_ave"]) <= TOL


@pytest.mark.parametrize("precision", [MM_PREC_FP64, MM_PREC_MIXED])
def test_affine_initial_guess(precision):
    """Starting the interior at the affine field reaches the same tangent (the
    tolerance stays relative to the zero-start right-hand side) in fewer PCG
    iterations than the reference's zero start."""
    specs = [configs.c2_spec(2, i, n=14) for i in range(3)]
    out = {}
    for guess in (MM_GUESS_ZERO, MM_GUESS_AFFINE):
        with Engine(nrhs=6, precision=precision, hbm_budget=GB, max_rves=8, initial_guess=guess,
                    inner_rtol=1e-10 if precision == MM_PREC_FP64 else 1e-5) as eng:
            out[guess] = _solve_batch(eng, specs)
    for s, a, b in zip(specs, out[MM_GUESS_ZERO], out[MM_GUESS_AFFINE]):
        c_ref = _oracle_tangent(s)
        assert a["status"] == 0 and b["status"] == 0
        assert normwise_rel(a["c_tangent"], c_ref) <= TOL
        assert normwise_rel(b["c_tangent"], c_ref) <= TOL
        assert b["pcg_its"] < a["pcg_its"]


def test_large_rve_mixed_vs_fp64():
    """One 56^3 RVE (0.55 M dofs, 4 fibers): the mixed-precision stress equals
    the FP64 solve within the bar, with either matrix storage (bit-identical
    to each other)."""
    spec = configs.c3_spec(n=56)
    eps = configs.c3_strain()
    out = {}
    for name, prec, storage in (("mixed_dict", MM_PREC_MIXED, MM_ELL_DICT),
                                ("mixed_full", MM_PREC_MIXED, MM_ELL_FULL),
                                ("fp64", MM_PREC_FP64, MM_ELL_DICT)):
        with Engine(nrhs=1, precision=prec, hbm_budget=2 * GB, max_rves=2, slab_nodes=256, rtol=1e-11,
                    inner_rtol=1e-5 if prec == MM_PREC_MIXED else 1e-11, max_inner_its=100000,
                    max_sweeps=12 if prec == MM_PREC_MIXED else 2, ell_storage=storage) as eng:
            assert eng.submit(1, spec, [eps], MM_TANGENT_NONE) == 0
            eng.drain()
            out[name] = eng.poll()[0]
        assert out[name]["status"] == 0
    assert np.array_equal(out["mixed_dict"]["stress"][0], out["mixed_full"]["stress"][0])
    assert normwise_rel(out["mixed_dict"]["stress"][0], out["fp64"]["stress"][0]) <= TOL


def test_lru_eviction_keeps_results_exact():
    """More warm-keyed RVEs than fit: the least recently used resident RVEs
    are evicted to admit new ones, and a key that comes back after eviction
    is assembled again and gives the same tangent."""
    specs = [rve_spec(10, [(0.5, 0.5, 0.2 + 0.01 * i)]) for i in range(6)]
    with Engine(nrhs=6, hbm_budget=48 << 20, max_rves=3) as eng:
        first = {}
        for i, s in enumerate(specs):
            assert eng.submit(i + 1, s, unit_loads(), MM_TANGENT_UNIT, warm_key=100 + i) == 0
            eng.drain()
            first[i] = eng.poll()[0]
        info = eng.info()
        assert info["resident"] <= 3
        # key 100 was evicted; resubmitting it assembles it again
        assert eng.submit(50, specs[0], unit_loads(), MM_TANGENT_UNIT, warm_key=100) == 0
        eng.drain()
        again = eng.poll()[0]
    assert all(r["status"] == 0 for r in first.values())
    assert np.array_equal(again["c_tangent"], first[0]["c_tangent"])


def test_warm_start_superposes_kept_solutions():
    """A warm start on a resident linear RVE combines the kept displacements:
    when six kept load cases span the strain space (a unit-load tangent
    block), the next strain starts from the superposed solution and needs
    far fewer PCG iterations than a cold start; with a single kept load it
    starts from the kept solution plus the affine field of the strain
    increment.  Every answer still meets the oracle bar."""
    spec = configs.c2_spec(2, 5, n=12)
    ref, _ = oracle_rve(spec)
    s1 = np.array([4e-4, -2e-4, 1e-4, 3e-4, -1e-4, 2e-4])
    s2 = 1.05 * s1 + np.array([0, 1e-5, 0, 0, 2e-5, 0])
    eye = unit_loads()
    with Engine(nrhs=7, hbm_budget=GB, max_rves=4) as eng:
        assert eng.submit(1, spec, eye + [s1], MM_TANGENT_UNIT, warm_key=9) == 0
        eng.drain()
        first = eng.poll()[0]
        out = {}
        for rid, warm in ((2, 1), (3, 0)):
            assert eng.submit(rid, spec, eye + [s2], MM_TANGENT_UNIT, warm_key=9, warm_start=warm) == 0
            eng.drain()
            out[warm] = eng.poll()[0]
            if warm:  # restore the kept solutions of s1 for the cold comparison
                assert eng.submit(10 + rid, spec, eye + [s1], MM_TANGENT_UNIT, warm_key=9) == 0
                eng.drain()
                eng.poll()
    c_ref, _ = ref.c_tangent_unit(1e-3, 1e-12)
    st_ref = ref.homog(s2, rtol=1e-12)["stress_ave"]
    for r in (first, out[0], out[1]):
        assert r["status"] == 0
        assert normwise_rel(r["c_tangent"], c_ref) <= TOL
    for r in (out[0], out[1]):
        assert normwise_rel(r["stress"][6], st_ref) <= TOL
    print("superposed warm its", out[1]["pcg_its"], "cold its", out[0]["pcg_its"], flush=True)
    assert out[1]["pcg_its"] < out[0]["pcg_its"] // 2

    # one kept load: start from x_old + E(s2 - s1) X
    with Engine(nrhs=1, hbm_budget=GB, max_rves=4) as eng:
        assert eng.submit(1, spec, [s1], MM_TANGENT_NONE, warm_key=5) == 0
        eng.drain()
        eng.poll()
        res = {}
        for rid, warm in ((2, 1), (3, 0)):
            assert eng.submit(rid, spec, [s2], MM_TANGENT_NONE, warm_key=5, warm_start=warm) == 0
            eng.drain()
            res[warm] = eng.poll()[0]
            if warm:
                assert eng.submit(10 + rid, spec, [s1], MM_TANGENT_NONE, warm_key=5) == 0
                eng.drain()
                eng.poll()
    for r in res.values():
        assert r["status"] == 0
        assert normwise_rel(r["stress"][0], st_ref) <= TOL
    print("increment warm its", res[1]["pcg_its"], "cold its", res[0]["pcg_its"], flush=True)
    assert res[1]["pcg_its"] < res[0]["pcg_its"]


def test_dictionary_gather_beyond_2_30_bytes():
    """The dictionary SpMV reads z through a buffer descriptor with 32-bit
    byte offsets.  A 190^3 RVE at 7 FP64 load cases has a 1.17 GB vector run
    (above 2^30 bytes): a few PCG iterations with dictionary storage give the
    same bits as full storage, whose gather uses 64-bit addresses."""
    spec = rve_spec(190, [(0.5, 0.5, 0.3)])
    loads = [1e-3 * np.eye(6)[i % 6] for i in range(7)]
    nodes = 191 ** 3
    run = -(-nodes // 256) * 7 * 3 * 256 * 8
    assert run > (1 << 30)
    out = {}
    for storage in (MM_ELL_DICT, MM_ELL_FULL):
        budget = int(3 * nodes * (81 * 8 + 40 + 7 * 40 + 8) * 1.15) + 2 * GB
        print("190^3 FP64 7 loads, storage", storage, flush=True)
        with Engine(nrhs=7, precision=MM_PREC_FP64, hbm_budget=budget, max_rves=2, slab_nodes=256,
                    max_inner_its=3, max_sweeps=1, inner_rtol=1e-14, rtol=1e-14, ell_storage=storage) as eng:
            assert eng.submit(1, spec, loads, MM_TANGENT_NONE, warm_key=1) == 0
            eng.drain()
            r = eng.poll()[0]
            x = eng.download_x(1, 6)
        out[storage] = (r, x)
    (rd, xd), (rf, xf) = out[MM_ELL_DICT], out[MM_ELL_FULL]
    assert rd["pcg_its"] == rf["pcg_its"] == 3
    assert np.array_equal(xd, xf)
    assert np.array_equal(rd["stress"], rf["stress"])
    assert np.any(xd != 0.0)
