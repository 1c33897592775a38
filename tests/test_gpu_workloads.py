"""The three large benchmark workloads at their full sizes on the GPU, checked
against FP64-oracle fixtures (tests/golden/wl_*.npz, written by
tests/golden/make_workload_fixtures.py) and against properties that hold at
any size:
  C3  one 100^3 RVE (3.09 M dofs, 4 fibers, vf 0.4), random strain: mixed and
      FP64 stresses vs the oracle, <eps> = eps_mac, full and dictionary
      storage give the same bits
  C4  the 512-RVE uneven batch (8^3 .. 32^3): every RVE converges, tangents
      are major-symmetric and lie between the Reuss and Voigt bounds, and
      one RVE of each size equals the oracle's tangent
  C5  the 32,768-request FE^2 stream (12^3 / 20^3 / 30^3 RVEs, strain + 6 unit
      loads per request, Poisson arrivals): every request converges, a
      64-request sample equals the oracle, a subset solved alone gives the
      same bits as in the stream, and a warm-started second Newton pass over
      the sample needs fewer PCG iterations and still meets the bar."""
import os
import time

import numpy as np
import pytest

from madmax_amd import (MM_ELL_DICT, MM_ELL_FULL, MM_PREC_FP64, MM_PREC_MIXED, MM_TANGENT_NONE,
                        MM_TANGENT_UNIT, Engine, configs, spec_fibers, unit_loads)
from oracle import pyoracle as po
from tests.helpers import normwise_rel

pytestmark = pytest.mark.gpu

TOL = 1e-8
GB = 1 << 30
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _spec_row(s):
    return np.array([s.nex, s.nfib] + [s.fib_x[i] for i in range(s.nfib)] + [s.fib_y[i] for i in range(s.nfib)]
                    + [s.fib_r[i] for i in range(s.nfib)] + [s.young[0], s.young[1], s.poisson[0], s.poisson[1]])


def test_c3_100cube_vs_oracle():
    g = np.load(os.path.join(GOLD, "wl_c3.npz"))
    spec, eps = configs.c3_spec(), configs.c3_strain()
    assert np.array_equal(_spec_row(spec), g["spec"]) and np.array_equal(eps, g["strain"])
    rows = 3 * 101 ** 3
    out = {}
    for name, prec, storage in (("mixed_dict", MM_PREC_MIXED, MM_ELL_DICT),
                                ("mixed_full", MM_PREC_MIXED, MM_ELL_FULL),
                                ("fp64_dict", MM_PREC_FP64, MM_ELL_DICT)):
        print("c3", name, flush=True)
        budget = int(rows * (81 * 8 + 200) * 1.3) + (512 << 20)
        with Engine(nrhs=1, precision=prec, hbm_budget=budget, max_rves=2, slab_nodes=256, rtol=1e-10,
                    inner_rtol=1e-5 if prec == MM_PREC_MIXED else 1e-10,
                    max_sweeps=12 if prec == MM_PREC_MIXED else 2, max_inner_its=200000,
                    ell_storage=storage) as eng:
            assert eng.submit(1, spec, [eps], MM_TANGENT_NONE) == 0
            eng.drain()
            out[name] = eng.poll()[0]
        r = out[name]
        print("c3", name, "its", r["pcg_its"], "sweeps", r["sweeps"], flush=True)
        assert r["status"] == 0
        assert normwise_rel(r["stress"][0], g["stress_ave"]) <= TOL, name
        assert normwise_rel(r["strain"][0], eps) <= 1e-12
    assert np.array_equal(out["mixed_dict"]["stress"][0], out["mixed_full"]["stress"][0])
    assert out["mixed_dict"]["pcg_its"] == out["mixed_full"]["pcg_its"]


def _bounds(spec):
    """Voigt / Reuss mixture bounds of an RVE from its phase fractions."""
    m = po.mesh(3, spec.nex, spec.ney, spec.nez)
    et = po.phase_map(m, spec_fibers(spec))
    vf = float(np.mean(et))
    c0 = po.material_c(3, spec.young[0], spec.poisson[0])
    c1 = po.material_c(3, spec.young[1], spec.poisson[1])
    voigt = (1 - vf) * c0 + vf * c1
    reuss = np.linalg.inv((1 - vf) * np.linalg.inv(c0) + vf * np.linalg.inv(c1))
    return voigt, reuss


def test_c4_512_uneven_batch():
    g = np.load(os.path.join(GOLD, "wl_c4.npz"))
    specs = configs.c4_specs(count=512)
    for i, row in zip(g["index"], g["spec"]):
        assert np.array_equal(_spec_row(specs[i]), row)
    rows = sum(3 * (s.nex + 1) ** 3 for s in specs)
    budget = int(rows * 250 * 1.3) + GB
    print("c4: 512 RVEs,", rows, "dofs", flush=True)
    t = time.perf_counter()
    with Engine(nrhs=6, precision=MM_PREC_MIXED, hbm_budget=budget, max_rves=520, slab_nodes=256,
                inner_rtol=1e-5, rtol=1e-9) as eng:
        for i, s in enumerate(specs):
            assert eng.submit(i + 1, s, unit_loads(), MM_TANGENT_UNIT) == 0
        eng.drain()
        res = {r["id"]: r for r in eng.poll(1024)}
    print("c4: %.2f s" % (time.perf_counter() - t), flush=True)
    assert len(res) == 512
    for i, s in enumerate(specs):
        r = res[i + 1]
        assert r["status"] == 0, i
        c = r["c_tangent"]
        assert normwise_rel(c, c.T) <= 1e-7, i
        voigt, reuss = _bounds(s)
        d = np.diag(c)
        assert np.all(d <= np.diag(voigt) * (1 + 1e-9)) and np.all(d >= np.diag(reuss) * (1 - 1e-9)), i
        # Voigt - C and C - Reuss are positive semi-definite (energy bounds)
        assert np.min(np.linalg.eigvalsh(0.5 * ((voigt - c) + (voigt - c).T))) >= -1e-7 * np.max(d), i
        assert np.min(np.linalg.eigvalsh(0.5 * ((c - reuss) + (c - reuss).T))) >= -1e-7 * np.max(d), i
    for i, c_ref in zip(g["index"], g["c_tangent"]):
        assert normwise_rel(res[i + 1]["c_tangent"], c_ref) <= TOL, (i, specs[i].nex)


def _c5_engine():
    return Engine(nrhs=7, precision=MM_PREC_MIXED, hbm_budget=48 * GB, max_rves=8192, slab_nodes=256,
                  inner_rtol=1e-5, rtol=1e-9, steps_per_launch=8)


def test_c5_full_stream():
    g = np.load(os.path.join(GOLD, "wl_c5.npz"))
    # below the engine's capacity for this mix (about 1,500 requests/s), so the
    # backlog stays short and the 64 kept RVEs are not evicted to make room
    rate = 1000.0
    stream = configs.c5_stream(rate=rate)
    assert len(stream) == 32768
    sample = [int(i) for i in g["index"]]
    for j, i in enumerate(sample):
        assert np.array_equal(_spec_row(stream[i]["spec"]), g["spec"][j])
        assert np.array_equal(stream[i]["strain"], g["strain"][j])
    keyed = set(sample)
    eye = unit_loads()
    got = {}
    print("c5: 32768 requests at %.0f/s" % rate, flush=True)
    with _c5_engine() as eng:
        eng.start()
        t0 = time.perf_counter()
        for i, rq in enumerate(stream):
            while time.perf_counter() - t0 < rq["arrival"]:
                for r in eng.poll():
                    got[r["id"]] = r
                time.sleep(0.0002)
            assert eng.submit(i + 1, rq["spec"], eye + [rq["strain"]], MM_TANGENT_UNIT,
                              warm_key=rq["key"] if i in keyed else 0) == 0
        while len(got) < len(stream) and time.perf_counter() - t0 < 240:
            for r in eng.poll():
                got[r["id"]] = r
            time.sleep(0.001)
        dt = time.perf_counter() - t0
        print("c5: %d results in %.1f s" % (len(got), dt), flush=True)
        assert len(got) == len(stream)
        # second Newton pass over the sample, warm-started from the resident RVEs
        warm = {}
        for j, i in enumerate(sample):
            rq = stream[i]
            assert eng.submit(100000 + i, rq["spec"], eye + [1.05 * rq["strain"]], MM_TANGENT_UNIT,
                              warm_key=rq["key"], warm_start=1) == 0
        t1 = time.perf_counter()
        while len(warm) < len(sample) and time.perf_counter() - t1 < 60:
            for r in eng.poll():
                warm[r["id"] - 100000] = r
            time.sleep(0.001)
        assert eng.stop() == 0
    bad = [i for i in range(len(stream)) if got[i + 1]["status"] != 0]
    assert not bad, bad[:10]
    for i, rq in enumerate(stream):
        r = got[i + 1]
        assert normwise_rel(r["strain"][6], rq["strain"]) <= 1e-12, i
    for j, i in enumerate(sample):
        r = got[i + 1]
        assert normwise_rel(r["stress"][6], g["stress_ave"][j]) <= TOL, i
        assert normwise_rel(r["c_tangent"], g["c_tangent"][j]) <= TOL, i
        w = warm[i]
        assert w["status"] == 0
        assert normwise_rel(w["stress"][6], 1.05 * g["stress_ave"][j]) <= TOL, i
        assert w["pcg_its"] < r["pcg_its"], (i, w["pcg_its"], r["pcg_its"])
    # a subset solved alone gives the same bits as in the stream
    with _c5_engine() as eng:
        for i in sample[::8]:
            rq = stream[i]
            assert eng.submit(i + 1, rq["spec"], eye + [rq["strain"]], MM_TANGENT_UNIT) == 0
            eng.drain()
            alone = eng.poll()[0]
            assert np.array_equal(alone["stress"], got[i + 1]["stress"]), i
            assert np.array_equal(alone["c_tangent"], got[i + 1]["c_tangent"]), i
            assert alone["pcg_its"] == got[i + 1]["pcg_its"]
