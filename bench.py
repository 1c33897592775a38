"""Benchmark: homogenized RVE tangents per second at 30^3 hex8 on one MI355X.

A step is one pass of the hot path over one batch: B seeded 30^3 two-phase
RVEs (fiber radius ~ U[0.15, 0.35]) are admitted, assembled on the GPU, their
six unit-strain load cases solved as a 6-RHS block by FP32 Jacobi-PCG inside
FP64 iterative refinement, and their 6x6 tangents averaged in FP64.

The JSON line also carries:
  roofline      the solve's dominant kernel, kernel A (the PCG step's SpMV +
                p/q update, dictionary ELL storage): the median duration of
                its launches inside a solve of the batch (HIP events around
                every launch, one stream, no graph), against the algorithmic
                bytes of one launch; the stand-alone replica of its work is
                reported beside it
  spmv          the ELL SpMV alone: dictionary storage and the full per-row
                ELL on the batch, and the HBM-resident figures of one 100^3
                RVE (2.0 GB FP32 ELL, 3.05 GB FP64 reference-layout ELL)
  c3            the 100^3 RVE: mixed solve time, iterations, sweeps, stress
                error vs the FP64 oracle fixture
  ceiling       hand-written streaming kernels (read-only and copy): the
                HBM rate a kernel of that access pattern reaches on the box
  cpu_baseline  the FP64 oracle (one core, one RVE at a time) on a sample
  cpu_baseline_threads  the same oracle on every usable core, one RVE per
                thread
  fp64_batched  the same batch in FP64 (FP64 PCG, same engine)
  plain_fp64_gpu  the plain FP64 single-RVE path (one load case at a time,
                zero start): with a host check after every PCG step, and with
                graph-replayed steps
  accuracy      max normwise relative tangent error vs the FP64 oracle over
                every RVE the CPU legs solved
  serving       p50 / p99 latency of a short FE^2 request stream

--kernels-only admits and solves the batch once, then runs the timing
replicas and the streaming probes (the workload of the counter runs).

--workload selects other seeded workloads, each printing its own JSON line:
  c3  one 100^3 RVE (3.09 M dofs, 4 fibers, random strain): solve time,
      PCG iterations, refinement sweeps, SpMV GB/s of both storages, and the
      mixed-precision stress against an FP64 solve and the oracle fixture
  c4  512 RVEs of uneven size (8^3..32^3), 6 load cases each: tangents/s
  ell the reference ELL API's device mirror (one 30^3 and one 100^3 RVE):
      FP64 SpMV, and FP64 Jacobi-PCG ms per iteration with the per-step
      kernels (two or four launches per iteration) and with the persistent
      kernel
  c5  FE^2 serving stream (12^3/20^3/30^3 RVEs, strain + tangent per request,
      Poisson arrivals): a sweep of arrival rates (throughput, p50 / p99,
      the saturation knee), then the full 32,768-request stream with a
      warm-started second Newton pass and the stress error of a 64-request
      sample against the oracle fixture
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_k_step_a_march_in_solve_128x30cube.md")
GOLD = os.path.join(ROOT, "tests", "golden")
METRIC = "RVE tangents/sec at 30³ hex8 (rel err≤1e-8); p99 latency; SpMV HBM GB/s"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def run_batch(eng, specs, mm, warm=False):
    for i, s in enumerate(specs):
        rc = eng.submit(i + 1, s, mm.unit_loads(), mm.MM_TANGENT_UNIT,
                        warm_key=(i + 1) if warm else 0)
        if rc:
            raise RuntimeError("submit failed: %d" % rc)
    eng.drain()
    out = {r["id"]: r for r in eng.poll(len(specs) + 8)}
    return [out[i + 1] for i in range(len(specs))]


def measured_traffic(batch, n):
    """HBM bytes per launch of kernel A in the solve from the committed
    counter runs (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), if
    they were taken on this workload."""
    if batch != 128 or n != 30 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def stream_ceilings(mm, nbytes=4 << 30, reps=10):
    """The hand-written read-only and copy kernels over nbytes: GB/s."""
    import ctypes as C
    out = {}
    for name, mode, factor in (("read", 1, 1), ("copy", 0, 2)):
        ms = C.c_double()
        rc = mm.lib().mm_hbm_stream_probe(0, nbytes, reps, mode, C.byref(ms))
        out[name + "_gbs"] = (factor * nbytes / (ms.value * 1e-3) / 1e9) if rc == 0 and ms.value > 0 else None
    out["bytes"] = nbytes
    return out


# Algorithmic bytes of one FP64 Jacobi-PCG iteration per row of the device
# mirror: the matrix (81 slots x (8 B value + 4 B column id)), q = A p with p
# formed from z and p_old (reads z, p_old; writes p, q), the update (reads x,
# p, r, q, dinv; writes x, r, z).
PCG_ROW_BYTES = 81 * 12 + 4 * 8 + 8 * 8


def pcg_ms_per_iteration(mm, h, rows, fused, n1, n2):
    """ms per FP64 Jacobi-PCG iteration of ell_device_solve_pcg on mirror h
    with `fused` iterations per persistent launch (0: the per-step kernels,
    two or four launches per iteration): the slope between solves capped at n1 and n2 iterations
    (zero start, random right-hand side, a target no solve reaches), so the
    transfers of b and x and the set-up cancel.  Best of two per cap."""
    import ctypes as C
    L = mm.lib()
    mm._lib.check(L.ell_device_set_pcg_fused(h, fused), "ell_device_set_pcg_fused")
    b = np.random.default_rng(3).standard_normal(rows)
    bp = b.ctypes.data_as(C.POINTER(C.c_double))

    def run(n):
        x = np.zeros(rows)
        s = mm._lib.EllSolver(n, 0, 1e-300, 0.0)
        t = time.perf_counter()
        rc = L.ell_device_solve_pcg(h, C.byref(s), bp, x.ctypes.data_as(C.POINTER(C.c_double)))
        dt = time.perf_counter() - t
        if rc != 5 or s.its != n:
            raise RuntimeError("pcg timing solve: rc %d after %d of %d iterations" % (rc, s.its, n))
        return dt

    run(n1)
    t1 = min(run(n1) for _ in range(2))
    t2 = min(run(n2) for _ in range(2))
    return (t2 - t1) / (n2 - n1) * 1e3


def ell_api_spmv(mm, spec, reps=20, pcg_its=(60, 260)):
    """The reference ELL API's y = A x on the device mirror (FP64 values,
    int32 column ids, 81 slots per row, bit-exact slot order) for the FP64
    matrix of one RVE, assembled on the GPU by an FP64 engine (no solve).
    Bytes per launch: N * 81 * (8 + 4) + 16 N (matrix, x once, y).  Then the
    FP64 Jacobi-PCG of ell_device_solve_pcg on the same mirror, ms per
    iteration with the per-step kernels and with the persistent kernel at 1,
    32 and all iterations per launch."""
    import ctypes as C
    rows = 3 * (spec.nex + 1) * (spec.ney + 1) * (spec.nez + 1)
    with mm.Engine(nrhs=1, precision=mm.MM_PREC_FP64, hbm_budget=int(rows * 81 * 16 * 1.3) + (256 << 20),
                   max_rves=2, slab_nodes=256, inner_rtol=1e-10, rtol=1e-10, max_sweeps=1,
                   max_inner_its=1) as eng:
        eng.submit(1, spec, [np.array([1e-3, 0, 0, 0, 0, 0])], mm.MM_TANGENT_NONE, warm_key=7)
        eng.drain()
        eng.poll()
        cols, vals = eng.download_ell(7)
    e = mm.Ell.from_arrays(cols, vals)
    del cols, vals
    h = C.c_void_p()
    L = mm.lib()
    mm._lib.check(L.ell_device_create(C.byref(e.m), 0, C.byref(h)), "ell_device_create")
    ms = C.c_double()
    pcg = None
    try:
        mm._lib.check(L.ell_device_time_mvp(h, reps, C.byref(ms)), "ell_device_time_mvp")
        if pcg_its:
            pcg = {"bytes_per_iteration": rows * PCG_ROW_BYTES}
            # per-step kernels: two launches per iteration up to 1024 slabs of
            # 256 rows (every workgroup sums the partials), four beyond
            pcg["per_step_launches_per_iteration"] = 2 if -(-rows // 256) <= 1024 else 4
            for name, f in (("per_step", 0), ("fused_1", 1), ("fused_32", 32), ("fused_all", 1 << 30)):
                t = pcg_ms_per_iteration(mm, h, rows, f, *pcg_its)
                pcg[name] = dict(ms_per_it=t, gbs=rows * PCG_ROW_BYTES / (t * 1e-3) / 1e9)
                log("fp64 pcg (%s), %d rows: %.4f ms/iteration -> %.0f GB/s"
                    % (name, rows, t, pcg[name]["gbs"]))
    finally:
        L.ell_device_destroy(h)
        e.free()
    nb = rows * 81 * 12 + 16 * rows
    gbs = nb / (ms.value * 1e-3) / 1e9
    return dict(ms=ms.value, bytes=nb, gbs=gbs, frac=gbs / HBM_PEAK_GBS, rows=rows, pcg=pcg)


def c3_fixture():
    path = os.path.join(GOLD, "wl_c3.npz")
    return np.load(path) if os.path.exists(path) else None


def c3_solves(mm, configs, names):
    """Solves the 100^3 RVE with the named engines; returns results, the
    stand-alone SpMV / step-kernel timings and the solve times."""
    spec, eps = configs.c3_spec(), configs.c3_strain()
    rows = 3 * 101 ** 3
    table = {"mixed_dict": (mm.MM_PREC_MIXED, mm.MM_ELL_DICT), "mixed_full": (mm.MM_PREC_MIXED, mm.MM_ELL_FULL),
             "fp64_dict": (mm.MM_PREC_FP64, mm.MM_ELL_DICT)}
    out = {}
    for name in names:
        prec, storage = table[name]
        budget = int(rows * (81 * 8 + 200) * 1.3) + (512 << 20)
        with mm.Engine(nrhs=1, precision=prec, hbm_budget=budget, max_rves=2, slab_nodes=256,
                       rtol=1e-10, inner_rtol=1e-5 if prec == mm.MM_PREC_MIXED else 1e-10,
                       max_sweeps=12 if prec == mm.MM_PREC_MIXED else 2, max_inner_its=200000,
                       ell_storage=storage) as eng:
            t = time.perf_counter()
            eng.submit(1, spec, [eps], mm.MM_TANGENT_NONE, warm_key=1)
            eng.drain()
            r = eng.poll()[0]
            dt = time.perf_counter() - t
            ms, nb = eng.time_spmv(reps=20)
            sms, sb = eng.time_step(reps=20)
            # the same request under a new key: admitted and assembled again,
            # solved by the graphs the first solve captured
            t = time.perf_counter()
            eng.submit(2, spec, [eps], mm.MM_TANGENT_NONE, warm_key=2)
            eng.drain()
            r2 = eng.poll()[0]
            dt2 = time.perf_counter() - t
        if not np.array_equal(r2["stress"][0], r["stress"][0]):
            raise RuntimeError("c3 %s: the second solve differs from the first" % name)
        out[name] = dict(result=r, seconds=dt, seconds_captured=dt2, spmv_ms=ms, spmv_bytes=nb, step_ms=sms,
                         step_bytes=sb)
        log("c3 %s: %.3f s (first, graph capture included), %.3f s (captured graphs), %d its, %d sweeps, "
            "SpMV %.3f ms -> %.0f GB/s" % (name, dt, dt2, r["pcg_its"], r["sweeps"], ms, nb / (ms * 1e-3) / 1e9))
    return out


def c3_report(res, fix):
    out = {}
    for name, v in res.items():
        r = v["result"]
        gbs = v["spmv_bytes"] / (v["spmv_ms"] * 1e-3) / 1e9
        out[name] = dict(seconds=v["seconds"], seconds_captured=v["seconds_captured"], status=r["status"],
                         pcg_its=r["pcg_its"], sweeps=r["sweeps"],
                         residual_rel=r["residual_rel"], spmv_ms=v["spmv_ms"], spmv_bytes=v["spmv_bytes"],
                         spmv_gbs=gbs, spmv_frac=gbs / HBM_PEAK_GBS, step_ms=v["step_ms"],
                         step_gbs=v["step_bytes"] / (v["step_ms"] * 1e-3) / 1e9,
                         stress_rel_err_vs_oracle=(rel(r["stress"][0], fix["stress_ave"])
                                                   if fix is not None else None))
    return out


def workload_c3(args, mm, configs):
    """One 100^3 RVE, one random strain."""
    rows = 3 * 101 ** 3
    out = {"workload": "c3: one 100^3 hex8 RVE (%d dofs), 4 fibers vf 0.4, strain ~ N(0, 1e-3)" % rows}
    res = c3_solves(mm, configs, ("mixed_dict", "mixed_full", "fp64_dict"))
    out.update(c3_report(res, c3_fixture()))
    out["ell_api_spmv"] = ell_api_spmv(mm, configs.c3_spec(), pcg_its=(20, 100))
    log("c3 reference-layout FP64 SpMV (ell_device_mvp): %.3f ms -> %.0f GB/s"
        % (out["ell_api_spmv"]["ms"], out["ell_api_spmv"]["gbs"]))
    a, b = res["mixed_dict"]["result"]["stress"][0], res["fp64_dict"]["result"]["stress"][0]
    out["mixed_vs_fp64_stress_rel"] = rel(a, b)
    out["full_equals_dict"] = bool(np.array_equal(res["mixed_full"]["result"]["stress"][0], a))
    out["ceiling"] = stream_ceilings(mm)
    return out


def workload_c4(args, mm, configs):
    """512 RVEs of uneven size, 6 unit load cases each."""
    specs = configs.c4_specs(count=512)
    rows = sum(3 * (s.nex + 1) ** 3 for s in specs)
    budget = int(rows * 250 * 1.3) + (1 << 30)
    with mm.Engine(nrhs=6, precision=mm.MM_PREC_MIXED, hbm_budget=budget, max_rves=520,
                   slab_nodes=256, inner_rtol=1e-5, rtol=1e-9) as eng:
        run_batch(eng, specs[:64], mm)  # warm-up: graph capture
        t = time.perf_counter()
        res = run_batch(eng, specs, mm)
        dt = time.perf_counter() - t
    its = [r["pcg_its"] for r in res]
    errs = []
    path = os.path.join(GOLD, "wl_c4.npz")
    if os.path.exists(path):
        g = np.load(path)
        errs = [rel(res[i]["c_tangent"], c) for i, c in zip(g["index"], g["c_tangent"])]
    out = {"workload": "c4: 512 RVEs, n in {8,12,16,20,24,32}, 6 load cases each", "seconds": dt,
           "tangents_per_s": len(specs) / dt, "dofs": rows, "pcg_its": [min(its), max(its)],
           "all_converged": all(r["status"] == 0 for r in res),
           "max_tangent_rel_err_vs_oracle": max(errs) if errs else None,
           "oracle_checked": len(errs)}
    log("c4: %.3f s -> %.1f tangents/s" % (dt, len(specs) / dt))
    return out


def workload_ell(args, mm, configs):
    """The reference ELL API's device mirror on one 30^3 RVE of the batch and
    on the 100^3 RVE: the bit-exact FP64 SpMV and the FP64 Jacobi-PCG per
    iteration, the per-step kernels against the persistent kernel."""
    out = {"workload": "ell: reference-layout FP64 ELL of one 30^3 and one 100^3 RVE, SpMV and Jacobi-PCG"}
    out["rve_30"] = ell_api_spmv(mm, configs.c2_spec(2, 0, n=30))
    out["rve_100"] = ell_api_spmv(mm, configs.c3_spec(), pcg_its=(20, 100))
    return out


def serve_stream(mm, se, stream, rid0, warm, strain_scale=1.0, limit_s=600):
    """Submits the requests at their arrival times to a started engine and
    collects the results; returns (results by request index, seconds)."""
    eye = [np.array([1e-3 if a == b else 0.0 for b in range(6)]) for a in range(6)]
    got = {}
    t0 = time.perf_counter()
    for i, rq in enumerate(stream):
        while time.perf_counter() - t0 < rq["arrival"]:
            for r in se.poll():
                got[r["id"] - rid0] = r
            time.sleep(0.0002)
        rc = se.submit(rid0 + i, rq["spec"], eye + [strain_scale * rq["strain"]], mm.MM_TANGENT_UNIT,
                       warm_key=rq.get("warm_key", 0), warm_start=1 if warm else 0)
        if rc:
            raise RuntimeError("submit failed: %d" % rc)
    while len(got) < len(stream) and time.perf_counter() - t0 < limit_s:
        for r in se.poll():
            got[r["id"] - rid0] = r
        time.sleep(0.0005)
    return got, time.perf_counter() - t0


def stream_stats(got, seconds):
    rs = list(got.values())
    lat = np.array([(r["t_done"] - r["t_submit"]) * 1e3 for r in rs])
    its = np.array([r["pcg_its"] for r in rs])
    return dict(requests=len(rs), seconds=seconds, tangents_per_s=len(rs) / seconds,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_pcg_its=float(its.mean()), all_ok=all(r["status"] == 0 for r in rs))


def c5_engine(mm, count, march_planes=-1):
    kw = {"march_planes": march_planes} if march_planes >= 0 else {}
    return mm.Engine(nrhs=7, precision=mm.MM_PREC_MIXED, hbm_budget=48 << 30, max_rves=min(count + 64, 4096),
                     slab_nodes=256, inner_rtol=1e-5, rtol=1e-9, steps_per_launch=8, **kw)


def workload_c5(args, mm, configs):
    """FE^2 serving stream: a sweep of arrival rates for the saturation knee,
    then the full stream with a warm-started second Newton pass."""
    sweep = []
    count = args.count or 4096
    rates = [float(r) for r in args.rates.split(",")] if args.rates else [1000.0, 2000.0, 3000.0, 4000.0, 6000.0]
    for rate in rates:
        stream = configs.c5_stream(count=count, rate=rate)
        with c5_engine(mm, count, args.march_planes) as se:
            se.start()
            got, ts = serve_stream(mm, se, stream, 1, False)
            se.stop()
        st = stream_stats(got, ts)
        st["rate"] = rate
        sweep.append(st)
        log("c5 rate %.0f: %s" % (rate, st))
    knee = max((s for s in sweep if s["tangents_per_s"] >= 0.95 * s["rate"] and s["all_ok"]),
               key=lambda s: s["rate"], default=None)
    capacity = max(s["tangents_per_s"] for s in sweep)
    out = {"workload": "c5: FE^2 stream (12^3/20^3/30^3 RVEs), strain + 6 unit loads per request, Poisson "
                       "arrivals", "rate_sweep": sweep, "capacity_tangents_per_s": capacity,
           "knee": knee}
    # the full stream, a 64-request sample kept resident and warm-started
    fix_path = os.path.join(GOLD, "wl_c5.npz")
    g = np.load(fix_path) if os.path.exists(fix_path) else None
    full_rate = args.rate or (knee["rate"] if knee else 2000.0)
    stream = configs.c5_stream(rate=full_rate)
    sample = [int(i) for i in g["index"]] if g is not None else list(range(37, len(stream), 512))
    for i in sample:
        stream[i]["warm_key"] = stream[i]["key"]
    with c5_engine(mm, len(stream), args.march_planes) as se:
        se.start()
        got, ts = serve_stream(mm, se, stream, 1, False)
        cold = stream_stats(got, ts)
        log("c5 full stream at %.0f/s: %s" % (full_rate, cold))
        sub = [dict(stream[i], arrival=0.0) for i in sample]
        wgot, wts = serve_stream(mm, se, sub, 1000000, True, strain_scale=1.05)
        warm = stream_stats(wgot, wts)
        se.stop()
    cold_its = float(np.mean([got[i]["pcg_its"] for i in sample]))
    warm["mean_pcg_its_cold_same_requests"] = cold_its
    log("c5 warm pass over the sample: %s" % warm)
    errs = []
    if g is not None:
        for j, i in enumerate(sample):
            errs.append(rel(got[i]["stress"][6], g["stress_ave"][j]))
            errs.append(rel(got[i]["c_tangent"], g["c_tangent"][j]))
            errs.append(rel(wgot[j]["stress"][6], 1.05 * g["stress_ave"][j]))
    out.update(full_stream=dict(cold, rate=full_rate), warm_pass=warm,
               max_rel_err_vs_oracle=max(errs) if errs else None, oracle_checked=len(sample) if errs else 0)
    # the macro side of the step: the 16^3 macro cube assembled on the GPU from
    # the tangents and stresses the stream returned (request 8 e + g is gauss
    # point g of element e), then one Jacobi-PCG solve with the x = 0 face held
    if len(got) == len(stream) == 32768:
        c_gp = np.zeros((4096, 8, 6, 6))
        s_gp = np.zeros((4096, 8, 6))
        for i in range(len(stream)):
            e, gp = divmod(i, 8)
            c_gp[e, gp] = got[i]["c_tangent"]
            s_gp[e, gp] = got[i]["stress"][6]
        with mm.Macro(16) as mac:
            mac.set_dirichlet([3 * k + d for k in range(17 ** 3) if k % 17 == 0 for d in range(3)])
            t = time.perf_counter()
            mac.assemble_a(c_gp)
            ta = time.perf_counter() - t
            t = time.perf_counter()
            b, nrm = mac.assemble_b(s_gp)
            tb = time.perf_counter() - t
            t = time.perf_counter()
            du, its, rc = mac.solve(b, 1e-8)
            ts_ = time.perf_counter() - t
        out["macro"] = dict(elements=4096, gauss_points=32768, assemble_a_ms=1e3 * ta, assemble_b_ms=1e3 * tb,
                            solve_ms=1e3 * ts_, solve_its=its, solve_status=rc, residual_norm=nrm)
        log("c5 macro assembly: A %.2f ms, b %.2f ms, solve %.2f ms (%d its)" % (1e3 * ta, 1e3 * tb, 1e3 * ts_, its))
    return out


def cpu_model():
    """lscpu's "Model name" verbatim, then its family / model numbers and
    topology (a virtual machine may report only the brand in the name)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return "unknown CPU"
    f = {}
    for ln in out.splitlines():
        if ":" in ln:
            k, v = ln.split(":", 1)
            f.setdefault(k.strip(), v.strip())
    name = f.get("Model name", "unknown CPU")
    extra = ["%s %s" % (k.lower(), f[k]) for k in ("CPU family", "Model", "Socket(s)", "Core(s) per socket",
                                                    "Thread(s) per core") if k in f]
    return "%s (lscpu: %s)" % (name, ", ".join(extra)) if extra else name


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--slab", type=int, default=0, help="nodes per ELL slab (64, 128, 256); 0 = autotune")
    ap.add_argument("--inner-rtol", type=float, default=0.0, help="FP32 PCG target per sweep; 0 = autotune")
    ap.add_argument("--rtol", type=float, default=1e-9, help="FP64 residual target (relative)")
    ap.add_argument("--tune-batch", type=int, default=0, help="RVEs the autotuner solves (0 = the batch)")
    ap.add_argument("--spl", type=int, default=16, help="PCG steps per graph launch")
    ap.add_argument("--streams", type=int, default=2, help="solver stream groups of the batch engine (1 or 2)")
    ap.add_argument("--cpu-sample", type=int, default=2, help="RVE tangents timed on one CPU core")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the 100^3 leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the batched FP64 leg")
    ap.add_argument("--serve", type=int, default=1024, help="requests in the serving sample")
    ap.add_argument("--serve-rate", type=float, default=1000.0, help="arrival rate of the serving sample")
    ap.add_argument("--no-full-ell", action="store_true", help="skip the full-ELL SpMV figure")
    ap.add_argument("--march-planes", type=int, default=-1,
                    help="kernel A's plane-march run length (0: slab by slab; -1: engine default)")
    ap.add_argument("--kernels-only", action="store_true")
    ap.add_argument("--workload", default="c2", choices=("c2", "c3", "c4", "c5", "ell"))
    ap.add_argument("--rate", type=float, default=0.0, help="c5 full-stream arrival rate (requests/s)")
    ap.add_argument("--rates", default="", help="c5 sweep rates, comma separated")
    ap.add_argument("--count", type=int, default=0, help="c5 requests per sweep point")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU; multi-GPU runs are not part of it "
              "(RVEs shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)

    import madmax_amd as mm
    from madmax_amd import configs

    if args.workload != "c2":
        line = {"c3": workload_c3, "c4": workload_c4, "c5": workload_c5,
                "ell": workload_ell}[args.workload](args, mm, configs)
        print(json.dumps(line), flush=True)
        return

    B, n = args.batch, args.n
    specs = [configs.c2_spec(2, i, n=n) for i in range(B)]
    # the batch engine's run length: as given, else the autotuner's, else the
    # engine default (the serving leg keeps the engine default unless given)
    batch_planes = args.march_planes
    if args.slab == 0 or args.inner_rtol == 0.0:
        from madmax_amd import autotune
        t = time.perf_counter()
        tuned = autotune.tune(specs[: args.tune_batch or B], rtol=args.rtol, log=log, use_cache=False)
        log("autotune (%.1f s): slab_nodes=%d march_planes=%d inner_rtol=%g"
            % (time.perf_counter() - t, tuned["slab_nodes"], tuned["march_planes"], tuned["inner_rtol"]))
        if args.slab == 0:
            args.slab = tuned["slab_nodes"]
        if args.inner_rtol == 0.0:
            args.inner_rtol = tuned["inner_rtol"]
        if batch_planes < 0:
            batch_planes = tuned["march_planes"]
    rows = 3 * (n + 1) ** 3
    per_rve = rows * 900 + (1 << 20)
    budget = int(per_rve * B * 1.3) + (256 << 20)
    log("batch %d x %d^3 (%d rows each), arena %.1f GB" % (B, n, rows, budget / 1e9))

    def engine(**kw):
        cfg = dict(nrhs=6, precision=mm.MM_PREC_MIXED, hbm_budget=budget, max_rves=B + 8, slab_nodes=args.slab,
                   inner_rtol=args.inner_rtol, rtol=args.rtol, steps_per_launch=args.spl, use_graph=True,
                   ell_storage=mm.MM_ELL_DICT, solver_streams=args.streams)
        if batch_planes >= 0:
            cfg["march_planes"] = batch_planes
        cfg.update(kw)
        return mm.Engine(**cfg)

    if args.kernels_only:
        # one solve of the batch with every launch outside graphs on one
        # stream (the in-solve kernels of the counter runs), then the replicas
        # and the streaming probes (1 GiB copy: the counter calibration)
        eng = engine(profile_step_kernel=True)
        run_batch(eng, specs, mm, warm=True)
        a_ms = eng.step_kernel_times()
        step_ms, step_bytes = eng.time_step(reps=20)
        spmv_ms, spmv_bytes = eng.time_spmv(reps=20)
        eng.close()
        ceil = stream_ceilings(mm, nbytes=1 << 30, reps=5)
        print(json.dumps({"in_solve_step_ms_median": float(np.median(a_ms)), "in_solve_launches": len(a_ms),
                          "step_ms": step_ms, "step_bytes": step_bytes, "spmv_ms": spmv_ms,
                          "spmv_bytes": spmv_bytes, "ceiling": ceil}), flush=True)
        return
    eng = engine()
    for w in range(args.warmup):
        t = time.perf_counter()
        run_batch(eng, specs, mm)
        log("warmup %d: %.3f s" % (w, time.perf_counter() - t))
    t0 = time.perf_counter()
    results = None
    for k in range(args.steps):
        results = run_batch(eng, specs, mm)
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = B * args.steps / elapsed
    its = [r["pcg_its"] for r in results]
    sweeps = [r["sweeps"] for r in results]
    status_ok = all(r["status"] == 0 for r in results)
    log("timed: %.3f s for %d steps -> %.1f tangents/s (its %d..%d, sweeps %d..%d)"
        % (elapsed, args.steps, value, min(its), max(its), min(sweeps), max(sweeps)))

    # stand-alone replicas of the step kernel and of the SpMV on the resident batch
    run_batch(eng, specs, mm, warm=True)
    rep_ms, step_bytes = eng.time_step(reps=50)
    spmv_ms, spmv_bytes = eng.time_spmv(reps=50)
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    log("step kernel replica: %.3f ms, %.3f GB/launch; spmv (dictionary ELL): %.3f ms -> %.0f GB/s"
        % (rep_ms, step_bytes / 1e9, spmv_ms, spmv_gbs))
    eng.close()

    # kernel A inside a solve of the batch: every launch timed (one stream, no graph)
    pe = engine(profile_step_kernel=True)
    pres = run_batch(pe, specs, mm)
    a_ms = pe.step_kernel_times()
    pe.close()
    in_solve_ms = float(np.median(a_ms))
    same_prof = all(np.array_equal(a["c_tangent"], b["c_tangent"]) for a, b in zip(pres, results))
    achieved = step_bytes / (in_solve_ms * 1e-3) / 1e9
    log("kernel A in the solve: median %.3f ms over %d launches -> %.0f GB/s (tangents equal: %s)"
        % (in_solve_ms, len(a_ms), achieved, same_prof))

    full = None
    if not args.no_full_ell:
        fb = int((rows * (81 * 4 + 27 * 4 / 3.0) + per_rve) * B * 1.3) + (256 << 20)
        fe = engine(hbm_budget=fb, ell_storage=mm.MM_ELL_FULL)
        t = time.perf_counter()
        fres = run_batch(fe, specs, mm, warm=True)
        tf = time.perf_counter() - t
        f_ms, f_bytes = fe.time_spmv(reps=50)
        fe.close()
        same = all(np.array_equal(a["c_tangent"], b["c_tangent"]) for a, b in zip(fres, results))
        full = dict(ms=f_ms, bytes=f_bytes, gbs=f_bytes / (f_ms * 1e-3) / 1e9,
                    frac=f_bytes / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    solve_tangents_per_s=B / tf, tangents_equal_to_dictionary_run=same,
                    note="%d x 30^3 RVEs, 6 loads; 4.7 GB per launch, above the 256 MB last-level cache" % B)
        log("spmv (full ELL): %.3f ms -> %.0f GB/s; batch in %.3f s, same tangents: %s"
            % (f_ms, full["gbs"], tf, same))

    # the same batch in FP64 (FP64 PCG to the same outer tolerance)
    fp64 = None
    if not args.no_fp64:
        de = engine(precision=mm.MM_PREC_FP64, inner_rtol=0.1 * args.rtol, max_sweeps=2)
        run_batch(de, specs, mm)
        t = time.perf_counter()
        nrep = 2
        for _ in range(nrep):
            dres = run_batch(de, specs, mm)
        td = time.perf_counter() - t
        de.close()
        fp64 = dict(value=B * nrep / td, unit="RVE tangents/s", pcg_its=[min(r["pcg_its"] for r in dres),
                                                                        max(r["pcg_its"] for r in dres)],
                    all_converged=all(r["status"] == 0 for r in dres),
                    max_rel_diff_vs_mixed=max(rel(a["c_tangent"], b["c_tangent"]) for a, b in zip(dres, results)))
        log("batched fp64: %.1f tangents/s" % fp64["value"])

    # the 100^3 RVE: HBM-resident SpMV of both layouts and the mixed solve
    c3 = None
    if not args.no_c3:
        res3 = c3_solves(mm, configs, ("mixed_full", "mixed_dict"))
        c3 = c3_report(res3, c3_fixture())
        c3["full_equals_dict"] = bool(np.array_equal(res3["mixed_full"]["result"]["stress"][0],
                                                     res3["mixed_dict"]["result"]["stress"][0]))
        c3["ell_api_fp64"] = ell_api_spmv(mm, configs.c3_spec(), pcg_its=(20, 100))
        log("c3 reference-layout FP64 SpMV (ell_device_mvp): %.3f ms -> %.0f GB/s"
            % (c3["ell_api_fp64"]["ms"], c3["ell_api_fp64"]["gbs"]))

    ceiling = stream_ceilings(mm)
    log("streaming kernels: read %.0f GB/s, copy %.0f GB/s" % (ceiling["read_gbs"] or 0, ceiling["copy_gbs"] or 0))

    # plain FP64 single-RVE GPU path: the reference's algorithm (zero interior
    # start, one load case at a time, Newton limits of micro/src/init.c:7-8)
    plain = None
    if not args.no_plain:
        plain = {}
        nplain = 2
        for name, kw in (("host_sync_per_step", dict(sync_every_step=True, use_graph=False)),
                         ("graph_replay", dict(sync_every_step=False, use_graph=True, steps_per_launch=16))):
            pe = mm.Engine(nrhs=1, precision=mm.MM_PREC_FP64, hbm_budget=per_rve * 2 + (64 << 20),
                           max_rves=2, slab_nodes=128, inner_rtol=1e-10, rtol=0.0, abs_tol=1e-4,
                           max_sweeps=2, initial_guess=mm.MM_GUESS_ZERO, **kw)
            for load in mm.unit_loads()[:1]:  # warm-up (graph capture)
                pe.submit(99, specs[0], [load], mm.MM_TANGENT_NONE, warm_key=99)
                pe.drain()
                pe.poll()
            t = time.perf_counter()
            cs = []
            for j in range(nplain):
                sig = []
                for i, load in enumerate(mm.unit_loads()):
                    pe.submit(10 * j + i + 1, specs[j], [load], mm.MM_TANGENT_NONE, warm_key=j + 1)
                    pe.drain()
                    sig.append(pe.poll()[0]["stress"][0] / 1e-3)
                cs.append(np.array(sig).T)
            tp = time.perf_counter() - t
            pe.close()
            plain[name] = dict(value=nplain / tp, unit="RVE tangents/s", n=nplain,
                               err_vs_batched=[rel(c, results[j]["c_tangent"]) for j, c in enumerate(cs)])
            log("plain fp64 (%s): %.2f tangents/s" % (name, plain[name]["value"]))

    # CPU baselines: the FP64 oracle, one core (bounded sample), then every
    # usable core with one RVE per thread; accuracy over all of them
    cpu = cpu_par = None
    errs = []
    if not args.no_cpu:
        from oracle import pyoracle as po
        model = cpu_model()

        def oracle_tangent(j):
            s = specs[j % len(specs)]
            m = po.mesh(3, n, n, n)
            ro = po.Rve(3, n, po.phase_map(m, mm.spec_fibers(s)), [s.young[0], s.young[1]],
                        [s.poisson[0], s.poisson[1]])
            c_ref, _ = ro.c_tangent_unit(1e-3, 1e-10)
            return c_ref

        t = time.perf_counter()
        for j in range(args.cpu_sample):
            errs.append(rel(results[j]["c_tangent"], oracle_tangent(j)))
        tc = time.perf_counter() - t
        cpu = dict(value=args.cpu_sample / tc, unit="RVE tangents/s", cores=1, kind="port",
                   sample="%d seeded %d^3 RVE tangents (assembly + 6 FP64 Jacobi-PCG solves to 1e-10 "
                          "+ averaging), oracle/oracle.c gcc -O3, one thread" % (args.cpu_sample, n),
                   cpu_model=model)
        log("cpu oracle: %.3f tangents/s" % cpu["value"])

        from concurrent.futures import ThreadPoolExecutor
        P = len(os.sched_getaffinity(0))
        idx = list(range(args.cpu_sample, args.cpu_sample + P))
        t = time.perf_counter()
        with ThreadPoolExecutor(P) as ex:
            refs = list(ex.map(oracle_tangent, idx))
        tp_ = time.perf_counter() - t
        for j, c_ref in zip(idx, refs):
            errs.append(rel(results[j % B]["c_tangent"], c_ref))
        cpu_par = dict(value=P / tp_, unit="RVE tangents/s", cores=P, kind="port",
                       sample="%d seeded %d^3 RVE tangents, one per thread on every usable core, same oracle"
                              % (P, n), cpu_model=model, os_cpu_count=os.cpu_count())
        log("cpu oracle, %d threads: %.3f tangents/s; max tangent err over %d RVEs %.2e"
            % (P, cpu_par["value"], len(errs), max(errs)))

    # the reference ELL API's SpMV (FP64, row-major slot order) on one RVE of
    # the batch; its 87 MB matrix fits in the last-level cache
    ell_api = ell_api_spmv(mm, specs[0])
    ell_api["note"] = "one 30^3 RVE, FP64 + int32, 81 slots/row; matrix (87 MB) cache-resident"
    log("reference-layout FP64 SpMV (ell_device_mvp), 30^3: %.3f ms -> %.0f GB/s" % (ell_api["ms"], ell_api["gbs"]))

    # serving sample (FE^2 stream, mixed RVE sizes, Poisson arrivals)
    serving = None
    if args.serve > 0:
        stream = configs.c5_stream(count=args.serve, rate=args.serve_rate)
        with c5_engine(mm, args.serve, args.march_planes) as se:
            se.start()
            got, ts = serve_stream(mm, se, stream, 1, False, limit_s=300)
            se.stop()
        serving = stream_stats(got, ts)
        serving["rate"] = args.serve_rate
        log("serving: %s" % serving)

    traffic = measured_traffic(B, n)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "RVE tangents/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 PCG in f64 refinement",
        "data": "synthetic (seeded two-phase fiber RVEs)",
        "config": {"workload": "%d x %d^3 hex8 RVE tangents per step, 6-RHS block, "
                               "FP32 PCG + FP64 IR" % (B, n), "rve_elements": n ** 3,
                   "batch": B, "slab_nodes": args.slab, "march_planes": batch_planes, "inner_rtol": args.inner_rtol,
                   "rtol": args.rtol, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": ("constant from profiles/pmc_k_step_a_march_in_solve_128x30cube.md "
                                        "(rocprofv3 FETCH_SIZE / WRITE_SIZE of kernel A's launches in a solve, "
                                        "calibrated on the copy kernel); not measured in this run"
                                        if traffic else None),
                     "kernel": "k_step_a_march_dict<float,6> (kernel A: SpMV + p/q update, dictionary ELL, "
                               "z-plane march through LDS) inside a solve of %d RVEs: median of %d launches"
                               % (B, len(a_ms)),
                     "avg_ms": in_solve_ms, "bytes_per_launch": step_bytes,
                     "replica_ms": rep_ms, "replica_frac": step_bytes / (rep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "ceiling_gbs": ceiling},
        "spmv": {"dictionary": {"ms": spmv_ms, "bytes": spmv_bytes, "gbs": spmv_gbs,
                                "frac": spmv_gbs / HBM_PEAK_GBS},
                 "full_ell": full, "ell_api_fp64_30cube": ell_api,
                 "c3_full_ell_fp32": ({k: c3["mixed_full"][k] for k in ("spmv_ms", "spmv_bytes", "spmv_gbs",
                                                                         "spmv_frac")} if c3 else None),
                 "c3_ell_api_fp64": c3["ell_api_fp64"] if c3 else None},
        "c3": c3,
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_par,
        "fp64_batched": fp64,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain": ({k: value / v["value"] for k, v in plain.items()} if plain else None),
        "speedup_vs_fp64_batched": (value / fp64["value"]) if fp64 else None,
        "accuracy": {"max_rel_err_tangent": max(errs) if errs else None, "rves_checked": len(errs),
                     "all_converged": status_ok, "pcg_its": [min(its), max(its)],
                     "ir_sweeps": [min(sweeps), max(sweeps)]},
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
