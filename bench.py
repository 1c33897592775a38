"""Benchmark: homogenized RVE tangents per second at 30^3 hex8 on one MI355X.

A step is one pass of the hot path over one batch: B seeded 30^3 two-phase
RVEs (fiber radius ~ U[0.15, 0.35]) are admitted, assembled on the GPU, their
six unit-strain load cases solved as a 6-RHS block by FP32 Jacobi-PCG inside
FP64 iterative refinement, and their 6x6 tangents averaged in FP64.

The JSON line also carries:
  roofline      the solve's dominant kernel (the PCG step's SpMV + p/q update,
                dictionary ELL storage), timed with HIP events on the engine
                stream over a stand-alone replica of its work
  spmv          the ELL SpMV alone: dictionary storage and the full
                per-row ELL (the reference's layout) on the same batch
  cpu_baseline  the FP64 oracle (one core, one RVE at a time) on a sample
  cpu_baseline_threads  the same oracle on P cores (P = usable cores, at
                most 16), one RVE per thread
  plain_fp64    the plain FP64 single-RVE GPU path (one load case at a time)
  accuracy      max normwise relative tangent error vs the FP64 oracle
  serving       p50 / p99 latency of a short FE^2 request stream

--kernels-only admits and solves the batch once, then runs the timing
replicas and the HBM copy probe (the workload of the counter runs).

--workload selects other seeded workloads, each printing its own JSON line:
  c3  one 100^3 RVE (3.09 M dofs, 4 fibers, random strain): solve time,
      PCG iterations, refinement sweeps, SpMV GB/s of both storages, and the
      mixed-precision stress against an FP64 solve
  c4  512 RVEs of uneven size (8^3..32^3), 6 load cases each: tangents/s
  c5  FE^2 serving stream (12^3/20^3/30^3 RVEs, strain + tangent per request,
      Poisson arrivals), two Newton passes over the same gauss points, the
      second warm-started from the resident RVEs: p50 / p99 latency
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_k_inner_time_128x30cube.md")
METRIC = "RVE tangents/sec at 30³ hex8 (rel err≤1e-8); p99 latency; SpMV HBM GB/s"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


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
    """HBM bytes per launch of the step kernel from the committed counter runs
    (FETCH_SIZE + WRITE_SIZE, calibrated on the copy kernel), if they were
    taken on this workload."""
    if batch != 128 or n != 30 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def ell_api_spmv(mm, spec, reps=20):
    """The reference ELL API's y = A x on the device mirror (FP64 values,
    int32 column ids, 81 slots per row, bit-exact slot order) for the FP64
    matrix of one RVE. Bytes per launch: N * 81 * (8 + 4) + 16 N (matrix,
    x once, y)."""
    import ctypes as C
    rows = 3 * (spec.nex + 1) * (spec.ney + 1) * (spec.nez + 1)
    with mm.Engine(nrhs=1, precision=mm.MM_PREC_FP64, hbm_budget=int(rows * 81 * 16 * 1.3) + (256 << 20),
                   max_rves=2, slab_nodes=256, inner_rtol=1e-10, rtol=1e-10, max_sweeps=2) as eng:
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
    try:
        mm._lib.check(L.ell_device_time_mvp(h, reps, C.byref(ms)), "ell_device_time_mvp")
    finally:
        L.ell_device_destroy(h)
        e.free()
    nb = rows * 81 * 12 + 16 * rows
    gbs = nb / (ms.value * 1e-3) / 1e9
    return dict(ms=ms.value, bytes=nb, gbs=gbs, frac=gbs / HBM_PEAK_GBS, rows=rows)


def workload_c3(args, mm, configs):
    """One 100^3 RVE, one random strain."""
    spec = configs.c3_spec()
    eps = configs.c3_strain()
    rows = 3 * 101 ** 3
    out = {"workload": "c3: one 100^3 hex8 RVE (%d dofs), 4 fibers vf 0.4, strain ~ N(0, 1e-3)" % rows}
    res = {}
    for name, prec, storage in (("mixed_dict", mm.MM_PREC_MIXED, mm.MM_ELL_DICT),
                                ("mixed_full", mm.MM_PREC_MIXED, mm.MM_ELL_FULL),
                                ("fp64_dict", mm.MM_PREC_FP64, mm.MM_ELL_DICT)):
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
        res[name] = r
        out[name] = dict(seconds=dt, status=r["status"], pcg_its=r["pcg_its"], sweeps=r["sweeps"],
                         residual_rel=r["residual_rel"], spmv_ms=ms, spmv_gbs=nb / (ms * 1e-3) / 1e9,
                         spmv_frac=nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, step_ms=sms,
                         step_gbs=sb / (sms * 1e-3) / 1e9)
        log("c3 %s: %.2f s, %d its, %d sweeps, SpMV %.3f ms -> %.0f GB/s"
            % (name, dt, r["pcg_its"], r["sweeps"], ms, out[name]["spmv_gbs"]))
    out["ell_api_spmv"] = ell_api_spmv(mm, spec)
    log("c3 reference-layout FP64 SpMV (ell_device_mvp): %.3f ms -> %.0f GB/s"
        % (out["ell_api_spmv"]["ms"], out["ell_api_spmv"]["gbs"]))
    a, b = res["mixed_dict"]["stress"][0], res["fp64_dict"]["stress"][0]
    out["mixed_vs_fp64_stress_rel"] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    out["full_equals_dict"] = bool(np.array_equal(res["mixed_full"]["stress"][0], a))
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
    out = {"workload": "c4: 512 RVEs, n in {8,12,16,20,24,32}, 6 load cases each", "seconds": dt,
           "tangents_per_s": len(specs) / dt, "dofs": rows, "pcg_its": [min(its), max(its)],
           "all_converged": all(r["status"] == 0 for r in res)}
    log("c4: %.3f s -> %.1f tangents/s" % (dt, len(specs) / dt))
    return out


def workload_c5(args, mm, configs):
    """FE^2 serving stream, two Newton passes over the same gauss points."""
    count = args.count or 4096
    rate = args.rate or 600.0
    stream = configs.c5_stream(count=count, rate=rate)
    eye = [np.array([1e-3 if a == b else 0.0 for b in range(6)]) for a in range(6)]
    passes = []
    with mm.Engine(nrhs=7, precision=mm.MM_PREC_MIXED, hbm_budget=64 << 30, max_rves=count + 64,
                   slab_nodes=256, inner_rtol=1e-5, rtol=1e-9, steps_per_launch=8) as se:
        se.start()
        for ps in range(2):
            got = []
            t0 = time.perf_counter()
            for i, rq in enumerate(stream):
                while time.perf_counter() - t0 < rq["arrival"]:
                    got.extend(se.poll())
                    time.sleep(0.0002)
                strain = rq["strain"] * (1.0 + 0.05 * ps)  # next Newton iterate
                se.submit(ps * count + i + 1, rq["spec"], eye + [strain], mm.MM_TANGENT_UNIT,
                          warm_key=rq["key"], warm_start=ps)
            while len(got) < count and time.perf_counter() - t0 < 600:
                got.extend(se.poll())
                time.sleep(0.0005)
            ts = time.perf_counter() - t0
            lat = np.array([(r["t_done"] - r["t_submit"]) * 1e3 for r in got])
            its = np.array([r["pcg_its"] for r in got])
            passes.append(dict(requests=len(got), seconds=ts, tangents_per_s=len(got) / ts,
                               p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                               mean_pcg_its=float(its.mean()),
                               all_ok=all(r["status"] == 0 for r in got)))
            log("c5 pass %d: %s" % (ps, passes[-1]))
        se.stop()
    return {"workload": "c5: FE^2 stream, %d gauss points (12^3/20^3/30^3 RVEs), Poisson %.0f req/s, "
                        "strain + 6 unit loads per request; pass 2 warm-started" % (count, rate),
            "cold": passes[0], "warm": passes[1]}


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
    ap.add_argument("--cpu-sample", type=int, default=2, help="RVE tangents timed on the CPU")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", type=int, default=1024, help="requests in the serving sample")
    ap.add_argument("--no-full-ell", action="store_true", help="skip the full-ELL SpMV figure")
    ap.add_argument("--kernels-only", action="store_true")
    ap.add_argument("--workload", default="c2", choices=("c2", "c3", "c4", "c5"))
    ap.add_argument("--rate", type=float, default=0.0, help="c5 arrival rate (requests/s)")
    ap.add_argument("--count", type=int, default=0, help="c5 requests per pass")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU; multi-GPU runs are not part of it "
              "(RVEs shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)

    import madmax_amd as mm
    from madmax_amd import configs

    if args.workload != "c2":
        line = {"c3": workload_c3, "c4": workload_c4, "c5": workload_c5}[args.workload](args, mm, configs)
        print(json.dumps(line), flush=True)
        return

    B, n = args.batch, args.n
    specs = [configs.c2_spec(2, i, n=n) for i in range(B)]
    tuned = None
    if args.slab == 0 or args.inner_rtol == 0.0:
        from madmax_amd import autotune
        t = time.perf_counter()
        tuned = autotune.tune(specs[: args.tune_batch or B], rtol=args.rtol, log=log, use_cache=False)
        log("autotune (%.1f s): slab_nodes=%d inner_rtol=%g" % (time.perf_counter() - t,
                                                                tuned["slab_nodes"], tuned["inner_rtol"]))
        if args.slab == 0:
            args.slab = tuned["slab_nodes"]
        if args.inner_rtol == 0.0:
            args.inner_rtol = tuned["inner_rtol"]
    rows = 3 * (n + 1) ** 3
    per_rve = rows * 900 + (1 << 20)
    budget = int(per_rve * B * 1.3) + (256 << 20)
    log("batch %d x %d^3 (%d rows each), arena %.1f GB" % (B, n, rows, budget / 1e9))
    eng = mm.Engine(nrhs=6, precision=mm.MM_PREC_MIXED, hbm_budget=budget, max_rves=B + 8,
                    slab_nodes=args.slab, inner_rtol=args.inner_rtol, rtol=args.rtol,
                    steps_per_launch=args.spl, use_graph=True, ell_storage=mm.MM_ELL_DICT)
    if args.kernels_only:
        import ctypes
        run_batch(eng, specs, mm, warm=True)
        step_ms, step_bytes = eng.time_step(reps=20)
        spmv_ms, spmv_bytes = eng.time_spmv(reps=20)
        eng.close()
        copy_ms = ctypes.c_double()
        mm.lib().mm_hbm_copy_probe(0, 1 << 30, 5, ctypes.byref(copy_ms))
        print(json.dumps({"step_ms": step_ms, "step_bytes": step_bytes, "spmv_ms": spmv_ms,
                          "spmv_bytes": spmv_bytes, "copy_bytes_each_way": 1 << 30,
                          "copy_ms": copy_ms.value}), flush=True)
        return
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

    # roofline of the dominant kernel (the PCG step's SpMV + update) and the
    # SpMV alone, on the resident batch
    run_batch(eng, specs, mm, warm=True)
    step_ms, step_bytes = eng.time_step(reps=50)
    achieved = step_bytes / (step_ms * 1e-3) / 1e9
    log("step kernel: %.3f ms, %.3f GB/launch -> %.0f GB/s" % (step_ms, step_bytes / 1e9, achieved))
    spmv_ms, spmv_bytes = eng.time_spmv(reps=50)
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    log("spmv (dictionary ELL): %.3f ms, %.3f GB/launch -> %.0f GB/s" % (spmv_ms, spmv_bytes / 1e9,
                                                                         spmv_gbs))
    eng.close()
    full = None
    if not args.no_full_ell:
        fb = int((rows * (81 * 4 + 27 * 4 / 3.0) + per_rve) * B * 1.3) + (256 << 20)
        fe = mm.Engine(nrhs=6, precision=mm.MM_PREC_MIXED, hbm_budget=fb, max_rves=B + 8,
                       slab_nodes=args.slab, inner_rtol=args.inner_rtol, rtol=args.rtol,
                       steps_per_launch=args.spl, use_graph=True, ell_storage=mm.MM_ELL_FULL)
        t = time.perf_counter()
        fres = run_batch(fe, specs, mm, warm=True)
        tf = time.perf_counter() - t
        f_ms, f_bytes = fe.time_spmv(reps=50)
        fe.close()
        same = all(np.array_equal(a["c_tangent"], b["c_tangent"]) for a, b in zip(fres, results))
        full = dict(ms=f_ms, bytes=f_bytes, gbs=f_bytes / (f_ms * 1e-3) / 1e9,
                    frac=f_bytes / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    solve_tangents_per_s=B / tf, tangents_equal_to_dictionary_run=same)
        log("spmv (full ELL): %.3f ms, %.3f GB/launch -> %.0f GB/s; batch in %.3f s, same tangents: %s"
            % (f_ms, f_bytes / 1e9, full["gbs"], tf, same))
    import ctypes
    copy_ms = ctypes.c_double()
    copy_bytes = 4 << 30
    mm.lib().mm_hbm_copy_probe(0, copy_bytes, 10, ctypes.byref(copy_ms))
    copy_gbs = 2 * copy_bytes / (copy_ms.value * 1e-3) / 1e9
    log("device copy: %.0f GB/s (read + write of 4 GiB)" % copy_gbs)

    # plain FP64 single-RVE GPU path: the reference's algorithm (zero interior
    # start, one load case at a time), host check per iteration
    plain = None
    if not args.no_plain:
        pe = mm.Engine(nrhs=1, precision=mm.MM_PREC_FP64, hbm_budget=per_rve * 2 + (64 << 20),
                       max_rves=2, slab_nodes=128, inner_rtol=1e-10, rtol=0.0, abs_tol=1e-4,
                       max_sweeps=2, sync_every_step=True, use_graph=False,
                       initial_guess=mm.MM_GUESS_ZERO)
        nplain = 2
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
        plain = dict(value=nplain / tp, unit="RVE tangents/s", n=nplain)
        log("plain fp64: %.2f tangents/s" % plain["value"])

    # CPU baseline: the FP64 oracle, one core, one RVE at a time (bounded sample)
    cpu = None
    cpu_par = None
    acc = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        model = "unknown CPU"
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
        except OSError:
            pass
        errs = []
        t = time.perf_counter()
        for j in range(args.cpu_sample):
            s = specs[j]
            m = po.mesh(3, n, n, n)
            ro = po.Rve(3, n, po.phase_map(m, mm.spec_fibers(s)), [s.young[0], s.young[1]],
                        [s.poisson[0], s.poisson[1]])
            c_ref, _ = ro.c_tangent_unit(1e-3, 1e-10)
            errs.append(float(np.max(np.abs(results[j]["c_tangent"] - c_ref)) / np.max(np.abs(c_ref))))
            if plain is not None and j < len(cs):
                plain.setdefault("err", []).append(
                    float(np.max(np.abs(cs[j] - c_ref)) / np.max(np.abs(c_ref))))
        tc = time.perf_counter() - t
        cpu = dict(value=args.cpu_sample / tc, unit="RVE tangents/s", cores=1, kind="port",
                   sample="%d seeded %d^3 RVE tangents (assembly + 6 FP64 Jacobi-PCG solves to 1e-10 "
                          "+ averaging), oracle/oracle.c gcc -O3, one thread" % (args.cpu_sample, n),
                   cpu_model=model)
        acc = max(errs)
        log("cpu oracle: %.3f tangents/s, max tangent err %.2e" % (cpu["value"], acc))

        # the same oracle on P cores, one RVE per thread (the C calls release
        # the GIL); P = the usable cores, at most 16
        from concurrent.futures import ThreadPoolExecutor
        P = max(1, min(16, len(os.sched_getaffinity(0))))

        def one(j):
            s = specs[j % len(specs)]
            m = po.mesh(3, n, n, n)
            ro = po.Rve(3, n, po.phase_map(m, mm.spec_fibers(s)), [s.young[0], s.young[1]],
                        [s.poisson[0], s.poisson[1]])
            ro.c_tangent_unit(1e-3, 1e-10)

        t = time.perf_counter()
        with ThreadPoolExecutor(P) as ex:
            list(ex.map(one, range(P)))
        tp_ = time.perf_counter() - t
        cpu_par = dict(value=P / tp_, unit="RVE tangents/s", cores=P, kind="port",
                       sample="%d seeded %d^3 RVE tangents, one per thread, same oracle" % (P, n),
                       cpu_model=model)
        log("cpu oracle, %d threads: %.3f tangents/s" % (P, cpu_par["value"]))

    # the reference ELL API's SpMV (FP64, row-major slot order) on one RVE of
    # the batch; its 87 MB matrix fits in the last-level cache
    ell_api = ell_api_spmv(mm, specs[0])
    ell_api["note"] = "one 30^3 RVE, FP64 + int32, 81 slots/row; matrix (87 MB) cache-resident"
    log("reference-layout FP64 SpMV (ell_device_mvp): %.3f ms -> %.0f GB/s" % (ell_api["ms"], ell_api["gbs"]))

    # serving sample (FE^2 stream, mixed RVE sizes, Poisson arrivals)
    serving = None
    if args.serve > 0:
        stream = configs.c5_stream(count=args.serve, rate=600.0)
        se = mm.Engine(nrhs=7, precision=mm.MM_PREC_MIXED, hbm_budget=8 << 30, max_rves=512,
                       slab_nodes=args.slab, inner_rtol=args.inner_rtol, rtol=args.rtol,
                       steps_per_launch=8)
        se.start()
        t0s = time.perf_counter()
        got = []
        for i, rq in enumerate(stream):
            while time.perf_counter() - t0s < rq["arrival"]:
                got.extend(se.poll())
                time.sleep(0.0002)
            loads = [np.array([1e-3 if a == b else 0.0 for b in range(6)]) for a in range(6)]
            se.submit(i + 1, rq["spec"], loads + [rq["strain"]], mm.MM_TANGENT_UNIT)
        while len(got) < len(stream) and time.perf_counter() - t0s < 300:
            got.extend(se.poll())
            time.sleep(0.0005)
        ts = time.perf_counter() - t0s
        se.stop()
        se.close()
        lat = np.array([(r["t_done"] - r["t_submit"]) * 1e3 for r in got])
        serving = dict(requests=len(got), tangents_per_s=len(got) / ts,
                       p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                       all_ok=all(r["status"] == 0 for r in got))
        log("serving: %s" % serving)

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
                   "batch": B, "slab_nodes": args.slab, "inner_rtol": args.inner_rtol,
                   "rtol": args.rtol, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(B, n),
                     "kernel": "k_inner_time<float,6,true>: the PCG step kernel's SpMV + p/q update "
                               "(dictionary ELL), %d resident RVEs" % B,
                     "avg_ms": step_ms, "bytes_per_launch": step_bytes,
                     "copy_gbs_measured": copy_gbs},
        "spmv": {"dictionary": {"ms": spmv_ms, "bytes": spmv_bytes, "gbs": spmv_gbs,
                                "frac": spmv_gbs / HBM_PEAK_GBS},
                 "full_ell": full, "ell_api_fp64": ell_api},
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_par,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "accuracy": {"max_rel_err_tangent": acc, "all_converged": status_ok,
                     "pcg_its": [min(its), max(its)], "ir_sweeps": [min(sweeps), max(sweeps)]},
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
