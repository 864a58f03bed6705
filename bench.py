"""Benchmark of the PDHG hot path on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on): one
synthetic LP with m=400,000 rows (half equality, half <=), n=1,200,000
columns of degree 2 + Poisson(6) (~9.6M nonzeros), N(0,1) values with 1% of
the columns scaled by 1e3, bounds [0, U[1,100]], feasible and dual feasible by
construction (seed 20261). A "step" is one PDHG iteration in the plain FP64
form of the reference loop (A^T y with the fused primal step, then A xbar
with the fused dual step), everything resident in HBM.

Prints one JSON line: iterations/s over the timed steps, the roofline of the
dominant kernel (A xbar + dual step) from HIP events on the engine's stream,
the time to relative KKT 1e-6 on the same instance of plain FP64 GPU PDHG
(the engine in FP64 with its restarts), of mixed precision and of the
unrestarted reference form (capped; a speedup over it is then a lower
bound), and the reference's own CPU loop
(oracle/_ref, compiled from the reference sources) timed on a bounded sample
of the same instance.

Other workloads: --workload small-lp (the reference's CPU-runnable 2,000 x
5,000 case, GPU and CPU times to 1e-6), --workload mixed (mixed precision on
the same instance),
--workload planning512 (512 LPs in the batched persistent engine),
--workload warmstart (learned warm start on a parametric family),
--workload pdlp-fp64 (the restarted engine in FP64 on the large LP) and
--workload dense-row (iteration time on LPs with one dense row).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-1e-6 KKT & PDHG iters/s vs plain FP64 GPU PDHG; % HBM roofline in SpMV"
HBM_PEAK_GBS = 8000.0
SEED = 20261


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(lp):
    """The reference PdhgSolver::solve (oracle/_ref/libref.so, threads=1, its
    default) on the same instance. Its loop can only stop on the wall-clock
    deadline, checked every 256 iterations, so the sample is exactly 256
    iterations: one run with time_limit=0 measures the setup (power iteration
    and the CSR copy), a second run stops at iteration 256."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Reference, have_ref
    if not have_ref():
        return None
    ref = Reference(lp)
    _, it0, setup = ref.pdhg_timed(0.0)
    _, it, total = ref.pdhg_timed(setup * 1.5 + 0.5)
    if it <= 0:
        return None
    out = {"value": it / (total - setup), "unit": "PDHG iterations/s", "cores": 1, "kind": "reference",
           "sample": f"{it} iterations of the reference loop on the full instance "
                     f"({total - setup:.1f} s after {setup:.1f} s setup), {os.cpu_count()} host cores "
                     f"available, CPU {cpu_model()}"}
    # The reference's own parallel setting (Options::threads = host cores):
    # only its primal update runs on the thread pool (pdhg_solver.hpp:71).
    # The box allots 16 cores to a GPU job whatever os.cpu_count() reports.
    nth = max(1, min(16, os.cpu_count() or 1))
    _, itp, totp = ref.pdhg_timed(setup * 1.5 + 0.5, threads=nth)
    if itp > 0:
        out["parallel"] = {"value": itp / (totp - setup), "threads": nth,
                           "sample": f"{itp} iterations, {totp - setup:.1f} s after setup"}
    return out


def plain_to_tolerance(eng, lp, cap_s, tol=1e-6, chunk=8192):
    """The unrestarted reference form on the GPU (the reference's loop,
    reference mode, autotuned layout) run in chunks of `chunk` iterations, each followed by a
    device evaluation of the relative KKT of its iterate, until the KKT
    reaches `tol`, the reference's own stop rule ends the loop, or `cap_s`
    seconds of iteration time have passed. `solve_s` counts the iteration
    chunks only (the evaluations are not part of the plain loop)."""
    import numpy as np
    from nirnaya_amd.engine import MODE_REFERENCE, default_params
    p = default_params(mode=MODE_REFERENCE, max_iters=2 ** 31 - 1, check_termination=1, tile_nnz=-1)
    s = eng.session(lp, p)
    solve = 0.0
    trace = []
    k = np.zeros(5)
    reached = stopped = False
    asked = 0
    try:
        while solve < cap_s:
            t0 = time.perf_counter()
            s.run(chunk)
            solve += time.perf_counter() - t0
            asked += chunk
            k = s.kkt()
            info = s.result().info
            reached = max(k[:3]) <= tol
            stopped = info.iterations < asked   # the reference's own stop rule ended the loop
            if not trace or solve - trace[-1][1] >= 10.0 or reached or stopped:
                trace.append([int(info.iterations), round(solve, 2), float(max(k[:3]))])
                log(f"[bench] reference form: {trace[-1][0]} iterations, {trace[-1][1]} s, relative KKT "
                    f"{trace[-1][2]:.3e}")
            if reached or stopped:
                break
        info = s.result().info
    finally:
        s.close()
    out = {"reached": bool(reached), "iterations": int(info.iterations), "solve_s": round(solve, 3),
           "stopped_by_reference_rule": bool(stopped), "status": int(info.status), "rel_kkt": [float(v) for v in k[:3]], "trace_it_s_kkt": trace}
    log(f"[bench] reference form (unrestarted) toward 1e-6: {out}")
    return out


def time_to_tolerance(eng_factory, lp, plain_cap_s):
    """Time to relative KKT 1e-6 on the same instance: plain FP64 GPU PDHG -
    the engine in FP64, one instance, cold start, with its restarts (the
    baseline the metric names) -, the same in re-centred mixed precision, and
    the unrestarted reference form on the GPU. The restarted solves include
    the session setup (upload, layout autotune, scaling, power iteration) in
    `total_s`; `solve_s` is the iteration loop with everything resident. The
    reference form stops at `plain_cap_s` of iteration time; if it has not
    reached the tolerance by then, the speedup over it is a lower bound."""
    from nirnaya_amd.engine import MODE_PDLP, PREC_MIXED, default_params
    out = {}
    eng = eng_factory()
    for name, prec in (("fp64", 0), ("mixed", PREC_MIXED)):
        p = default_params(mode=MODE_PDLP, precision=prec, max_iters=2_000_000, tol=1e-6)
        t0 = time.perf_counter()
        r = eng.solve(lp, p)
        total = time.perf_counter() - t0
        i = r.info
        out[name] = {"status": r.status, "iterations": r.iterations, "fp64_tail_iterations": i.polish_iterations,
                     "restarts": i.restarts, "solve_s": round(i.solve_seconds, 3), "total_s": round(total, 3),
                     "rel_kkt": [i.rel_primal, i.rel_dual, i.rel_gap], "objective": r.objective}
        log(f"[bench] time to 1e-6 ({name}): {out[name]}")
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["solve_s"] / out["mixed"]["solve_s"], 3)
    if plain_cap_s > 0:
        pl = plain_to_tolerance(eng, lp, plain_cap_s)
        out["reference_form_fp64"] = pl
        out["speedup_mixed_vs_reference_form"] = round(pl["solve_s"] / out["mixed"]["solve_s"], 3)
        out["speedup_fp64_vs_reference_form"] = round(pl["solve_s"] / out["fp64"]["solve_s"], 3)
        out["speedup_vs_reference_form_is_lower_bound"] = not pl["reached"]
    eng.close()
    return out


def batched_leg():
    """All 512 planning LPs (seeds 20263*1000 + i) to relative KKT 1e-6, all
    measured: the batched persistent engine in FP64 and in mixed precision
    (setup + solve), and plain FP64 GPU PDHG one LP at a time (the single-LP
    engine, every one of the 512 solved and timed)."""
    import numpy as np
    from nirnaya_amd.engine import Engine, MODE_PDLP, PREC_MIXED, default_params, gen_planning_lp
    lps = [gen_planning_lp(20263 * 1000 + i) for i in range(512)]
    eng = Engine(0)
    out = {}
    for name, prec in (("batch_fp64", 0), ("batch_mixed", PREC_MIXED)):
        p = default_params(mode=MODE_PDLP, precision=prec, max_iters=400000, tol=1e-6)
        t0 = time.perf_counter()
        b = eng.batch(lps, p)
        b.solve()
        dt = time.perf_counter() - t0
        res = [b.result(i) for i in range(len(lps))]
        b.close()
        out[name] = {"seconds": round(dt, 3), "optimal": int(sum(r.status == 0 for r in res)),
                     "iterations": int(sum(r.iterations for r in res))}
        log(f"[bench] 512 planning LPs, {name}: {out[name]}")
    p = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6)
    t0 = time.perf_counter()
    opt = its = 0
    for lp in lps:
        r = eng.solve(lp, p)
        opt += r.status == 0
        its += r.iterations
    one = time.perf_counter() - t0
    eng.close()
    out["one_at_a_time_fp64"] = {"seconds": round(one, 3), "optimal": int(opt), "iterations": int(its)}
    out["speedup_batch_fp64"] = round(one / out["batch_fp64"]["seconds"], 2)
    out["speedup_batch_mixed"] = round(one / out["batch_mixed"]["seconds"], 2)
    log(f"[bench] 512 planning LPs one at a time (FP64 single-LP engine): {one:.1f}s; batch speedups "
        f"{out['speedup_batch_fp64']}x (FP64), {out['speedup_batch_mixed']}x (mixed)")
    return out


def planning512(args):
    """Batched workload: 512 planning LPs of uneven size (T ~ U{4..48}
    periods x K ~ U{5..40} products, seeds 20263*1000 + i), restarted PDHG to
    relative KKT 1e-6. Times the batched persistent engine (setup + solve) in
    FP64 and in mixed precision, and the FP64 single-LP engine one LP at a
    time on every 32nd instance; the one-at-a-time total is that sample's
    time per iteration times the exact iteration counts of all 512 (the batch
    FP64 trajectory of each instance is the single-LP one, bit for bit)."""
    import numpy as np
    from nirnaya_amd.engine import Engine, MODE_PDLP, PREC_MIXED, default_params, gen_planning_lp
    t0 = time.time()
    lps = [gen_planning_lp(20263 * 1000 + i) for i in range(512)]
    ms = np.array([lp.m for lp in lps])
    log(f"[bench] 512 planning LPs generated in {time.time() - t0:.1f}s, m {ms.min()}..{ms.max()}, "
        f"nnz {sum(lp.nnz for lp in lps)}")
    eng = Engine(0)
    out = {}
    for name, prec in (("fp64", 0), ("mixed", PREC_MIXED)):
        p = default_params(mode=MODE_PDLP, precision=prec, max_iters=400000, tol=1e-6)
        t0 = time.perf_counter()
        b = eng.batch(lps, p)
        t1 = time.perf_counter()
        b.solve()
        t2 = time.perf_counter()
        res = [b.result(i) for i in range(len(lps))]
        st = b.stats()
        b.close()
        out[name] = {"setup_s": t1 - t0, "solve_s": t2 - t1, "total_s": t2 - t0,
                     "optimal": int(sum(r.status == 0 for r in res)), "iterations": int(st.total_iterations),
                     "threads": st.threads, "launches": int(st.launches),
                     "nnz_iterations_per_s": st.nnz_iterations / (t2 - t1)}
        if name == "fp64":
            its = np.array([r.iterations for r in res])
        log(f"[bench] batch {name}: {out[name]}")
    # Mixed precision with the FP64 polish tail to 1e-8 (relative KKT of every
    # instance under the FP64 evaluation).
    p = default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=2_000_000, tol=1e-6, polish_tol=1e-8)
    t0 = time.perf_counter()
    b = eng.batch(lps, p)
    t1 = time.perf_counter()
    b.solve()
    t2 = time.perf_counter()
    res = [b.result(i) for i in range(len(lps))]
    st = b.stats()
    b.close()
    worst = max(max(r.info.rel_primal, r.info.rel_dual, r.info.rel_gap) for r in res)
    out["mixed_polish_1e-8"] = {"setup_s": t1 - t0, "solve_s": t2 - t1, "total_s": t2 - t0,
                                "optimal": int(sum(r.status == 0 for r in res)),
                                "iterations": int(st.total_iterations),
                                "fp64_tail_iterations": int(sum(r.info.polish_iterations for r in res)),
                                "worst_rel_kkt": worst}
    log(f"[bench] batch mixed + FP64 polish to 1e-8: {out['mixed_polish_1e-8']}")
    # The critical path: the instance with the most nonzero-iterations, alone
    # in a batch (one workgroup), against the whole batch's solve time.
    work = its * np.array([lp.nnz for lp in lps], dtype=np.float64)
    big = int(np.argmax(work))
    for name, prec in (("fp64", 0), ("mixed", PREC_MIXED)):
        b = eng.batch([lps[big]], default_params(mode=MODE_PDLP, precision=prec, max_iters=400000, tol=1e-6))
        t0 = time.perf_counter()
        b.solve()
        out[name]["largest_alone_s"] = round(time.perf_counter() - t0, 3)
        b.close()
    out["largest"] = {"instance": big, "m": int(lps[big].m), "n": int(lps[big].n), "nnz": int(lps[big].nnz),
                      "fp64_iterations": int(its[big]), "work_share": round(float(work[big] / work.sum()), 4)}
    log(f"[bench] largest instance: {out['largest']}, alone fp64 {out['fp64']['largest_alone_s']} s, "
        f"mixed {out['mixed']['largest_alone_s']} s")
    p = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6)
    samp = [] if args.batch_only else list(range(0, 512, 32))
    ts, si = 0.0, 0
    for i in samp:
        t0 = time.perf_counter()
        r = eng.solve(lps[i], p)
        ts += time.perf_counter() - t0
        si += r.iterations
        assert r.iterations == its[i]
    one = ts / si * its.sum() if samp else None
    # FP64 one at a time to 1e-8 on the same sample, scaled by the instance
    # count (the batch's iteration counts to 1e-8 differ: mixed phase + FP64
    # tail), against the mixed + polish batch.
    one8 = None
    if samp:
        p8 = default_params(mode=MODE_PDLP, max_iters=2_000_000, tol=1e-8)
        t8 = 0.0
        for i in samp:
            t0 = time.perf_counter()
            r = eng.solve(lps[i], p8)
            t8 += time.perf_counter() - t0
            assert r.status == 0
        one8 = t8 / len(samp) * len(lps)
    eng.close()
    hist, edges = np.histogram(ms, bins=[0, 1000, 2000, 5000, 10000, 20000])
    line = {
        "metric": "time-to-1e-6 relative KKT, 512 planning LPs (batched persistent engine)",
        "value": round(out["fp64"]["total_s"], 3), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "512 multi-period planning LPs, T~U{4..48} x K~U{5..40}, seeds 20263*1000+i",
                   "rows_histogram": {f"{int(a)}-{int(b)}": int(h) for a, b, h in zip(edges[:-1], edges[1:], hist)}},
        "batch_fp64": out["fp64"], "batch_mixed": out["mixed"], "batch_mixed_polish_1e-8": out["mixed_polish_1e-8"],
        "largest_instance": out["largest"],
        "one_at_a_time_fp64_s": round(one, 3) if one else None,
        "one_at_a_time_sample": f"{len(samp)} instances (every 32nd), {ts:.2f}s for {si} iterations",
        "speedup_vs_one_at_a_time": round(one / out["fp64"]["total_s"], 2) if one else None,
        "one_at_a_time_fp64_1e-8_s": round(one8, 3) if one8 else None,
        "speedup_mixed_polish_1e-8_vs_one_at_a_time": round(one8 / out["mixed_polish_1e-8"]["total_s"], 2)
        if one8 else None,
    }
    print(json.dumps(line), flush=True)


def small_lp_workload(args):
    """The reference's CPU-runnable case: the 2,000 x 5,000 equality LP (5
    nonzeros per column at uniform rows, N(0,1) values, seed 20260), time to
    relative KKT 1e-6.
    GPU: plain FP64 GPU PDHG - the engine in FP64, one instance, cold start,
    with its restarts (the baseline the metric names) -, the same in mixed
    precision, and the unrestarted reference form (its relative KKT evaluated
    on the device every 64 iterations, timer paused, capped at 50,000
    iterations, the reference's default limit).
    CPU (the host of the GPU, one thread): the reference loop's restatement in
    the oracle run the same way (KKT every 64 iterations, timer paused, cap
    50,000), the restarted FP64 oracle, and the reference's own loop
    (oracle/_ref) iterations/s at threads=1 and at the host's cores."""
    import numpy as np
    from nirnaya_amd.engine import Engine, MODE_PDLP, MODE_REFERENCE, PREC_MIXED, default_params, gen_equality_lp
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, Reference, have_ref
    lp = gen_equality_lp(20260, 2000, 5000, 5)
    tol, cap, chunk = 1e-6, 50_000, 64
    eng = Engine(0)
    gpu = {}
    for name, prec in (("fp64", 0), ("mixed", PREC_MIXED)):
        q = default_params(mode=MODE_PDLP, precision=prec, max_iters=2_000_000, tol=tol)
        eng.solve(lp, q)   # warm the code paths
        t0 = time.perf_counter()
        r = eng.solve(lp, q)
        gpu[name] = {"status": r.status, "iterations": r.iterations, "solve_s": round(r.info.solve_seconds, 5),
                     "total_s": round(time.perf_counter() - t0, 5),
                     "rel_kkt": [r.info.rel_primal, r.info.rel_dual, r.info.rel_gap]}
        log(f"[bench] GPU {name} to 1e-6: {gpu[name]}")
        # The same solve as a batch of one: one persistent workgroup runs the
        # whole solve (the same iterates, bit for bit), no per-iteration
        # launches; setup timed apart.
        t0 = time.perf_counter()
        b = eng.batch([lp], q)
        t1 = time.perf_counter()
        b.solve()
        t2 = time.perf_counter()
        rb = b.result(0)
        b.close()
        gpu[name + "_batch_of_one"] = {"status": rb.status, "iterations": rb.iterations,
                                       "setup_s": round(t1 - t0, 5), "solve_s": round(t2 - t1, 5),
                                       "same_iterates": bool(np.array_equal(rb.x, r.x) and np.array_equal(rb.y, r.y))}
        log(f"[bench] GPU {name} as a batch of one: {gpu[name + '_batch_of_one']}")
    # reference form on the GPU, KKT on the device between chunks
    s = eng.session(lp, default_params(mode=MODE_REFERENCE, max_iters=2 ** 31 - 1, check_termination=0))
    solve, it, k = 0.0, 0, np.ones(5)
    while it < cap and max(k[:3]) > tol:
        step = min(chunk, cap - it)
        t0 = time.perf_counter()
        s.run(step)
        solve += time.perf_counter() - t0
        it += step
        k = s.kkt()
    s.close()
    eng.close()
    gpu["reference_form"] = {"reached": bool(max(k[:3]) <= tol), "iterations": it, "solve_s": round(solve, 4),
                             "rel_kkt": [float(v) for v in k[:3]]}
    log(f"[bench] GPU reference form: {gpu['reference_form']}")
    # CPU: the reference loop's restatement, chunked with paused KKT checks
    o = Oracle(lp)
    x, y, tau, _, _ = o.pdhg(0, check=False)
    solve, it, k = 0.0, 0, np.ones(5)
    while it < cap and max(k[:3]) > tol:
        step = min(chunk, cap - it)
        t0 = time.perf_counter()
        o.steps(step, tau, x, y)
        solve += time.perf_counter() - t0
        it += step
        k = o.kkt(x, y)
    cpu_ref_form = {"reached": bool(max(k[:3]) <= tol), "iterations": it, "solve_s": round(solve, 3),
                    "rel_kkt": [float(v) for v in k[:3]], "kind": "port", "cores": 1}
    log(f"[bench] CPU reference form (oracle): {cpu_ref_form}")
    t0 = time.perf_counter()
    _, _, info = o.pdlp(max_iters=2_000_000, tol=tol)
    cpu_pdlp = {"status": info.status, "iterations": info.iterations, "total_s": round(time.perf_counter() - t0, 3),
                "kind": "port", "cores": 1}
    log(f"[bench] CPU restarted FP64 (oracle): {cpu_pdlp}")
    cpu = None
    if have_ref():
        ref = Reference(lp)
        nth = max(1, min(16, os.cpu_count() or 1))
        rates = {}
        for th in (1, nth):
            _, it0, setup = ref.pdhg_timed(0.0, threads=th)
            _, it1, total = ref.pdhg_timed(setup + 1.0, threads=th)
            rates[th] = {"iterations": it1, "seconds": round(total - setup, 3),
                         "value": it1 / max(total - setup, 1e-9)}
        cpu = {"value": rates[1]["value"], "unit": "PDHG iterations/s", "cores": 1, "kind": "reference",
               "sample": f"{rates[1]['iterations']} iterations of the reference loop in {rates[1]['seconds']} s "
                         f"after setup, CPU {cpu_model()}",
               "parallel": {"threads": nth, **rates[nth]},
               "reference_form_time_to_1e-6": cpu_ref_form, "restarted_fp64_time_to_1e-6": cpu_pdlp}
    line = {
        "metric": "time to relative KKT 1e-6, 2,000 x 5,000 equality LP", "value": gpu["fp64"]["solve_s"],
        "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "2,000 equality rows x 5,000 columns, 5 nonzeros per column at uniform rows, "
                               "N(0,1) values, bounds [0,10], seed 20260", "m": lp.m, "n": lp.n, "nnz": lp.nnz},
        "gpu": gpu,
        "speedup_gpu_fp64_vs_cpu_restarted": round(cpu_pdlp["total_s"] / gpu["fp64"]["total_s"], 2),
        "cpu_baseline": cpu or {"reference_form_time_to_1e-6": cpu_ref_form, "restarted_fp64_time_to_1e-6": cpu_pdlp},
    }
    print(json.dumps(line), flush=True)


def mixed_workload(args):
    """Mixed precision on the 400k x 1.2M instance: the restarted engine's FP32
    phase (FP32 values, iterates and accumulation in the iteration kernels;
    FP64 evaluations, restarts, primal weight and objective). `value` is its
    iteration rate over `steps` iterations of the production path (graph
    replays, evaluation blocks included); the roofline is the FP32 A xbar
    kernel's, from the events sweep. Then both precisions are solved to
    relative KKT 1e-6 and the mixed solution's error against the FP64 one is
    reported (the FP64 engine's trajectory is the oracle's, bit for bit, in
    the parity tests)."""
    import numpy as np
    from nirnaya_amd.engine import Engine, MODE_PDLP, PREC_MIXED, default_params, gen_mixed_lp
    t0 = time.time()
    lp = gen_mixed_lp(SEED, 400_000, 1_200_000, 6.0, 0.01, 1e3)
    log(f"[bench] instance m={lp.m} n={lp.n} nnz={lp.nnz} generated in {time.time() - t0:.1f}s")
    eng = Engine(0)
    if args.calibrate:
        ms = eng.calibrate_copy(256 << 20)
        log(f"[bench] calibration copy of 256 MiB: {ms:.3f} ms")
    # tol far below reach: the timed iterations stay in the FP32 phase
    p = default_params(mode=MODE_PDLP, precision=PREC_MIXED, max_iters=10 ** 9, tol=1e-14, tile_nnz=args.tile,
                       tile_rows=args.tile_rows, seg_width=args.seg, seg_width_t=args.seg_t)
    sess = eng.session(lp, p)
    shape = sess.shape()
    log(f"[bench] mixed session ready in {sess.result().info.setup_seconds:.2f}s; shape A x: tile {shape[0]} "
        f"segments {shape[2]}, A^T y: tile {shape[4]} segments {shape[6]}")
    sess.run(args.warmup)
    t0 = time.perf_counter()
    sess.run(args.steps)
    wall = time.perf_counter() - t0
    t = sess.time(args.steps)
    sess.close()
    it_s = args.steps / wall
    spmv_avg_s = t.spmv_ms / t.iterations / 1e3
    spmtv_avg_s = t.spmtv_ms / t.iterations / 1e3
    achieved = t.spmv_bytes / spmv_avg_s / 1e9
    log(f"[bench] FP32 phase: {it_s:.0f} it/s; A xbar {spmv_avg_s * 1e6:.1f} us ({achieved:.0f} GB/s), "
        f"A^T y {spmtv_avg_s * 1e6:.1f} us")
    sol = {}
    ttt = {}
    for name, prec in (() if args.no_ttt else (("fp64", 0), ("mixed", PREC_MIXED))):
        q = default_params(mode=MODE_PDLP, precision=prec, max_iters=2_000_000, tol=1e-6)
        t0 = time.perf_counter()
        r = eng.solve(lp, q)
        sol[name] = r
        ttt[name] = {"status": r.status, "iterations": r.iterations, "solve_s": round(r.info.solve_seconds, 3),
                     "total_s": round(time.perf_counter() - t0, 3),
                     "rel_kkt": [r.info.rel_primal, r.info.rel_dual, r.info.rel_gap], "objective": r.objective}
        log(f"[bench] to 1e-6 ({name}): {ttt[name]}")
    eng.close()
    err = None
    if sol:
        a, b = sol["mixed"], sol["fp64"]
        err = {"x_rel_l2": float(np.linalg.norm(a.x - b.x) / max(np.linalg.norm(b.x), 1e-300)),
               "y_rel_l2": float(np.linalg.norm(a.y - b.y) / max(np.linalg.norm(b.y), 1e-300)),
               "objective_rel": abs(a.objective - b.objective) / (1.0 + abs(b.objective))}
        ttt["speedup_mixed_vs_fp64"] = round(ttt["fp64"]["solve_s"] / ttt["mixed"]["solve_s"], 3)
    cpu = None if args.no_cpu_baseline else cpu_baseline(lp)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "spmv_cfg2_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            for e in json.load(f).get("layouts", []):
                if e.get("tile") == shape[0] and e.get("segments") == shape[2]:
                    traffic = e.get("hbm_bytes_per_launch")
    line = {
        "metric": METRIC, "value": round(it_s, 2), "unit": "PDHG iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (iterations) + f64 (evaluations)",
        "data": "synthetic",
        "config": {"workload": "mixed precision on the single large LP: m=400000, n=1200000, "
                               f"nnz={lp.nnz}, seed {SEED}, restarted PDHG FP32 phase",
                   "m": lp.m, "n": lp.n, "nnz": lp.nnz},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_dual_pdlp<float> (FP32 A xbar + dual step); one launch = its segment passes",
                     "bytes_per_launch": t.spmv_bytes, "avg_launch_us": spmv_avg_s * 1e6,
                     "launches_averaged": t.iterations,
                     "dispatches_per_launch": t.spmv_passes},
        "kernels": {"spmtv_primal_avg_us": spmtv_avg_s * 1e6, "spmtv_bytes_per_launch": t.spmtv_bytes,
                    "iteration_bytes": t.iter_bytes},
        "time_to_1e-6": ttt or None,
        "mixed_error_vs_fp64": err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def pdlp_fp64_workload(args):
    """Plain FP64 GPU PDHG - the restarted engine in FP64, one instance, cold
    start, the baseline the headline metric's time-to-tolerance ratios are
    taken against - on the 400k x 1.2M instance: its iteration rate over
    `steps` iterations of the production path (graph replays, evaluation
    blocks included; tolerance out of reach), and the roofline of its FP64
    A xbar kernel (k_dual_pdlp<double>) from the events sweep."""
    from nirnaya_amd.engine import Engine, MODE_PDLP, default_params, gen_mixed_lp
    lp = gen_mixed_lp(SEED, 400_000, 1_200_000, 6.0, 0.01, 1e3)
    eng = Engine(0)
    p = default_params(mode=MODE_PDLP, precision=0, max_iters=10 ** 9, tol=1e-14, tile_nnz=args.tile,
                       tile_rows=args.tile_rows, seg_width=args.seg, seg_width_t=args.seg_t)
    sess = eng.session(lp, p)
    shape = sess.shape()
    log(f"[bench] FP64 pdlp session ready in {sess.result().info.setup_seconds:.2f}s; shape A x: tile {shape[0]} "
        f"segments {shape[2]}, A^T y: tile {shape[4]} segments {shape[6]}")
    sess.run(args.warmup)
    t0 = time.perf_counter()
    sess.run(args.steps)
    wall = time.perf_counter() - t0
    t = sess.time(args.steps)
    sess.close()
    eng.close()
    it_s = args.steps / wall
    spmv_avg_s = t.spmv_ms / t.iterations / 1e3
    spmtv_avg_s = t.spmtv_ms / t.iterations / 1e3
    achieved = t.spmv_bytes / spmv_avg_s / 1e9
    log(f"[bench] FP64 pdlp: {it_s:.0f} it/s; A xbar {spmv_avg_s * 1e6:.1f} us ({achieved:.0f} GB/s), "
        f"A^T y {spmtv_avg_s * 1e6:.1f} us")
    print(json.dumps({
        "metric": "restarted FP64 PDHG iterations/s (plain FP64 GPU PDHG)", "value": round(it_s, 2),
        "unit": "PDHG iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"single large LP: m=400000, n=1200000, nnz={lp.nnz}, seed {SEED}, pdlp mode FP64",
                   "shape": list(shape)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "kernel": "k_dual_pdlp<double> (FP64 A xbar + dual step); one launch = its segment passes",
                     "bytes_per_launch": t.spmv_bytes, "avg_launch_us": spmv_avg_s * 1e6,
                     "launches_averaged": t.iterations, "dispatches_per_launch": t.spmv_passes},
        "kernels": {"spmtv_primal_avg_us": spmtv_avg_s * 1e6, "spmtv_bytes_per_launch": t.spmtv_bytes}}),
        flush=True)


def warmstart_workload(args):
    """Learned warm start on the parametric family (fixed A as the 2,000 x
    5,000 equality LP scaled to m x n, 5 nonzeros per column; b = b0 (1 + 0.1
    tanh(B theta)), c = c0 + 0.05 C theta, theta in R^64): `ws_train`
    training solves on the family engine (64 members per wave, one copy of
    A) to `ws_train_tol`, POD bases of rank 256 (x) + 256 (y) from the
    snapshot Gram matrices on the device, the 64 -> 512 -> 512 -> 512 bf16 MLP
    trained on the device, then `ws_test` held-out thetas to relative KKT 1e-6
    cold and warm-started (family engine). Time to 1e-6 is reported for the
    warm-started family solve, the cold family solve, and plain FP64 GPU PDHG
    one LP at a time (the single-LP engine; a sample of held-out LPs timed and
    scaled by the exact iteration counts of all of them, which the family
    engine reproduces bit for bit)."""
    import numpy as np
    from nirnaya_amd.engine import Engine, MODE_PDLP, default_params
    from nirnaya_amd.warmstart import ParametricFamily, Pod, WarmStart, family_solve
    m, n = args.ws_m, args.ws_n
    t_all = time.perf_counter()
    fam = ParametricFamily(20264, m, n, 5)
    eng = Engine(0)
    prec = 1 if args.precision == "mixed" else 0
    params = default_params(mode=MODE_PDLP, precision=prec, max_iters=400000, tol=1e-6)
    tparams = default_params(mode=MODE_PDLP, precision=prec, max_iters=400000, tol=args.ws_train_tol)
    th_train = fam.thetas(args.ws_train, 1)
    th_test = fam.thetas(args.ws_test, 2)
    N = len(th_train)
    Sx, Sy = np.empty((N, n)), np.empty((N, m))
    st, om, its = np.zeros(N, np.int64), np.zeros(N), np.zeros(N, np.int64)
    t0 = time.perf_counter()
    count = [0]

    def keep(i, x, y, info):
        Sx[i], Sy[i] = x, y
        st[i], om[i], its[i] = info.status, info.primal_weight, info.iterations
        count[0] += 1
        if count[0] % 1000 == 0:
            log(f"[bench] training solves: {count[0]}/{N} in {time.perf_counter() - t0:.0f}s")
    _, t_train, fst = family_solve(eng, fam, th_train, tparams, on_result=keep, lanes=1024, with_stats=True)
    log(f"[bench] training family engine: {fst.blocks} blocks, {fst.refills} refills, {fst.compactions} compactions, "
        f"{fst.handoffs} members handed to the single-LP engine ({fst.handoff_seconds:.1f}s), {fst.tails} FP64 tails, "
        f"host set-up {fst.prepare_seconds:.1f}s, host waiting for blocks {fst.wait_seconds:.1f}s")
    ok = st == 0
    train_iters = int(its.sum())
    log(f"[bench] {int(ok.sum())}/{N} training solves reached {args.ws_train_tol:g} in {t_train:.1f}s "
        f"({train_iters} member-iterations, median {np.median(its):.0f}); "
        f"{t_train / max(1, train_iters) * 1e6:.3f} us per member-iteration")
    t0 = time.perf_counter()
    pod_x = Pod(eng, Sx[ok], 256, log=lambda msg: log("[bench] " + msg))
    del Sx
    log(f"[bench] POD x rank {pod_x.rank} in {time.perf_counter() - t0:.1f}s")
    pod_y = Pod(eng, Sy[ok], 256, log=lambda msg: log("[bench] " + msg))
    del Sy
    t_pod = time.perf_counter() - t0
    omega = float(np.median(om[ok]))
    log(f"[bench] POD y rank {pod_y.rank}; POD {t_pod:.1f}s; median primal weight {omega:.4g}")
    ws = WarmStart(eng, pod_x, pod_y, hidden=512, max_rows=256, primal_weight=omega, out_width=512)
    t0 = time.perf_counter()
    losses = ws.fit(th_train[ok], epochs=args.ws_epochs, batch=256, lr=1e-3, seed=3)
    t_fit = time.perf_counter() - t0
    rows, ms, flops = ws.net.train_stats()
    log(f"[bench] MLP loss {losses[0]:.3e} -> {losses[-1]:.3e} in {t_fit:.1f}s "
        f"({flops / (ms * 1e-3) / 1e12:.2f} TFLOP/s in the training GEMMs)")
    T = len(th_test)
    cold_x, cold_y = np.empty((T, n)), np.empty((T, m))
    ci, cs = np.zeros(T, np.int64), np.zeros(T, np.int64)
    wi, wsts = np.zeros(T, np.int64), np.zeros(T, np.int64)

    t0 = time.perf_counter()
    cnt = [0]

    def keep_cold(i, x, y, info):
        cold_x[i], cold_y[i], ci[i], cs[i] = x, y, info.iterations, info.status
        cnt[0] += 1
        if cnt[0] % 100 == 0:
            log(f"[bench] held-out cold: {cnt[0]}/{T} in {time.perf_counter() - t0:.0f}s")
    _, t_cold = family_solve(eng, fam, th_test, params, on_result=keep_cold)
    log(f"[bench] held-out cold: {T} LPs to 1e-6 in {t_cold:.1f}s, iterations median {np.median(ci):.0f}")
    # How far the rank-256 bases can represent the held-out solutions at all:
    # the projection error, relative to the distance from the mean.
    dx = np.linalg.norm(cold_x - pod_x.mean, axis=1)
    dy = np.linalg.norm(cold_y - pod_y.mean, axis=1)
    px_err = np.linalg.norm(cold_x - pod_x.decode(pod_x.encode(cold_x)), axis=1) / np.maximum(dx, 1e-30)
    py_err = np.linalg.norm(cold_y - pod_y.decode(pod_y.encode(cold_y)), axis=1) / np.maximum(dy, 1e-30)
    log(f"[bench] held-out POD projection error / distance from the mean: x {np.median(px_err):.3f}, "
        f"y {np.median(py_err):.3f}")
    # Start points: the network's coefficients decoded on the device, where the
    # family engine reads them (host copies only for the error figures).
    t0 = time.perf_counter()
    X0, Y0 = ws.predict_batch(th_test, to_host=True)
    t_pred = time.perf_counter() - t0
    ex = np.linalg.norm(X0 - cold_x, axis=1) / np.maximum(dx, 1e-30)
    ey = np.linalg.norm(Y0 - cold_y, axis=1) / np.maximum(dy, 1e-30)
    del X0, Y0
    log(f"[bench] held-out prediction error / distance from the mean: x {np.median(ex):.3f}, y {np.median(ey):.3f}")

    t0 = time.perf_counter()
    cnt[0] = 0

    def keep_warm(i, x, y, info):
        wi[i], wsts[i] = info.iterations, info.status
        cnt[0] += 1
        if cnt[0] % 100 == 0:
            log(f"[bench] held-out warm: {cnt[0]}/{T} in {time.perf_counter() - t0:.0f}s")
    _, t_warm = family_solve(eng, fam, th_test, params, primal_weight=omega, on_result=keep_warm, device_starts=ws)
    log(f"[bench] held-out warm: {T} LPs to 1e-6 in {t_warm:.1f}s, iterations median {np.median(wi):.0f}")
    # A warm solve that misses the tolerance falls back to its cold solve.
    fallback = [i for i in range(T) if wsts[i] != 0]
    t_fallback = t_cold / T * len(fallback)
    eff = wi.copy()
    for i in fallback:
        eff[i] += ci[i]
    plain = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6)
    samp = list(range(0, T, max(1, T // args.ws_plain_sample)))
    ts, si = 0.0, 0
    for i in samp:
        t0 = time.perf_counter()
        r = eng.solve(fam.instance(th_test[i]), plain)
        ts += time.perf_counter() - t0
        si += r.iterations
        if not prec:
            assert r.iterations == ci[i], (i, r.iterations, ci[i])
    # FP64 family: the sample's time per iteration times the exact counts of
    # all held-out LPs (the same in both engines); mixed family: the sample's
    # mean time per LP times the number of LPs.
    plain_total = ts / si * ci.sum() if not prec else ts / len(samp) * T
    ws.net.close()
    eng.close()
    nnz = fam.base.nnz
    # Algorithmic bytes of one member-iteration on the family engine (FP64):
    # primal x, anchor, cost read + x, xbar written, y gathered once; dual y,
    # anchor, row bounds read + y written, xbar gathered once; the matrix and
    # the column bounds shared by the 64 members of a wave.
    if prec:
        b_member = 4 * (6 * n + m) + 4 * (5 * m + n) + (2 * nnz * 8 + 4 * (m + n + 2)) / 64
    else:
        b_member = 8 * (5 * n + m) + 8 * (5 * m + n) + (2 * nnz * 12 + 4 * (m + n + 2) + 16 * n) / 64
    us_member = t_train / max(1, train_iters) * 1e6
    line = {
        "metric": "time-to-1e-6 relative KKT on held-out parametric LPs, learned warm start",
        "value": round(t_warm + t_pred + t_fallback, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f32 (iterations) + f64 (evaluations)" if prec else "f64", "data": "synthetic",
        "config": {"workload": f"parametric equality LP family m={m} n={n} ({nnz} nonzeros), theta in R^64, "
                               f"b +-10%, c +-5%, {N} training solves to {args.ws_train_tol:g}, {T} held-out to 1e-6",
                   "pod_rank_x": pod_x.rank, "pod_rank_y": pod_y.rank, "mlp": "64-512-512-512 bf16 GELU"},
        "training": {"solves": N, "optimal": int(ok.sum()), "seconds": round(t_train, 3),
                     "family_engine": {"blocks": int(fst.blocks), "refills": int(fst.refills),
                                       "compactions": int(fst.compactions),
                                       "handoffs": int(fst.handoffs), "fp64_tails": int(fst.tails),
                                       "handoff_s": round(fst.handoff_seconds, 2),
                                       "host_setup_s": round(fst.prepare_seconds, 2),
                                       "host_wait_s": round(fst.wait_seconds, 2)},
                     "member_iterations": train_iters, "iterations_median": float(np.median(its)),
                     "us_per_member_iteration": round(us_member, 4),
                     "algorithmic_bytes_per_member_iteration": b_member,
                     "achieved_GBs": round(b_member / us_member / 1e3, 1)},
        "held_out_iterations": {"cold_median": float(np.median(ci)), "warm_median": float(np.median(wi)),
                                "cold_mean": float(ci.mean()), "warm_mean": float(eff.mean()),
                                "saved_fraction_mean": float(1 - eff.sum() / ci.sum()),
                                "saved_fraction_median": float(1 - np.median(wi) / np.median(ci)),
                                "fallbacks": len(fallback)},
        "warm_primal_weight": omega,
        "held_out_pod_projection_error": {"x_median": float(np.median(px_err)), "y_median": float(np.median(py_err))},
        "held_out_prediction_error": {"x_median": float(np.median(ex)), "y_median": float(np.median(ey)),
                                      "note": "||prediction - solution|| / ||solution - POD mean||"},
        "mlp_loss": {"first_epoch": float(losses[0]), "last_epoch": float(losses[-1])},
        "time_s": {"plain_fp64_one_at_a_time": round(plain_total, 3), "family_cold": round(t_cold, 3),
                   "family_warm": round(t_warm, 3), "prediction": round(t_pred, 3),
                   "training_solves": round(t_train, 3), "pod": round(t_pod, 3), "mlp_training": round(t_fit, 3),
                   "whole_workload": round(time.perf_counter() - t_all, 1)},
        "speedup_warm_family_vs_plain_fp64": round(plain_total / (t_warm + t_pred + t_fallback), 2),
        "speedup_cold_family_vs_plain_fp64": round(plain_total / t_cold, 2),
        "plain_sample": (f"{len(samp)} held-out LPs solved one at a time by the single-LP engine in {ts:.2f}s "
                         f"({si} iterations, each equal to its family count); scaled by all {T} held-out counts")
                        if not prec else
                        (f"{len(samp)} held-out LPs solved one at a time by the FP64 single-LP engine in {ts:.2f}s; "
                         f"mean time per LP x {T}"),
        "mlp_train_tflops": round(flops / (ms * 1e-3) / 1e12, 3),
    }
    print(json.dumps(line), flush=True)


def family_workload(args):
    """The parametric-family engine alone on the 50,000 x 120,000 family:
    `fam_lanes` members (one per lane) iterated for `steps` iterations each
    (an iteration cap, tolerance 1e-12 so that none stops earlier): time per
    member-iteration and algorithmic bandwidth, the workload the family
    kernels' rocprofv3 summaries are taken on."""
    import numpy as np
    from nirnaya_amd.engine import Engine, MODE_PDLP, default_params
    from nirnaya_amd.warmstart import ParametricFamily, family_solve
    m, n = args.ws_m, args.ws_n
    fam = ParametricFamily(20264, m, n, 5)
    eng = Engine(0)
    if args.calibrate:
        ms = eng.calibrate_copy(256 << 20)
        log(f"[bench] calibration copy of 256 MiB: {ms:.3f} ms")
    th = fam.thetas(args.fam_lanes, 1)
    p = default_params(mode=MODE_PDLP, precision=1 if args.precision == "mixed" else 0, max_iters=args.steps,
                       tol=1e-12)
    its = []
    _, dt = family_solve(eng, fam, th, p, lanes=args.fam_lanes, on_result=lambda i, x, y, info: its.append(info.iterations))
    eng.close()
    total = int(sum(its))
    nnz = fam.base.nnz
    if args.precision == "mixed":
        # FP32 phase: primal reads the delta, re-based cost and bounds, writes
        # the delta and the reflected point, gathers the dual delta; dual
        # reads the delta, re-based row bounds and -Y / sigma, writes the
        # delta, gathers the reflected point; FP32 matrix values shared.
        b_member = 4 * (6 * n + m) + 4 * (5 * m + n) + (2 * nnz * 8 + 4 * (m + n + 2)) / 64
    else:
        b_member = 8 * (5 * n + m) + 8 * (5 * m + n) + (2 * nnz * 12 + 4 * (m + n + 2) + 16 * n) / 64
    us = dt / max(1, total) * 1e6
    mixed = args.precision == "mixed"
    line = {"metric": "parametric-family PDHG member-iterations/s (family engine, "
                      + ("mixed precision, FP32 phase)" if mixed else "FP64)"),
            "value": round(total / dt, 1), "unit": "member-iterations/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32 (iterations) + f64 (evaluations)" if mixed else "f64", "data": "synthetic",
            "config": {"workload": f"parametric equality LP family m={m} n={n} ({nnz} nonzeros), "
                                   f"{args.fam_lanes} members x {args.steps} iterations"},
            "us_per_member_iteration": round(us, 4), "algorithmic_bytes_per_member_iteration": b_member,
            "achieved_GBs": round(b_member / us / 1e3, 1),
            "note": "includes member set-up on the host and the start-point evaluations"}
    print(json.dumps(line), flush=True)


def dense_row_lp(dense_len, m=200_000, n=600_000, seed=11):
    """Equality LP (x in [0, 10], feasible by construction) with 3 nonzeros
    per column at uniform rows plus one row (row 3) of `dense_len` nonzeros
    at distinct random columns; N(0,1) values."""
    import numpy as np
    import scipy.sparse as sp
    from nirnaya_amd.engine import LP
    rng = np.random.default_rng(seed)
    ri = rng.integers(0, m, 3 * n)
    ci = np.repeat(np.arange(n), 3)
    if dense_len:
        ri = np.concatenate([ri, np.full(dense_len, 3)])
        ci = np.concatenate([ci, rng.choice(n, dense_len, replace=False)])
    A = sp.csc_matrix((rng.standard_normal(ri.size), (ri, ci)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    b = A @ rng.uniform(0, 1, n)
    return LP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float64),
              rng.uniform(0, 1, n), np.zeros(n), np.full(n, 10.0), b.copy(), b.copy())


def dense_row_workload(args):
    """Iteration time on LPs with one dense row (200,000 x 600,000, 3
    nonzeros per column, one row of 0 / 10,000 / 100,000 / 600,000): reference
    mode sums that row sequentially on one workgroup, as the reference's loop
    does; pdlp mode splits it over workgroups in its chunked order. Each mode
    runs `warmup` then `steps` iterations with the stop rule off (pdlp mode
    still evaluates, restarts and tests certificates every 64 iterations);
    the line reports pdlp mode's iterations/s on the longest row and both
    modes' time per iteration for every row length."""
    from nirnaya_amd.engine import Engine, MODE_PDLP, MODE_REFERENCE, default_params
    eng = Engine(0)
    rows = {}
    for d in args.dense_len:
        lp = dense_row_lp(d)
        per = {}
        for name, mode in (("reference", MODE_REFERENCE), ("pdlp", MODE_PDLP)):
            p = default_params(mode=mode, max_iters=args.warmup + args.steps + 1, check_termination=0, tol=1e-300)
            t0 = time.perf_counter()
            s = eng.session(lp, p)
            per[name + "_setup_s"] = round(time.perf_counter() - t0, 3)
            try:
                s.run(args.warmup)
                t0 = time.perf_counter()
                s.run(args.steps)
                s.result()
                per[name + "_us_per_iteration"] = round((time.perf_counter() - t0) / args.steps * 1e6, 1)
            finally:
                s.close()
        rows[str(d)] = per
        log(f"[bench] dense row of {d}: {per}")
    eng.close()
    longest = rows[str(args.dense_len[-1])]
    print(json.dumps({
        "metric": "PDHG iterations/s on an LP with one dense row (pdlp mode, FP64)",
        "value": round(1e6 / longest["pdlp_us_per_iteration"], 1), "unit": "iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"equality LP m=200000 n=600000, 3 nonzeros per column + one row of "
                               f"{args.dense_len[-1]} nonzeros, seed 11"},
        "per_dense_row_length": rows,
        "note": "wall-clock over the timed iterations; set-up (layout autotuning, upload, scaling, step sizes) "
                "timed apart; reference mode keeps the reference's sequential order in the dense row"}), flush=True)


def gemm_workload(args):
    """The warm start's bf16 MFMA GEMM through nir_mlp_gemm_check (FP32
    inputs rounded to bf16 on the device, C = A B^T in FP32) at a large shape
    (4096^3: the 128 x 128-tile kernel) and at the network's training shape
    (256 x 512 x 512: the 64 x 64-tile kernel), three calls each. The kernel
    durations come from a kernel trace of this run (the call also uploads,
    converts and downloads); the line reports the checks' errors."""
    import numpy as np
    from nirnaya_amd.engine import Engine
    eng = Engine(0)
    rng = np.random.default_rng(0)
    shapes = {}
    for (M, N, K) in ((256, 512, 512), (4096, 4096, 4096)):
        A = rng.standard_normal((M, K)).astype(np.float32)
        B = rng.standard_normal((N, K)).astype(np.float32)
        for _ in range(3):
            C = eng.gemm_check(A, B)
        rows = slice(0, 8)
        err = float(np.abs(C[rows] - A[rows].astype(np.float64) @ B.T.astype(np.float64)).max())
        shapes[f"{M}x{N}x{K}"] = {"flop": 2.0 * M * N * K, "max_abs_diff_first_rows_vs_unrounded_fp64": err}
        log(f"[bench] gemm {M}x{N}x{K}: {shapes[f'{M}x{N}x{K}']}")
    eng.close()
    print(json.dumps({"metric": "bf16 MFMA GEMM (kernel durations from a kernel trace of this run)",
                      "value": None, "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "bf16",
                      "data": "synthetic", "config": {"workload": "C = A B^T, N(0,1) inputs", "shapes": shapes}}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ttt", action="store_true", help="skip the time-to-1e-6 solves (pdlp mode, FP64 and mixed)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 512 planning LPs (batched engine vs one LP at a time, all measured)")
    ap.add_argument("--plain-cap", type=float, default=100.0,
                    help="seconds of iteration time for the unrestarted reference form toward 1e-6 (0 = skip)")
    ap.add_argument("--tile", type=int, default=-1, help="LDS tile (-1: autotune tile, rows and segments)")
    ap.add_argument("--seg", type=int, default=0, help="column-segment width of A xbar when not autotuning")
    ap.add_argument("--seg-t", type=int, default=-1, help="column-segment width of A^T y (-1: as --seg)")
    ap.add_argument("--fused", type=int, default=0, help="walk all segments in one launch when not autotuning")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run the 256 MiB copy kernel (reference for rocprofv3 byte counters)")
    ap.add_argument("--workload", default="large-lp",
                    choices=["large-lp", "small-lp", "mixed", "pdlp-fp64", "planning512", "warmstart", "family",
                             "gemm", "dense-row"],
                    help="large-lp: the headline line (default); small-lp: the reference's CPU-runnable "
                         "2,000 x 5,000 case; mixed: mixed precision on the large instance; "
                         "planning512: the batched workload; warmstart: the learned warm start on a parametric "
                         "family")
    ap.add_argument("--dense-len", type=lambda t: [int(v) for v in t.split(",")], default=[0, 10000, 100000, 600000],
                    help="dense-row: lengths of the dense row, comma-separated (the last one gives the value)")
    ap.add_argument("--batch-only", action="store_true",
                    help="planning512: skip the one-at-a-time sample (batch solves only)")
    ap.add_argument("--ws-m", type=int, default=50000)
    ap.add_argument("--ws-n", type=int, default=120000)
    ap.add_argument("--ws-train", type=int, default=20000)
    ap.add_argument("--ws-train-tol", type=float, default=1e-4,
                    help="relative KKT of the training solves (the held-out solves run to 1e-6)")
    ap.add_argument("--ws-test", type=int, default=1000)
    ap.add_argument("--ws-epochs", type=int, default=300)
    ap.add_argument("--fam-lanes", type=int, default=1024, help="family: members iterated together")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="family and warmstart workloads: the family engine's precision")
    ap.add_argument("--ws-plain-sample", type=int, default=32,
                    help="held-out LPs solved one at a time by the single-LP engine for the plain baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is design-only here.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "planning512":
        planning512(args)
        return
    if args.workload == "warmstart":
        warmstart_workload(args)
        return
    if args.workload == "mixed":
        mixed_workload(args)
        return
    if args.workload == "family":
        family_workload(args)
        return
    if args.workload == "gemm":
        gemm_workload(args)
        return
    if args.workload == "dense-row":
        dense_row_workload(args)
        return
    if args.workload == "pdlp-fp64":
        pdlp_fp64_workload(args)
        return
    if args.workload == "small-lp":
        small_lp_workload(args)
        return

    from nirnaya_amd.engine import Engine, MODE_REFERENCE, default_params, gen_mixed_lp

    t0 = time.time()
    lp = gen_mixed_lp(SEED, 400_000, 1_200_000, 6.0, 0.01, 1e3)
    log(f"[bench] instance m={lp.m} n={lp.n} nnz={lp.nnz} generated in {time.time() - t0:.1f}s")

    eng = Engine(0)
    calibration = None
    if args.calibrate:
        ms = eng.calibrate_copy(256 << 20)
        log(f"[bench] calibration copy of 256 MiB: {ms:.3f} ms ({2 * 256 * 1.048576 / ms:.0f} GB/s read+write)")
        # Random FP64 gathers at the instance's nonzero count, by operand size.
        st_us = eng.calibrate_gather(lp.nnz, 1, with_values=2)
        calibration = {"copy_256MiB_ms": ms, "gather_rate_G_per_s": {},
                       "index_value_stream_us": st_us, "index_value_stream_GBs": lp.nnz * 12 / st_us / 1e3}
        log(f"[bench] index + value streams alone: {st_us:.2f} us ({lp.nnz * 12 / st_us / 1e3:.0f} GB/s)")
        for length in (1_200_000, 400_000, 240_000, 100_000, 4_000):
            g = lp.nnz / eng.calibrate_gather(lp.nnz, length, with_values=0) / 1e3
            gv = lp.nnz / eng.calibrate_gather(lp.nnz, length, with_values=1) / 1e3
            calibration["gather_rate_G_per_s"][str(length * 8)] = {"gather": round(g, 1), "value_x_gather": round(gv, 1)}
            log(f"[bench] random FP64 gathers, operand {length * 8 / 1e6:.2f} MB: {g:.0f} G/s alone, "
                f"{gv:.0f} G/s with the value stream")
    params = default_params(mode=MODE_REFERENCE, max_iters=args.warmup + 2 * args.steps + 6064, check_termination=0,
                            tile_nnz=args.tile, tile_rows=args.tile_rows, seg_width=args.seg,
                            seg_width_t=args.seg_t, seg_fused=args.fused)
    sess = eng.session(lp, params)
    shape = sess.shape()
    log(f"[bench] session ready (upload, tile autotune, power iteration) in {sess.result().info.setup_seconds:.2f}s; "
        f"shape A x: tile {shape[0]} rows {shape[1]} segments {shape[2]} fused {shape[3]}, A^T y: tile {shape[4]} "
        f"rows {shape[5]} segments {shape[6]} fused {shape[7]}")
    sess.run(args.warmup)
    log(f"[bench] warmup {args.warmup} iterations done")
    # Two sweeps of `steps` iterations: the production path (graph replays)
    # for the throughput, then the same kernels with events around every
    # dispatch for the per-kernel durations.
    t = sess.time(args.steps)
    # Per-kernel durations (the roofline) from at least 1,000 iterations of
    # the per-dispatch sweep, so that a short run still averages over many
    # launches; `value` is the production sweep of exactly `steps`.
    kt = t if args.steps >= 1000 else sess.time(1000)
    k_iters = max(args.steps, 1000)
    # A longer production sweep beside the timed one, so that a short --steps
    # run still shows the steady rate it samples (reported, not `value`).
    long_steps = 2000
    lt = t if args.steps >= long_steps else sess.time(long_steps)
    long_it_s = max(args.steps, long_steps) / (lt.total_ms / 1e3)
    log(f"[bench] production sweep of {max(args.steps, long_steps)} iterations: {long_it_s:.0f} iterations/s")
    log(f"[bench] timed {args.steps} iterations: {t.total_ms:.1f} ms; per-dispatch sweep over {k_iters}: A xbar "
        f"{kt.spmv_ms / k_iters * 1e3:.1f} us in {kt.spmv_passes} passes, A^T y {kt.spmtv_ms / k_iters * 1e3:.1f} us "
        f"in {kt.spmtv_passes} passes")
    # Access floors: the same layouts' value/index streams and operand gathers
    # with nothing else (no row sums, no epilogue).
    floor_a, floor_t = sess.gather_floor(0), sess.gather_floor(1)
    log(f"[bench] access floor: A xbar {floor_a:.1f} us, A^T y {floor_t:.1f} us")
    sess.close()
    eng.close()

    it_s = args.steps / (t.total_ms / 1e3)
    spmv_avg_s = kt.spmv_ms / k_iters / 1e3
    spmtv_avg_s = kt.spmtv_ms / k_iters / 1e3
    achieved = t.spmv_bytes / spmv_avg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "spmv_cfg1_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        # Measured traffic of the layout this run uses (the autotuner may pick
        # another one on another box; then there is no measurement to quote).
        for e in tr.get("layouts", []):
            if e.get("tile") == shape[0] and e.get("segments") == shape[2] and not shape[3]:
                traffic = e.get("hbm_bytes_per_launch")

    ttt = None
    if not args.no_ttt:
        ttt = time_to_tolerance(eng_factory=lambda: Engine(0), lp=lp, plain_cap_s=args.plain_cap)

    batched = None
    if not args.no_batched:
        batched = batched_leg()

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] timing the reference CPU loop on a bounded sample ...")
        cpu = cpu_baseline(lp)

    out = {
        "metric": METRIC,
        "value": round(it_s, 2),
        "unit": "PDHG iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t.total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single large LP, plain FP64 PDHG (reference form): m=400000 rows "
                               "(200000 equality, 200000 <=), n=1200000 columns, degree 2+Poisson(6), "
                               f"nnz={lp.nnz}, 1% columns x1e3, seed {SEED}",
                   "m": lp.m, "n": lp.n, "nnz": lp.nnz},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_dual_ref (A xbar + fused dual step); one launch = its column-segment passes",
                     "bytes_per_launch": t.spmv_bytes, "avg_launch_us": spmv_avg_s * 1e6,
                     "launches_averaged": k_iters,
                     "dispatches_per_launch": t.spmv_passes,
                     "avg_dispatch_us": spmv_avg_s * 1e6 / t.spmv_passes,
                     "shape": {"tile": shape[0], "rows": shape[1], "segments": shape[2], "fused": shape[3]},
                     "access_floor_us": round(floor_a, 2),
                     "frac_of_access_floor": round(floor_a / (spmv_avg_s * 1e6), 4)},
        "kernels": {"spmtv_primal_avg_us": spmtv_avg_s * 1e6, "spmtv_dispatches_per_launch": t.spmtv_passes,
                    "spmtv_access_floor_us": round(floor_t, 2),
                    "spmtv_bytes_per_launch": t.spmtv_bytes,
                    "spmtv_gbs": t.spmtv_bytes / spmtv_avg_s / 1e9,
                    "iteration_bytes": t.iter_bytes,
                    "iteration_gbs": t.iter_bytes * args.steps / (t.total_ms / 1e3) / 1e9},
        "production_sweep": {"iterations": max(args.steps, long_steps), "iterations_per_s": round(long_it_s, 2)},
        "time_to_1e-6": ttt,
        "batched_512_planning_lps": batched,
        "calibration": calibration,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
