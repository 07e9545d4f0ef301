"""Benchmark: LocVolCalib options calibrated per second on one MI355X.

Workload (the plain method at 256x256x64): main(128, 256, 256, 64, s0=0.03,
t=5, alpha=0.2, nu=0.6, beta=0.5) -> 128 prices.  One "step" is one call of
the entry point over the 128 options (setup, 63 ADI time steps, price
gather), inputs resident on the device (the entry takes scalars only), the
result left in device memory.

Reported:
  value       options/s of the FP32 + FP64-residual-refinement form
              (futhark_entry_main, the headline form);
  fp64        options/s of the FP64 form (futhark_entry_main_f64) and the
              ratio; p50/p99 per-call latency of both;
  roofline    the dominant kernel (the x sweep of the FP32+refine form):
              algorithmic bytes per launch (read R + write Y, 2*N*4 bytes per
              option, 128 options) / its mean launch duration, against 8 TB/s.
              The duration: the event-free timed calls' time per call times
              the kernel's share of the GPU time per call, the share from HIP
              events recorded around every launch on the library's stream in
              a second pass of the same calls (a call is its kernels back to
              back, with no gaps between them, so the event-free call time is
              their sum; the events themselves lengthen each launch);
              `valu` beside it from the committed rocprofv3 counters;
  accuracy    per workload, max abs / relative price error against the FP64
              oracle (the C restatement, Thomas form);
  workloads   main(16, 32, 256, 256) (the CPU-sized plain method),
              main(256, 256, 256, 128) (the precision form), the batch of 512
              uneven problems, the served Poisson stream at six rates and at
              saturation;
  cpu_baseline  the oracle (C restatement of LocVolCalib.fut, scan-form
              tridiagonal solve as the reference) on the host cores, OpenMP
              over options, on a time-bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "futhark-benchmarks_amd"))
sys.path.insert(0, ROOT)

OUTER, NX, NY, NT = 128, 256, 256, 64
PARAMS = (0.03, 5.0, 0.2, 0.6, 0.5)
HBM_PEAK_GBS = 8000.0


def log(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def time_calls(ctx, raw, free, steps, warmup):
    for _ in range(warmup):
        free(raw(OUTER, NX, NY, NT, *PARAMS))
    ctx.sync()
    lat = []
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        h = raw(OUTER, NX, NY, NT, *PARAMS)
        ctx.sync()
        lat.append(time.perf_counter() - t0)
        free(h)
    ctx.sync()
    total = time.perf_counter() - t_start
    return total, lat


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` measured by rocprofv3 (FETCH_SIZE +
    WRITE_SIZE, scaled by a device copy of known size) for this workload,
    from the committed profile summary; None when there is none."""
    path = os.path.join(ROOT, "profiles",
                        "lvc_sweeps_main_128x256x256x64_fp32refine.traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["hbm_bytes_per_launch"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


def profiled_valu(kernel):
    """VALU issue fraction of `kernel` from the committed rocprofv3 counters
    (SQ_INSTS_VALU over the SIMDs' issue slots in its measured duration)."""
    path = os.path.join(ROOT, "profiles",
                        "lvc_sweeps_main_128x256x256x64_fp32refine.traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["valu"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


def accuracy(got, ref):
    """Max abs and relative (per price, and to the largest price) error of
    `got` against the oracle's `ref`."""
    import numpy as np
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    fin = np.isfinite(ref)
    d = np.abs(got[fin] - ref[fin])
    nz = ref[fin] != 0
    return {"max_abs": float(d.max()) if d.size else 0.0,
            "max_rel": float((d[nz] / np.abs(ref[fin][nz])).max()) if nz.any() else 0.0,
            "max_rel_to_max_price": float(d.max() / np.abs(ref[fin]).max()) if d.size else 0.0,
            "prices": int(fin.sum())}


def oracle_main(outer, nx, ny, nt, params, f32_entry, nthreads=16):
    """The FP64 oracle's prices of main's options: with `f32_entry` in the
    f32 entry's form (f32 parameters and strikes 0.001f*i, initGrid's scalars
    in f32), else the FP64 entry's."""
    import numpy as np
    from oracle import oracle
    idx = np.arange(outer)
    if f32_entry:
        p = [float(np.float32(v)) for v in params]
        K = (np.float32(0.001) * idx.astype(np.float32)).astype(np.float64)
    else:
        p = list(params)
        K = 0.001 * idx.astype(np.float64)
    return oracle.batch([nx] * outer, [ny] * outer, [nt] * outer, [p[0]] * outer, K,
                        [p[1]] * outer, [p[2]] * outer, [p[3]] * outer, [p[4]] * outer,
                        dtype=np.float64, thomas=True, nthreads=nthreads, grid32=f32_entry)


def bench_main_workload(lvc, outer, nx, ny, nt, params=PARAMS, reps=10, check=True):
    """main and main_f64 on one workload: ms per call, options/s, and their
    errors against the FP64 oracle."""
    import numpy as np
    out = {"workload": f"main({outer}, {nx}, {ny}, {nt})"}
    res = {}
    with lvc.LocVolCalib(device=0) as ctx:
        for key, raw, free, get in (("fp32refine", ctx.main_raw, ctx.free_f32, ctx.main),
                                    ("fp64", ctx.main_f64_raw, ctx.free_f64, ctx.main_f64)):
            for _ in range(2):
                free(raw(outer, nx, ny, nt, *params))
            ctx.sync()
            lat = []
            for _ in range(reps):
                t0 = time.perf_counter()
                h = raw(outer, nx, ny, nt, *params)
                ctx.sync()
                lat.append(time.perf_counter() - t0)
                free(h)
            res[key] = get(outer, nx, ny, nt, *params)
            out[key] = {"ms_per_call": round(1e3 * float(np.median(lat)), 4),
                        "options_per_s": round(outer / float(np.median(lat)), 1)}
    out["ratio_fp32refine_vs_fp64"] = round(out["fp32refine"]["options_per_s"] /
                                            out["fp64"]["options_per_s"], 3)
    log(f"{out['workload']}: {out['fp32refine']['ms_per_call']} / {out['fp64']['ms_per_call']} ms")
    if check:
        log(f"{out['workload']}: FP64 oracle")
        out["fp32refine"]["accuracy_vs_fp64_oracle"] = accuracy(
            res["fp32refine"], oracle_main(outer, nx, ny, nt, params, True))
        out["fp64"]["accuracy_vs_fp64_oracle"] = accuracy(
            res["fp64"], oracle_main(outer, nx, ny, nt, params, False))
    return out


def pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v) * 1e3, q))


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s=12.0, single_budget_s=4.0):
    """The oracle (f32, the reference's scan-form tridagPar) on the host:
    OpenMP over options on all the cores this process may use (the analogue
    of futhark multicore), then one thread (futhark c), each on a
    time-bounded sample of the headline workload's options."""
    import numpy as np
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))

    def sample(nthreads, budget):
        chunk = 2 * nthreads
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget:
            idx = np.arange(done, done + chunk) % OUTER
            K = (np.float32(0.001) * idx.astype(np.float32)).astype(np.float32)
            n = len(idx)
            oracle.batch([NX] * n, [NY] * n, [NT] * n, [PARAMS[0]] * n, K, [PARAMS[1]] * n,
                         [PARAMS[2]] * n, [PARAMS[3]] * n, [PARAMS[4]] * n,
                         dtype=np.float32, thomas=False, nthreads=nthreads)
            done += n
        return done, time.perf_counter() - t0

    done, el = sample(threads, budget_s)
    done1, el1 = sample(1, single_budget_s)
    return {"value": round(done / el, 3), "unit": "options/s", "cores": threads,
            "kind": "port",
            "sample": f"{done} options of main({OUTER},{NX},{NY},{NT}) (f32, scan-form "
                      f"tridagPar as LocVolCalib.fut:68-124), OpenMP over options, {el:.1f} s",
            "single_thread": {"value": round(done1 / el1, 3), "unit": "options/s",
                              "sample": f"{done1} options, 1 thread, {el1:.1f} s"},
            "host": host_cpu_model()}


def uneven_problems(n, seed, sizes=(32, 64, 128, 256), nt=(32, 256)):
    """Seeded uneven problems: numX, numY from `sizes`, numT uniform in `nt`,
    s0 U[0.02,0.05], K U[0,0.1], t U[1,10], alpha U[0.1,0.3], nu U[0.4,0.8],
    beta U[0.3,0.7]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return dict(numX=rng.choice(sizes, n).astype(np.int32),
                numY=rng.choice(sizes, n).astype(np.int32),
                numT=rng.integers(nt[0], nt[1] + 1, n).astype(np.int32),
                s0=rng.uniform(0.02, 0.05, n), strike=rng.uniform(0.0, 0.1, n),
                t=rng.uniform(1, 10, n), alpha=rng.uniform(0.1, 0.3, n),
                nu=rng.uniform(0.4, 0.8, n), beta=rng.uniform(0.3, 0.7, n))


def bench_batched(lvc, reps=3, n=512, seed=42, check=True):
    """512 uneven problems in one batch call (FP32 + refinement), and its
    errors against the FP64 oracle."""
    import numpy as np
    import numpy as np
    pr = uneven_problems(n, seed)
    times = {}
    for graph in (0, 1):
        with lvc.LocVolCalib(device=0, tuning={"lvc.graph": graph}) as ctx:
            got = ctx.main_batch(**pr)             # warm-up (tuning; the graph kept)
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.main_batch(**pr)
            times[graph] = (time.perf_counter() - t0) / reps
    dt = times[0]
    elems = float((pr["numX"].astype(float) * pr["numY"] * (pr["numT"] - 1)).sum())
    segs = {}
    for nx, ny in zip(pr["numX"], pr["numY"]):
        segs[f"{nx}x{ny}"] = segs.get(f"{nx}x{ny}", 0) + 1
    out = {"problems": n, "seed": seed, "ms_per_batch": round(dt * 1e3, 3),
           "problems_per_s": round(n / dt, 2),
           "grid_point_steps_per_s": round(elems / dt, 1),
           "graph_replay": {"ms_per_batch": round(times[1] * 1e3, 3),
                            "problems_per_s": round(n / times[1], 2),
                            "how": "the step launches replayed from the context's kept "
                                   "HIP graph (lvc.graph), against direct launches"},
           "size_classes": len(segs)}
    log(f"batched: {out['ms_per_batch']} ms per batch")
    if check:
        log("batched: FP64 oracle")
        from oracle import oracle
        keys = ("numX", "numY", "numT", "s0", "strike", "t", "alpha", "nu", "beta")
        ref = oracle.batch(*[pr[k] for k in keys], dtype=np.float64, thomas=True,
                           nthreads=16, allow_nan=True)
        out["accuracy_vs_fp64_oracle"] = accuracy(got, ref)
    return out


def bench_ocean(lvc, n=57600, m=115, reps=20, seed=3):
    """The ocean-sim tridiagonal entries (misc/ocean-sim/tridiag-test.fut) on
    their `random input` shape, [57600][115] f32 x 4 (values uniform in
    [0, 1), as futhark's random inputs), inputs resident on the device:
    ms per call and algorithmic bytes (read a, b, c, y, write the result:
    5*n*m*4) per second against 8 TB/s."""
    import numpy as np
    rng = np.random.default_rng(seed)
    arrs = [rng.random((n, m), dtype=np.float32) for _ in range(4)]
    out = {"shape": [n, m], "seed": seed}
    with lvc.LocVolCalib(device=0) as ctx:
        hs = [ctx.upload_f32_2d(v) for v in arrs]
        try:
            for seq, key in ((True, "tridagNestedSeq"), (False, "tridagNested")):
                for _ in range(3):
                    ctx.free_f32_2d(ctx.tridag_nested_raw(*hs, seq=seq))
                ctx.sync()
                t0 = time.perf_counter()
                for _ in range(reps):
                    h = ctx.tridag_nested_raw(*hs, seq=seq)
                    ctx.sync()
                    ctx.free_f32_2d(h)
                dt = (time.perf_counter() - t0) / reps
                gbs = 5 * n * m * 4 / dt / 1e9
                out[key] = {"ms_per_call": round(dt * 1e3, 4), "achieved_GBs": round(gbs, 1),
                            "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4)}
        finally:
            for h in hs:
                ctx.free_f32_2d(h)
    return out


def bench_option_pricing(reps=10, cpu=True):
    """finpar/OptionPricing.fut's `main` on the reference's three datasets
    (tests/golden/OptionPricing-data: contracts 1, 2, 3), inputs resident on
    the device: ms per call (entry + sync) and Sobol paths per second; with
    `cpu`, the oracle's time on the same dataset (OpenMP, 16 threads) and the
    relative difference of the two results."""
    import numpy as np
    from locvolcalib import futhark_values as fv
    from locvolcalib import optionpricing as op
    data = os.path.join(ROOT, "tests", "golden", "OptionPricing-data")
    res = {}
    with op.OptionPricing(device=0) as ctx:
        for name in ("small", "medium", "large"):
            c, n, *arrays = op.main_arguments(fv.load(os.path.join(data, f"{name}.in")))
            hs = ctx.upload_arguments(arrays)
            try:
                for _ in range(3):
                    ctx.free_f32(ctx.main_raw(c, n, hs))
                ctx.sync()
                t0 = time.perf_counter()
                for _ in range(reps):
                    h = ctx.main_raw(c, n, hs)
                    ctx.sync()
                    ctx.free_f32(h)
                dt = (time.perf_counter() - t0) / reps
                h = ctx.main_raw(c, n, hs)
                got = ctx.values_f32(h)
                ctx.free_f32(h)
            finally:
                ctx.free_arguments(hs)
            r = {"contract": c, "paths": n, "ms_per_call": round(dt * 1e3, 4),
                 "paths_per_s": round(n / dt, 1)}
            if cpu:
                from oracle import oracle
                t0 = time.perf_counter()
                want = oracle.op_main(c, n, *arrays, nthreads=16)
                ct = time.perf_counter() - t0
                r["cpu_oracle"] = {"ms": round(ct * 1e3, 2), "paths_per_s": round(n / ct, 1),
                                   "cores": 16, "kind": "port"}
                r["max_rel_err_vs_oracle"] = float(np.max(np.abs(got - want) / np.abs(want)))
            res[name] = r
    return res


def bench_interest_calib(reps=5, cpu=True):
    """finpar/InterestCalib.fut's `main` on the reference's three datasets
    (tests/golden/InterestCalib-data), inputs resident on the device: ms per
    call (the entry returns its six scalars to the host, so a call ends
    synchronised), MCMC iterations and genome-swaption evaluations per
    second; the winner's log-likelihood beside the dataset's expected one
    (the trajectories differ in their accept decisions, see the tests).
    With `cpu`, the oracle's rate on a bounded sample: one population's
    log-likelihoods (pop genomes on every swaption, OpenMP, 16 threads)."""
    import numpy as np
    from locvolcalib import futhark_values as fv
    from locvolcalib import interestcalib as icl
    data = os.path.join(ROOT, "tests", "golden", "InterestCalib-data")
    # VALU instructions of one ic_eval_kernel launch on small (rocprofv3 PMC,
    # the profile beside): the issue fraction below divides it by the launch's
    # live duration
    counters = json.load(open(os.path.join(ROOT, "profiles",
                                           "ic_eval_kernel_interestcalib_small.counters.json")))
    res = {}
    with icl.InterestCalib(device=0, profiling=True) as ctx:
        for name in ("small", "medium", "large"):
            a = icl.main_arguments(fv.load(os.path.join(data, f"{name}.in")))
            expected = float(fv.load(os.path.join(data, f"{name}.out"))[5])
            hs = ctx.upload_inputs(a["swaptions"], a["herm_x"], a["herm_w"], a["dir_vct"])
            try:
                ctx.pause_profiling()
                out, h = ctx.main_raw(a["pop"], a["mcmc_conv"], hs)
                ctx.free_summary(h)
                t0 = time.perf_counter()
                for _ in range(reps):
                    out, h = ctx.main_raw(a["pop"], a["mcmc_conv"], hs)
                    ctx.free_summary(h)
                dt = (time.perf_counter() - t0) / reps
                # one more call with events around every launch (the report
                # accumulates: this call's share is the difference)
                rep0 = json.loads(ctx.report())
                ctx.unpause_profiling()
                _, h = ctx.main_raw(a["pop"], a["mcmc_conv"], hs)
                ctx.free_summary(h)
                rep1 = json.loads(ctx.report())
                ctx.pause_profiling()
                zero = {"count": 0, "total_us": 0.0}
                rep = {k: {"count": v["count"] - rep0.get(k, zero)["count"],
                           "total_us": v["total_us"] - rep0.get(k, zero)["total_us"]}
                       for k, v in rep1.items() if isinstance(v, dict)}
            finally:
                ctx.free_inputs(hs)
            nsw = len(a["swaptions"])
            evals = a["pop"] * nsw * (a["mcmc_conv"] + 1)
            gpu_us = sum(v["total_us"] for v in rep.values())
            ev = rep.get("ic_eval", {"count": 0, "total_us": 0.0})
            # the eval launches' share of the event-timed GPU time, applied to
            # the event-free call time (the launches run back to back)
            eval_us = (dt * 1e6 * ev["total_us"] / gpu_us / max(ev["count"], 1)
                       if gpu_us > 0 else 0.0)
            r = {"pop": a["pop"], "mcmc_conv": a["mcmc_conv"], "swaptions": nsw,
                 "ms_per_call": round(dt * 1e3, 3),
                 "iterations_per_s": round(a["mcmc_conv"] / dt, 1),
                 "evaluations_per_s": round(evals / dt, 1),
                 "winner_loglik": float(out[5]), "expected_loglik": expected,
                 "kernels_us_with_events": {k: round(v["total_us"] / max(v["count"], 1), 3)
                                            for k, v in rep.items()},
                 "eval_kernel_us": round(eval_us, 3)}
            if name == "small" and eval_us > 0:
                peak = 1024 * 2.4e9 / 4  # wave64 VALU instructions per second
                achieved = counters["SQ_INSTS_VALU_per_launch"] / (eval_us * 1e-6)
                r["roofline"] = {"bound": "valu", "kernel": "ic_eval_kernel",
                                 "achieved": round(achieved, 1), "peak": peak,
                                 "unit": "wave64 VALU instructions/s",
                                 "frac": round(achieved / peak, 4),
                                 "how": counters["how"]}
            if cpu:
                from oracle import oracle
                g = np.array([[np.float32(oracle.ic_sobol(5 * i + 1 + j, a["dir_vct"]))
                               for j in range(5)] for i in range(a["pop"])], np.float32)
                lo = np.array([1e-3, 1e-3, np.float32(1e-3) - np.float32(1), 1e-3, 1e-3],
                              np.float32)
                hi = np.array([np.float32(1) - np.float32(1e-3)] * 3 + [0.2, 0.2], np.float32)
                g = (g * (hi - lo) + lo).astype(np.float32)
                t0 = time.perf_counter()
                oracle.ic_logliks(g, a["swaptions"], a["herm_x"], a["herm_w"], nthreads=16)
                ct = time.perf_counter() - t0
                r["cpu_oracle"] = {"sample": f"the initial population's log-likelihoods "
                                             f"({a['pop']} genomes x {nsw} swaptions)",
                                   "evaluations_per_s": round(a["pop"] * nsw / ct, 1),
                                   "cores": 16, "kind": "port"}
            res[name] = r
    return res


def bench_serving(lvc, rate, duration_s=1.5, seed=7, max_slots=1024, budget=64 << 30):
    """Poisson arrivals at `rate` req/s (None: every request queued at time 0,
    the saturated server) of single problems (sizes and parameters as the
    batched set), continuous batching, FP32 + refinement.  Latency =
    completion - arrival (host clock); completions/s over the whole run."""
    import heapq
    import numpy as np
    rng = np.random.default_rng(seed)
    n = max(1, int((rate or 20000) * duration_s))
    arrivals = (np.cumsum(rng.exponential(1.0 / rate, n)) if rate
                else np.zeros(n))
    pr = uneven_problems(n, seed + 1)
    keys = ("numX", "numY", "numT", "s0", "strike", "t", "alpha", "nu", "beta")
    reqs = [tuple(int(pr[k][i]) if k.startswith("num") else float(pr[k][i]) for k in keys)
            for i in range(n)]
    with lvc.LocVolCalib(device=0) as ctx:
        srv = lvc.Server(ctx, max_slots=max_slots, mem_budget=budget)
        srv.submit(*reqs[0])
        srv.drain()
        ids, done_at = [None] * n, [None] * n
        # requests are admitted in submission order: those before `adm` are
        # admitted; `due` holds (earliest tick it can complete, index) for them
        due = []
        nxt, adm, tick, left = 0, 0, 0, n
        base = srv.stats()["completed"]
        t0 = time.perf_counter()
        t_log = t0
        while left:
            now = time.perf_counter() - t0
            if now + t0 - t_log > 10.0:
                t_log = now + t0
                log(f"serving at {rate or 'saturation'}: {n - left} of {n} done, tick {tick}")
            while nxt < n and arrivals[nxt] <= now:
                ids[nxt] = srv.submit(*reqs[nxt])
                nxt += 1
            if nxt == n - left:                # nothing outstanding: wait
                time.sleep(max(0.0, min(arrivals[nxt] - now, 0.001)))
                continue
            srv.step()
            st = srv.stats()
            admitted = st["completed"] - base + st["resident"]
            while adm < admitted:          # admitted by this tick: n_t - 1 ticks from it
                heapq.heappush(due, (tick + reqs[adm][2] - 2, adm))
                adm += 1
            # requests due by the previous tick: their prices were copied
            # behind that tick's launches, which the GPU has run (or is
            # finishing) while this tick was enqueued
            now = time.perf_counter() - t0
            while due and due[0][0] <= tick - 1:
                _, i = heapq.heappop(due)
                if srv.poll(ids[i]) is not None:
                    done_at[i] = now
                    left -= 1
                else:
                    heapq.heappush(due, (tick, i))
            tick += 1
        total = time.perf_counter() - t0
        srv.free()
    lat = (np.array(done_at) - arrivals) * 1e3
    log(f"serving at {rate or 'saturation'}: {n} requests, {n / total:.0f} completed/s")
    return {"rate_req_per_s": rate if rate else "saturated (all queued at t=0)",
            "requests": n, "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "options_per_s": round(n / total, 2), "wall_s": round(total, 3),
            "ticks": tick}


def bench_serving_sweep(lvc, rates):
    """The given rates, then doubling from the last one until completions/s
    falls below 95% of the arrival rate, then the saturated server."""
    runs = [bench_serving(lvc, float(r)) for r in rates]
    r = float(rates[-1])
    while runs[-1]["options_per_s"] >= 0.95 * r and r < 64e3:
        r *= 2
        runs.append(bench_serving(lvc, r))
    sat = bench_serving(lvc, None)
    keeps = [x["rate_req_per_s"] for x in runs if x["options_per_s"] >= 0.95 * x["rate_req_per_s"]]
    return {"runs": runs, "saturated": sat,
            "highest_rate_kept_up": max(keeps) if keeps else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batched and serving measurements")
    ap.add_argument("--serve-rates", default="100,250,500,1000,2500,5000",
                    help="Poisson arrival rates (req/s) of the serving runs")
    ap.add_argument("--only-fp32refine", action="store_true",
                    help="run only the FP32+refinement calls (for profiling)")
    ap.add_argument("--only-ocean", action="store_true",
                    help="run only the ocean-sim tridiagonal entries (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain device copy of 256 MiB (read) + 256 MiB "
                         "(write) three times, to scale HBM traffic counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(options shard across GPUs with no exchange); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import locvolcalib as lvc

    if args.copy_calibration:
        import torch
        a = torch.empty(64 << 20, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    if args.only_ocean:
        print(json.dumps({"ocean_tridiag": bench_ocean(lvc)}), flush=True)
        return

    with lvc.LocVolCalib(device=0, profiling=True) as ctx:
        # FP32 + refinement (headline): timed calls without per-launch events
        ctx.pause_profiling()
        total32, lat32 = time_calls(ctx, ctx.main_raw, ctx.free_f32, args.steps,
                                    args.warmup)
        # the same calls again with HIP events around every launch on the
        # library's stream: per-kernel durations for the roofline (the
        # report accumulates; nothing was recorded while paused, so the next
        # report holds these calls alone)
        ctx.unpause_profiling()
        total32p, _ = time_calls(ctx, ctx.main_raw, ctx.free_f32, args.steps, 0)
        rep32 = json.loads(ctx.report())
        ctx.pause_profiling()
        if args.only_fp32refine:
            print(json.dumps({"ms_per_step": round(1e3 * total32 / args.steps, 4),
                              "kernels": rep32}), flush=True)
            return
        # FP64
        total64, lat64 = time_calls(ctx, ctx.main_f64_raw, ctx.free_f64, args.steps,
                                    args.warmup)
        p32 = ctx.main(OUTER, NX, NY, NT, *PARAMS).astype(np.float64)
        p64 = ctx.main_f64(OUTER, NX, NY, NT, *PARAMS)

    v32 = OUTER * args.steps / total32
    v64 = OUTER * args.steps / total64
    xs = rep32.get("lvc_xsweep", {"count": 0, "total_us": 0.0})
    ys = rep32.get("lvc_ysweep", {"count": 0, "total_us": 0.0})
    dom, other = (xs, ys) if xs["total_us"] >= ys["total_us"] else (ys, xs)
    dom_name = "lvc_xsweep" if dom is xs else "lvc_ysweep"
    avg_us_events = dom["total_us"] / max(dom["count"], 1)
    # the kernel's share of the event-timed GPU time, applied to the
    # event-free call time (the calls' kernels run back to back)
    gpu_us = sum(v["total_us"] for v in rep32.values())
    share = dom["total_us"] / gpu_us if gpu_us > 0 else 0.0
    launches_per_call = dom["count"] / max(args.steps, 1)
    avg_us = (1e6 * total32 / args.steps) * share / max(launches_per_call, 1e-9)
    bytes_launch = OUTER * 2 * NX * NY * 4
    achieved = bytes_launch / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0
    rel = float(np.max(np.abs(p32 - p64) / np.abs(p64)))

    out = {
        "metric": "options calibrated/sec @256x256x64 FP32+refine vs FP64 plain; p50/p99 latency",
        "value": round(v32, 2),
        "unit": "options/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total32 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "LocVolCalib main(128, 256, 256, 64) plain method, "
                               "s0=0.03 t=5 alpha=0.2 nu=0.6 beta=0.5",
                   "outer": OUTER, "numX": NX, "numY": NY, "numT": NT,
                   "precision": "FP32 state and sweeps + one FP64-residual refinement "
                                "per implicit solve; FP64 setup and price gather"},
        "p50_ms": round(pct(lat32, 50), 4),
        "p99_ms": round(pct(lat32, 99), 4),
        "fp64": {"value": round(v64, 2), "unit": "options/s",
                 "ms_per_step": round(1e3 * total64 / args.steps, 4),
                 "p50_ms": round(pct(lat64, 50), 4), "p99_ms": round(pct(lat64, 99), 4)},
        "ratio_fp32refine_vs_fp64": round(v32 / v64, 3),
        "max_rel_err_fp32refine_vs_fp64": rel,
        "kernels_us_per_launch_with_events": {k: round(v["total_us"] / max(v["count"], 1), 3)
                                              for k, v in rep32.items()},
        "ms_per_step_with_events": round(1e3 * total32p / args.steps, 4),
        "roofline": {"bound": "hbm", "kernel": dom_name,
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(dom_name),
                     "bytes_per_launch": bytes_launch, "avg_launch_us": round(avg_us, 3),
                     "avg_launch_us_how": "event-free time per call x the kernel's share of "
                                          "the event-timed GPU time / its launches per call",
                     "avg_launch_us_with_events": round(avg_us_events, 3),
                     "valu": profiled_valu(dom_name)},
    }
    log(f"headline: {out['ms_per_step']} ms per call (FP32+refine), "
        f"{out['fp64']['ms_per_step']} ms (FP64)")
    check = not args.no_cpu_baseline
    if check:
        log("headline: FP64 oracle")
        out["accuracy_vs_fp64_oracle"] = {
            "fp32refine": accuracy(p32, oracle_main(OUTER, NX, NY, NT, PARAMS, True)),
            "fp64": accuracy(p64, oracle_main(OUTER, NX, NY, NT, PARAMS, False))}
    if not args.no_forms:
        out["plain_small"] = bench_main_workload(lvc, 16, 32, 256, 256, check=check)
        out["precision_form"] = bench_main_workload(lvc, 256, 256, 256, 128, check=check)
        out["batched"] = bench_batched(lvc, check=check)
        out["serving"] = bench_serving_sweep(lvc, args.serve_rates.split(","))
        log("ocean-sim tridiagonal entries")
        out["ocean_tridiag"] = bench_ocean(lvc)
        log("OptionPricing")
        out["option_pricing"] = bench_option_pricing(cpu=check)
        log("InterestCalib")
        out["interest_calib"] = bench_interest_calib(cpu=check)
    if not args.no_cpu_baseline:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
