#!/usr/bin/env python3
"""Benchmark of the GPU pricing engine on one MI355X.

Headline workload (one "step"): the 1,000-contract book of
kontracts_amd.book (250 each of European, monthly-fixing Asian, up-and-out
barrier and 5-asset basket calls; maturities 0.25-3 y on daily grids; 2^16 to
2^20 paths per contract; single-asset contracts half GBM, half Heston; 1.63e11
path-steps), priced with the pathwise Greeks of its Lipschitz contracts
(European, Asian, basket) in persistent launches over (contract, block)
queues under the autotuned tile, and the barriers in the price-only
persistent launch.  `value` = the book's path-steps per second of wall time
over the timed steps, inputs (contracts, models, seeds) resident: nothing
but the program tables crosses PCIe.

The JSON line also carries, as sub-objects of "configs", the other
workloads of the metric:
  config1  European call, GBM, 1,000,000 paths x 252 steps
  config2  up-and-out call CRN Greeks (7 scenario pricings), 4,194,304 x 252
  config3  Heston call, 2,097,152 x 365: six pathwise Greeks vs 13 bump
           pricings at equal standard error
  config5  American put by Longstaff-Schwartz, 2^20 paths x 50 dates (HBM)
each with its own roofline.  Roofline figures use device times from HIP
events the engine records on its own stream around every path-kernel launch
group; cpu_baseline objects time the oracle (the C++ FP64 restatement of the
reference pricer) on this host's cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "path-steps/s on 1 MI355X vs host CPU; adjoint vs bump Greeks/s at equal stderr"
PEAK_BASIS = ("FP32 vector 157.3 TFLOP/s = 78.6e12 FMA lane-ops/s; lane-ops are the minimal gfx950 instruction "
              "sequence per path-step (kontracts_amd/roofline.py); no MFMA: no step is a dense contraction, the "
              "work is Philox integer ops, inverse normals and FMAs")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    # the box's CPU share for one GPU is OMP_NUM_THREADS (16); use all of it
    share = int(os.environ.get("OMP_NUM_THREADS", affinity) or affinity)
    return model, affinity, max(1, min(affinity, share))


def timed_steps(eng, fn, steps):
    """Wall time and engine device time of `steps` calls (each call returns
    after the engine's stream has drained)."""
    eng.set_profiling(True)
    eng.kernel_time(reset=True)
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    wall = time.perf_counter() - t0
    dev_ms, launches, path_steps = eng.kernel_time(reset=True)
    eng.set_profiling(False)
    return out, wall, dev_ms, launches, path_steps


def median_call(eng, fn, reps=5):
    """Median wall and device time of reps calls after a warm one."""
    out = fn()
    walls, devs = [], []
    for _ in range(reps):
        out, w, d, _, _ = timed_steps(eng, fn, 1)
        walls.append(w)
        devs.append(d)
    return out, statistics.median(walls), statistics.median(devs)


def interleaved_calls(eng, fns, reps=5):
    """Median wall and device time of each function over reps rounds that
    call every function once, after one warm call of each."""
    outs = [fn() for fn in fns]
    walls = [[] for _ in fns]
    devs = [[] for _ in fns]
    for _ in range(reps):
        for i, fn in enumerate(fns):
            outs[i], w, d, _, _ = timed_steps(eng, fn, 1)
            walls[i].append(w)
            devs[i].append(d)
    return [(outs[i], statistics.median(walls[i]), statistics.median(devs[i])) for i in range(len(fns))]


def roof(lane_ops_per_launch, avg_launch_ms, kernel, extra=None):
    from kontracts_amd import roofline as R
    achieved = lane_ops_per_launch / (avg_launch_ms / 1e3) / 1e12 if avg_launch_ms > 0 else 0.0
    peak = R.PEAK_LANE_OPS / 1e12
    out = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 3), "unit": "Tlane-op/s",
           "frac": round(achieved / peak, 4), "traffic": None, "kernel": kernel,
           "avg_launch_ms": round(avg_launch_ms, 4), "lane_ops_per_launch": lane_ops_per_launch,
           "peak_basis": PEAK_BASIS}
    if extra:
        out.update(extra)
    return out


def traffic_from_profiles(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


# ---------------------------------------------------------------------------
# Config 4: the 1,000-contract book (headline)
# ---------------------------------------------------------------------------
def bench_book(args, eng, K, B, R, T, threads):
    jobs = B.book(1000)
    profiles = [B.job_profile(i, j) for i, j in enumerate(jobs)]
    path_steps = B.path_steps(jobs)
    # algorithmic lane-ops of one step: each job's paths x steps x its kernel's count
    lane_ops = sum(job[2] * (len(job[5]) - 1) * R.book_job(kind, model, n, nf, greeks=True)
                   for job, (kind, model, n, nf) in zip(jobs, profiles))
    if args.no_tune:
        tuned = {"best": T.DEFAULT_TILE, "device_ms": {}}
    else:
        tuned = T.autotune_portfolio(jobs, repeats=1, greeks=True)
    log(f"config4: autotuned tile {tuned['best']} ({', '.join(f'{k}: {v:.1f} ms' for k, v in tuned['device_ms'].items())})")

    def step():
        return K.price_portfolio(jobs, greeks=True)

    for _ in range(args.warmup):
        step()
    res, wall, dev_ms, launches, _ = timed_steps(eng, step, args.steps)
    value = path_steps * args.steps / wall
    dev_per_step = dev_ms / args.steps
    log(f"config4: {1e3 * wall / args.steps:.1f} ms/step ({dev_per_step:.1f} ms device), {value:.4e} path-steps/s")
    n_greeks = sum(1 for _, g in res if g is not None)
    lane_ops_po = sum(job[2] * (len(job[5]) - 1) * R.book_job(kind, model, n, nf, greeks=False)
                      for job, (kind, model, n, nf) in zip(jobs, profiles))
    roofline = roof(lane_ops, dev_per_step,
                    "portfolio_adjoint_kernel<GBM | Heston | Heston+feeds | basket5> + portfolio_kernel<16> "
                    "(the launch sequence of one step, device time summed)",
                    {"launch_groups_per_step": launches / args.steps,
                     "lane_ops_per_path_step_mean": round(lane_ops / path_steps, 3),
                     "traffic": traffic_from_profiles("portfolio_book_traffic.json")})
    kinds = {}
    for (kind, model, _, _) in profiles:
        kinds[f"{kind}/{model}"] = kinds.get(f"{kind}/{model}", 0) + 1
    info = {
        "workload": "1,000-contract book (kontracts_amd.book, seed 42): 250 each of European, monthly-fixing Asian, "
                    "up-and-out barrier and 5-asset basket (rho 0.3) calls; K 80-120% of 100, vol 0.15-0.45, T "
                    "0.25-3 y on daily grids (63-756 steps), 2^16-2^20 paths, single-asset contracts GBM/Heston "
                    "50/50; pathwise Greeks on the 750 Lipschitz contracts, barriers priced; persistent launches "
                    "under the autotuned tile",
        "contracts": len(jobs), "path_steps": path_steps, "kinds": kinds, "with_greeks": n_greeks,
        "tile": {"best": dict(zip(("blocks_per_cu", "queue_chunk", "price_waves_per_simd", "greeks_waves_per_simd",
                                   "groups_per_iteration", "lanes_per_workgroup"), tuned["best"])),
                 "paths_per_lane": 256 // tuned["best"][5],
                 "device_ms": tuned["device_ms"]},
        "parallelism": "single GPU",
    }
    if args.no_compare:
        return value, wall / args.steps, info, roofline, {}, None

    # the reference's way: one call per contract (price_gbm loop), pathwise
    # Greeks per contract where the payoff admits them
    def per_contract():
        for i, job in enumerate(jobs):
            kw = dict(n_paths=job[2], seed=job[3], steps_per_year=job[4], grid=job[5])
            if profiles[i][0] == "barrier":
                job[0].price(job[1], **kw)
            else:
                job[0].pathwise_greeks(job[1], **kw)

    _, wall_pc, dev_pc, _, _ = timed_steps(eng, per_contract, 1)
    # price only, the same book, with the price kernel's register budget
    # tuned on this call (it runs every job there; the Greeks book's tile
    # was tuned on a call where the price kernel runs only the barriers)
    if args.no_tune:
        po_tuned = {"best": tuned["best"], "device_ms": {}}
    else:
        po_tuned = T.autotune_portfolio(jobs, candidates=[tuple(tuned["best"][:2])], repeats=1, greeks=False,
                                        unroll_lanes=[])
    _, wall_po, dev_po, _, _ = timed_steps(eng, lambda: K.price_portfolio(jobs), 2)
    T.set_tile(eng, tuned["best"])
    log(f"config4 price only: {1e3 * wall_po / 2:.1f} ms/call ({dev_po / 2:.1f} ms device), price kernel tile "
        f"{po_tuned['best'][2]} waves ({', '.join(f'{k}: {v:.1f} ms' for k, v in po_tuned['device_ms'].items())})")
    extra = {
        "per_contract_s": wall_pc, "per_contract_device_ms": dev_pc,
        "speedup_vs_per_contract": wall_pc / (wall / args.steps),
        "price_only_ms": 1e3 * wall_po / 2, "price_only_device_ms": dev_po / 2,
        "price_only_path_steps_per_s": path_steps * 2 / wall_po,
        "price_only_roofline_frac": round(lane_ops_po / (dev_po / 2 / 1e3) / R.PEAK_LANE_OPS, 4),
        "price_only_tile": {"price_waves_per_simd": po_tuned["best"][2], "device_ms": po_tuned["device_ms"]},
        "first_prices": [round(r.price, 6) for r, _ in res[:4]],
        "speedup_target_note": "the 3x over per-contract launches is out of reach for this book: its contracts "
                               "(2^16-2^20 paths x 63-756 steps) already fill the GPU one launch at a time, so "
                               "both ways are VALU-bound and the persistent launch saves only the launch gaps "
                               "and tails; the launch-bound book below shows the gain where launches dominate",
    }
    extra["launch_bound_book"] = bench_small_book(eng, K, B)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_book(jobs, threads, args.cpu_paths_book)
    return value, wall / args.steps, info, roofline, extra, cpu


def bench_small_book(eng, K, B):
    """The launch-bound regime: a book of the same kinds at 2^14 paths and
    3-month maturities (63 steps), one persistent launch (with the pathwise
    Greeks) against one call per contract; medians of 5 after a warm call."""
    small = B.book(1000, paths_exp=(14, 14), maturity=(0.25, 0.25))
    kinds = [B.job_profile(i, j)[0] for i, j in enumerate(small)]

    def per_contract():
        for kind, job in zip(kinds, small):
            kw = dict(n_paths=job[2], seed=job[3], steps_per_year=job[4], grid=job[5])
            if kind == "barrier":
                job[0].price(job[1], **kw)
            else:
                job[0].pathwise_greeks(job[1], **kw)

    _, t_book, d_book = median_call(eng, lambda: K.price_portfolio(small, greeks=True))
    _, t_loop, d_loop = median_call(eng, per_contract, reps=3)
    _, t_book_p, _ = median_call(eng, lambda: K.price_portfolio(small))

    def per_contract_price():
        for job in small:
            job[0].price(job[1], n_paths=job[2], seed=job[3], steps_per_year=job[4], grid=job[5])

    _, t_loop_p, _ = median_call(eng, per_contract_price, reps=3)
    log(f"launch-bound book: {t_book * 1e3:.1f} ms persistent vs {t_loop * 1e3:.1f} ms per contract "
        f"({t_loop / t_book:.2f}x; price only {t_loop_p / t_book_p:.2f}x)")
    return {"workload": "1,000 contracts of the same kinds at 2^14 paths x 63 daily steps (T = 0.25)",
            "path_steps": B.path_steps(small), "persistent_ms": t_book * 1e3, "persistent_device_ms": d_book,
            "per_contract_ms": t_loop * 1e3, "speedup": t_loop / t_book,
            "price_only_persistent_ms": t_book_p * 1e3, "price_only_per_contract_ms": t_loop_p * 1e3,
            "price_only_speedup": t_loop_p / t_book_p}


def cpu_book(jobs, threads, n_paths):
    """The oracle (reference algorithm in f64: materialised paths, recursive
    contract evaluation, two-pass summarize) on every 25th job of the book
    (10 of each kind) at n_paths paths each."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from test_gpu_parity import oracle_model

    sample = jobs[::25]
    ps = 0.0
    t0 = time.perf_counter()
    for c, m, _, seed, spy, grid in sample:
        O.price(c.to_json(), oracle_model(m), O.cfg(n_paths, seed, spy, m.rate, grid=grid, threads=threads))
        ps += n_paths * (len(grid) - 1)
    dt = time.perf_counter() - t0
    model, affinity, _ = cpu_info()
    log(f"cpu baseline (book sample) {ps / dt:.4e} path-steps/s on {threads} threads ({dt:.1f} s)")
    return {"value": round(ps / dt, 1), "unit": "path-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle price_gbm restatement (f64 paths, recursive AST evaluation) on every 25th contract of "
                      f"the book (40 contracts, 10 per kind) at {n_paths} paths each, {ps:.3e} path-steps in "
                      f"{dt:.1f} s; prices only (the reference has no pathwise Greeks)",
            "cpu_model": model, "affinity_cores": affinity,
            "threads_note": "threads = the host's CPU share per GPU (OMP_NUM_THREADS), capped by the affinity set"}


# ---------------------------------------------------------------------------
# Config 2: up-and-out CRN Greeks
# ---------------------------------------------------------------------------
def bench_config2(args, eng, K, N, R, threads):
    contract = K.up_and_out_call("X", 100.0, 200.0, 1.0)
    model = K.GBM(180.0, 0.25, 0.05, asset="X")
    n, steps = 4_194_304, 252

    def step():
        return contract.greeks(model, n_paths=n, seed=42, steps_per_year=steps, bumps=(1.8, 0.01, 1e-4))

    # the engine clock ramps up over ~40 launches of this kernel from idle
    for _ in range(40):
        step()
    g, wall, dev_ms, launches, _ = timed_steps(eng, step, 20)
    fma = valu_probe(N, eng)
    gen = rng_probe(N, eng, True)
    avg_launch_ms = dev_ms / max(launches, 1)
    lane_ops = R.gbm_barrier(scenarios=7, states=5)
    per_launch = float(n) * steps
    rate = per_launch / (avg_launch_ms / 1e3)
    ceiling = 1.0 / (1.0 / gen["normals_per_s"] + (lane_ops - R.NORMAL) / fma["fma_lane_ops_per_s"])
    rl = roof(per_launch * lane_ops, avg_launch_ms, "path_kernel<GBM, 7 CRN scenarios>",
              {"lane_ops_per_path_step": lane_ops, "path_steps_per_launch": per_launch,
               "traffic": traffic_from_profiles("path_kernel_gbm_barrier_traffic.json"),
               "fma_issue_measured": round(fma["fma_lane_ops_per_s"] / 1e12, 3),
               "frac_of_measured_fma_issue": round(per_launch * lane_ops / (avg_launch_ms / 1e3)
                                                   / fma["fma_lane_ops_per_s"], 4),
               "normals_measured": round(gen["normals_per_s"] / 1e12, 4),
               "mix_ceiling_path_steps_per_s": round(ceiling, 1),
               "frac_of_mix_ceiling": round(rate / ceiling, 4)})
    value = 7.0 * n * steps * 20 / wall
    log(f"config2: {1e3 * wall / 20:.3f} ms/step, {value:.4e} scenario path-steps/s, kernel {avg_launch_ms:.3f} ms, "
        f"frac {rl['frac']}")
    out = {"workload": "up-and-out call (S-100)^+ KO at S>=200, GBM s0=180 sigma=0.25 r=0.05, 4,194,304 paths x 252 "
                       "steps, CRN bump Greeks (7 scenario pricings sharing the normals)",
           "value": value, "unit": "scenario path-steps/s", "ms_per_step": 1e3 * wall / 20,
           "pricing_path_steps_per_s": n * steps * 20 / wall, "steps": 20, "warmup": 40, "roofline": rl,
           "greeks": {"price": g.price, "delta": g.delta, "gamma": g.gamma, "vega": g.vega, "rho": g.rho,
                      "std_errors": list(g.std_errors)}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        npaths = args.cpu_paths
        t0 = time.perf_counter()
        O.greeks_gbm(contract.to_json(), "X", 180.0, 0.25, O.cfg(npaths, 42, 252, 0.05, threads=threads),
                     1.8, 0.01, 1e-4)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(7 * npaths * 252 / dt, 1), "unit": "scenario path-steps/s",
                               "cores": threads, "kind": "port",
                               "sample": f"oracle greeks_gbm (7 CRN pricings, f64) on {npaths} paths x 252 steps, "
                                         f"{dt:.1f} s"}
    return out


def valu_probe(N, eng, reps=20, iters=4096):
    """FP32 FMA issue rate of the chip, measured on a plain FMA loop."""
    import ctypes as C
    ms, ops = C.c_double(), C.c_double()
    N.check(N.lib().kt_debug_valu(eng.handle, iters, reps, C.byref(ms), C.byref(ops)), eng.handle)
    return {"kernel": "debug_valu_kernel", "iters": iters, "reps": reps, "avg_ms": ms.value,
            "fma_lane_ops_per_s": ops.value}


def rng_probe(N, eng, normals, reps=20, iters=256):
    """Rate of the normal generator alone."""
    import ctypes as C
    ms, words = C.c_double(), C.c_double()
    N.check(N.lib().kt_debug_rng_rate(eng.handle, 1 if normals else 0, iters, reps, C.byref(ms), C.byref(words)),
            eng.handle)
    return {"kernel": f"debug_rng_rate_kernel<{'true' if normals else 'false'}>", "iters": iters, "reps": reps,
            "avg_ms": ms.value, ("normals_per_s" if normals else "words_per_s"): words.value}


# ---------------------------------------------------------------------------
# Config 3: Heston pathwise Greeks vs bump-and-reprice at equal stderr
# ---------------------------------------------------------------------------
def bench_config3(eng, K, R):
    c = K.european_call("X", 100.0, 1.0)
    hm = K.heston(100, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X")
    n = 2_097_152
    hb = [1.0, 1e-3, 1e-2, 1e-3, 1e-2, 1e-2]

    def tangent():
        eng.set_option("pathwise_mode", 1)
        try:
            return c.pathwise_greeks(hm, n_paths=n, steps_per_year=365)
        finally:
            eng.set_option("pathwise_mode", 0)

    # the engine clock settles after some tens of launches: warm it up, then
    # time the four calls in interleaved rounds so that each sees the same
    # clock conditions
    for _ in range(40):
        c.price(hm, n_paths=n, steps_per_year=365)
    (rf, tf, kf), (gA, tA, kA), (_, tB, kB), (gT, tT, kT) = interleaved_calls(eng, [
        lambda: c.price(hm, n_paths=n, steps_per_year=365),
        lambda: c.pathwise_greeks(hm, n_paths=n, steps_per_year=365),
        lambda: c.bump_greeks(hm, bumps=hb, n_paths=n, steps_per_year=365, std_errors=False),
        tangent])
    gB = c.bump_greeks(hm, bumps=hb, n_paths=n, steps_per_year=365)
    eq = {nm: (tB / tA) * (sb / sa) ** 2 for nm, sa, sb in zip(gA.names, gA.std_errors, gB.std_errors)}
    eqT = {nm: (tB / tT) * (sb / sa) ** 2 for nm, sa, sb in zip(gT.names, gT.std_errors, gB.std_errors)}
    ps = float(n) * 365
    log(f"config3: pathwise {tA * 1e3:.2f} ms vs 13 bumps {tB * 1e3:.2f} ms: {tB / tA:.2f}x raw, "
        f"min {min(eq.values()):.2f}x at equal stderr")
    return {
        "workload": "Heston call s0=K=100 v0=theta=0.04 kappa=2 sigma_v=0.3 rho=-0.5 r=0.05 T=1, 2,097,152 paths x "
                    "365 steps; d price / d {s0, v0, kappa, theta, sigma_v, rho}: pathwise reverse sweep with a "
                    "checkpoint every 16 steps vs 13 CRN bump pricings",
        "timing": "after 40 warm-up pricings, median of 5 interleaved rounds of (price, pathwise, 13 bumps, "
                  "tangent)",
        "forward_ms": tf * 1e3, "forward_kernel_ms": kf, "forward_path_steps_per_s": ps / tf,
        "forward_roofline": roof(ps * R.heston_european(), kf, "path_kernel<Heston>"),
        "pathwise_ms": tA * 1e3, "pathwise_kernel_ms": kA,
        "pathwise_roofline": roof(ps * R.heston_pathwise(), kA, "adjoint_heston_kernel (forward, recompute, sweep)"),
        "bump13_ms": tB * 1e3, "bump13_kernel_ms": kB,
        "speedup_raw": tB / tA, "speedup_equal_stderr": eq, "speedup_equal_stderr_min": min(eq.values()),
        "pathwise": dict(zip(gA.names, gA.values)), "pathwise_se": dict(zip(gA.names, gA.std_errors)),
        "bump": dict(zip(gB.names, gB.values)), "bump_se": dict(zip(gB.names, gB.std_errors)),
        "price": gA.price, "price_se": gA.price_std_error,
        "tangent_ms": tT * 1e3, "tangent_kernel_ms": kT, "tangent_speedup_equal_stderr": eqT,
        "greeks_per_s_pathwise": 6 / tA, "greeks_per_s_bump": 6 / tB,
    }


# ---------------------------------------------------------------------------
# Config 1 and config 5
# ---------------------------------------------------------------------------
def bench_config1(eng, K, R):
    grid = [i / 252 for i in range(253)]
    c1 = K.european_call("X", 100.0, 1.0)
    m1 = K.GBM(100.0, 0.2, 0.05, asset="X")
    r1, t1, k1 = median_call(eng, lambda: c1.price(m1, n_paths=1_000_000, grid=grid))
    ps = 1_000_000 * 252.0
    # the book's 5-asset basket call priced alone (pairwise rho 0.3)
    fs = [K.GbmFactor(f"A{i}", 100.0, 0.05, 0.15 + 0.05 * i) for i in range(5)]
    b5 = K.correlated_gbm(fs, [[1.0 if i == j else 0.3 for j in range(5)] for i in range(5)], 0.05)
    avg5 = (K.spot("A0") + K.spot("A1") + K.spot("A2") + K.spot("A3") + K.spot("A4")) / 5.0
    cb = K.scale((avg5 - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)
    rb, tb, kb = median_call(eng, lambda: cb.price(b5, n_paths=1 << 20, grid=grid))
    psb = (1 << 20) * 252.0
    return {"workload": "European call GBM s0=K=100 sigma=0.2 r=0.05 T=1, 1,000,000 paths x 252 steps (explicit "
                        "grid), seed 42", "price": r1.price, "std_error": r1.std_error,
            "black_scholes": 10.450583572185565, "ms": t1 * 1e3, "kernel_ms": k1, "path_steps_per_s": ps / t1,
            "roofline": roof(ps * R.gbm_european(), k1, "path_kernel<GBM>"),
            "basket5": {"workload": "5-asset basket call (vols 0.15-0.35, pairwise rho 0.3), 1,048,576 paths x 252 "
                                    "steps", "price": rb.price, "std_error": rb.std_error, "ms": tb * 1e3,
                        "kernel_ms": kb, "path_steps_per_s": psb / tb,
                        "roofline": roof(psb * R.basket_european(5), kb, "path_kernel<basket K=5>")}}


def bench_config5(eng, K):
    mL = K.GBM(36.0, 0.2, 0.06, asset="X")
    put = K.scale((K.const_(40.0) - K.spot("X")).max(0.0), K.one("USD"))
    dates = [i / 50 for i in range(1, 51)]
    out = {"workload": "American put by LSM, GBM s0=36 K=40 sigma=0.2 r=0.06 T=1, 50 exercise dates, basis "
                       "{1, S, S^2}, paths stored [date][path] in HBM", "timing": "median of 5 calls after a warm one"}
    for label, nL in (("n_1048576", 1 << 20), ("n_8388608", 1 << 23)):
        rL, tL, kL = median_call(eng, lambda: put.price_american(mL, dates, n_paths=nL, seed=42))
        # algorithmic bytes: the paths kernel writes payoff + spot (8 B per
        # path and date); each backward pass reads the payoff/spot pairs of two
        # dates and reads + writes the f64 value (32 B per path and date)
        nbytes = nL * len(dates) * (8 + 32)
        out[label] = {"paths": nL, "price": rL.price, "std_error": rL.std_error, "ms": tL * 1e3, "device_ms": kL,
                      "path_dates_per_s": nL * len(dates) / tL, "algorithmic_bytes": nbytes,
                      "roofline": {"bound": "hbm", "achieved": round(nbytes / (kL / 1e3) / 1e9, 1),
                                   "peak": 8000.0, "unit": "GB/s",
                                   "frac": round(nbytes / (kL / 1e3) / 8e12, 4)}}
    out["reference_ls2001"] = {"lsm": 4.472, "se": 0.010, "fd": 4.478}
    return out


# ---------------------------------------------------------------------------
# The next rows built on the same kernels: Greek surface, variance
# reduction, calibration-sized calls
# ---------------------------------------------------------------------------
def bench_next(eng, K, R):
    out = {}
    # Greek surface: 20 x 20 (spot, vol) cells of the config-2 barrier at
    # 2^16 paths x 252 steps in one launch (src/surface.rs:104-138), against a
    # loop of greeks calls over 100 of the cells (scaled to 400)
    bar = K.up_and_out_call("X", 100.0, 200.0, 1.0)
    spots = [150.0 + 2.5 * i for i in range(20)]
    vols = [0.15 + 0.01 * j for j in range(20)]
    surf = lambda: K.greek_surface(bar, "X", spots, vols, n_paths=1 << 16, seed=42, steps_per_year=252)
    # the same kernel (7 CRN scenarios of one barrier) over the surface's
    # 26,214,400 paths in one launch of one cell, timed in the same rounds
    single = lambda: bar.greeks(K.GBM(180.0, 0.25, 0.05, asset="X"), n_paths=400 << 16, seed=42,
                                steps_per_year=252)
    for _ in range(8):  # the engine clock settles after some tens of ms of launches
        surf()
    (_, t_s, k_s), (_, t_1, k_1) = interleaved_calls(eng, [surf, single], reps=5)
    t0 = time.perf_counter()
    for s_ in spots[:5]:
        for v_ in vols:
            bar.greeks(K.GBM(s_, v_, 0.05, asset="X"), n_paths=1 << 16, seed=42, steps_per_year=252)
    t_loop = (time.perf_counter() - t0) * 4
    sps = 400 * 7 * (1 << 16) * 252
    out["greek_surface"] = {
        "workload": "20 x 20 (spot 150-197.5, vol 0.15-0.34) cells of the up-and-out call, 65,536 paths x 252 "
                    "steps, seven CRN scenarios per cell, one launch",
        "ms": t_s * 1e3, "kernel_ms": k_s, "loop_of_greeks_calls_ms": t_loop * 1e3,
        "single_launch_same_paths_kernel_ms": k_1,
        "scenario_path_steps_per_s": sps / t_s,
        "roofline": roof(400 * (1 << 16) * 252 * R.gbm_barrier(7, 5), k_s, "surface_kernel (7 CRN scenarios per cell)"),
    }
    # variance reduction on a daily-average Asian call (1,000,000 paths x 252)
    m1 = K.GBM(100.0, 0.2, 0.05, asset="X")
    asian = K.scale((K.average(K.spot("X")) - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)
    g252 = [i / 252 for i in range(253)]
    vr = {"workload": "arithmetic-average Asian call (every grid point, incl. t = 0), GBM s0=K=100 sigma=0.2 "
                      "r=0.05 T=1, 1,000,000 paths x 252 steps (src/variance_reduction.rs estimators)"}
    for label, anti, cv in (("plain", False, False), ("antithetic", True, False), ("control_variate", False, True),
                            ("both", True, True)):
        r, t, k = median_call(eng, lambda: asian.price(m1, n_paths=1_000_000, seed=42, grid=g252, antithetic=anti,
                                                       control_variate=cv), reps=3)
        vr[label] = {"price": r.price, "std_error": r.std_error, "ms": t * 1e3, "kernel_ms": k}
    out["variance_reduction"] = vr
    # calibration-sized calls: a 1,000-path x 50-step knock-out price, and ten
    # such quotes as one portfolio call against a loop of ten calls
    ko = K.up_and_out_call("X", 100.0, 130.0, 1.0)
    for _ in range(20):
        ko.price(K.GBM(100.0, 0.2, 0.05, asset="X"), n_paths=1000, seed=1)
    t0 = time.perf_counter()
    for i in range(200):
        ko.price(K.GBM(100.0, 0.2 + 1e-4 * i, 0.05, asset="X"), n_paths=1000, seed=1)
    single_us = (time.perf_counter() - t0) / 200 * 1e6
    quotes = [(ko, K.GBM(float(s_), 0.2, 0.05, asset="X"), 1000, 1, 50) for s_ in range(90, 110, 2)]
    for _ in range(10):
        K.price_portfolio(quotes)
    t0 = time.perf_counter()
    for _ in range(100):
        K.price_portfolio(quotes)
    pf_us = (time.perf_counter() - t0) / 100 * 1e6
    t0 = time.perf_counter()
    for _ in range(20):
        for q in quotes:
            q[0].price(q[1], n_paths=q[2], seed=q[3], steps_per_year=q[4])
    loop_us = (time.perf_counter() - t0) / 20 * 1e6
    out["calibration_calls"] = {"workload": "up-and-out call, 1,000 paths x 50 steps per quote",
                                "single_price_us": single_us, "ten_quotes_portfolio_us": pf_us,
                                "ten_quotes_loop_us": loop_us}
    log(f"next rows: surface {t_s * 1e3:.1f} ms ({k_s:.2f} ms kernel; one launch of the same paths {k_1:.2f} ms) "
        f"vs loop {t_loop * 1e3:.0f} ms; "
        f"single price {single_us:.0f} us; 10 quotes {pf_us:.0f} us vs {loop_us:.0f} us")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-paths", type=int, default=262_144, help="config-2 CPU sample paths")
    ap.add_argument("--cpu-paths-book", type=int, default=4096, help="paths per contract of the book's CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="book: skip the per-contract and launch-bound comparisons (profiling runs)")
    ap.add_argument("--no-tune", action="store_true", help="book: default tile instead of the autotuner (profiling runs)")
    ap.add_argument("--only", choices=["book", "config2", "config3", "config1", "config5", "next"], default=None,
                    help="run one workload only (profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on MIB MiB (profiler byte-counter calibration)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only run the FMA issue-rate kernel (eight independent FMA chains per lane)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (paths would shard by index with one "
              "f64 statistics reduction per contract); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    from kontracts_amd import _native as N
    eng = N.engine(0)
    if args.valu_calibration:
        print(json.dumps(valu_probe(N, eng, reps=50)))
        print(json.dumps(rng_probe(N, eng, False, reps=50)))
        print(json.dumps(rng_probe(N, eng, True, reps=50)))
        return
    if args.copy_calibration:
        import ctypes as C
        ms = C.c_double()
        nbytes = args.copy_calibration << 20
        N.check(N.lib().kt_debug_copy(eng.handle, nbytes, 5, C.byref(ms)), eng.handle)
        print(json.dumps({"copy_bytes": nbytes, "reps": 5, "avg_ms": ms.value,
                          "GBps": 2 * nbytes / (ms.value / 1e3) / 1e9}))
        return

    import kontracts_amd as K
    from kontracts_amd import book as B
    from kontracts_amd import roofline as R
    from kontracts_amd import tune as T

    model, affinity, threads = cpu_info()
    configs = {}
    if args.only in (None, "config2"):
        configs["config2"] = bench_config2(args, eng, K, N, R, threads)
    if args.only in (None, "config3"):
        configs["config3"] = bench_config3(eng, K, R)
    if args.only in (None, "config1"):
        configs["config1"] = bench_config1(eng, K, R)
    if args.only in (None, "config5"):
        configs["config5"] = bench_config5(eng, K)
    nxt = bench_next(eng, K, R) if args.only in (None, "next") else None
    if args.only not in (None, "book"):
        print(json.dumps({"configs": configs, "next_rows": nxt}))
        return
    value, step_s, info, roofline, extra, cpu = bench_book(args, eng, K, B, R, T, threads)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 paths, f64 payoff/stderr reduction",
        "data": "synthetic (seeded contract book, models and seeds; counter-based Philox normals generated on the GPU)",
        "config": info,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "book": extra,
        "configs": configs,
        "next_rows": nxt,
        "host": {"cpu_model": model, "affinity_cores": affinity, "threads_used": threads},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
