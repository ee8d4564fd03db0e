#!/usr/bin/env python3
"""Benchmark of the GPU pricing engine on one MI355X.

Workload (one "step"): the up-and-out barrier call ((S-100)^+ at T=1, knocked
out when S >= 200 on any of 253 daily monitoring points) under GBM s0=180,
sigma=0.25, r=0.05, 4,194,304 paths x 252 steps, priced with its Greeks by CRN
bump-and-reprice exactly as greeks_gbm does (price, S+-1.8, sigma+-0.01,
r+-1e-4: seven scenario pricings sharing one random stream).  The metric is
scenario path-steps per second: 7 x paths x steps per step, divided by the
wall time of the timed steps.  Inputs are a contract, a model and a seed, so
nothing crosses PCIe inside the timed region except the program tables
(a few KB) and the results.

Prints one JSON line.  The roofline object uses HIP events recorded by the
engine on its own stream around every path-kernel launch; cpu_baseline times
the oracle (the C++ f64 restatement of the reference pricer) on this host's
cores on a bounded number of paths.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "path-steps/s on 1 MI355X vs host CPU; adjoint vs bump Greeks/s at equal stderr"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n_paths: int, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    import kontracts_amd as K

    c = K.up_and_out_call("X", 100.0, 200.0, 1.0)
    cfg = O.cfg(n_paths, 42, 252, 0.05, threads=threads)
    t0 = time.perf_counter()
    O.greeks_gbm(c.to_json(), "X", 180.0, 0.25, cfg, 1.8, 0.01, 1e-4)
    dt = time.perf_counter() - t0
    return 7 * n_paths * 252 / dt, dt


def traffic_from_profiles():
    path = os.path.join(ROOT, "profiles", "path_kernel_gbm_barrier_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def valu_probe(N, eng, reps=20, iters=4096):
    """FP32 FMA issue rate of the chip, measured on a plain FMA loop."""
    import ctypes as C
    ms, ops = C.c_double(), C.c_double()
    N.check(N.lib().kt_debug_valu(eng.handle, iters, reps, C.byref(ms), C.byref(ops)), eng.handle)
    return {"kernel": "debug_valu_kernel", "iters": iters, "reps": reps, "avg_ms": ms.value,
            "fma_lane_ops_per_s": ops.value}


def rng_probe(N, eng, normals, reps=20, iters=256):
    """Rate of the normal generator alone: Philox blocks with the path kernels'
    counter layout, folded as raw words or as their FP32 inverse normals."""
    import ctypes as C
    ms, words = C.c_double(), C.c_double()
    N.check(N.lib().kt_debug_rng_rate(eng.handle, 1 if normals else 0, iters, reps, C.byref(ms), C.byref(words)),
            eng.handle)
    return {"kernel": f"debug_rng_rate_kernel<{'true' if normals else 'false'}>", "iters": iters, "reps": reps,
            "avg_ms": ms.value, ("normals_per_s" if normals else "words_per_s"): words.value}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the engine clock ramps up over the first ~40 launches from idle
    # (1.3 ms -> 1.05 ms per launch of this kernel): warm up past that
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--paths", type=int, default=4_194_304)
    ap.add_argument("--cpu-paths", type=int, default=262_144)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--extras", action="store_true", help="also time configs 1 and 3 (single pricings)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on MIB MiB (profiler byte-counter calibration)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only run the FMA issue-rate kernel (eight independent FMA chains per lane)")
    args = ap.parse_args()
    if args.valu_calibration:
        from kontracts_amd import _native as N
        eng = N.engine(0)
        print(json.dumps(valu_probe(N, eng, reps=50)))
        print(json.dumps(rng_probe(N, eng, False, reps=50)))
        print(json.dumps(rng_probe(N, eng, True, reps=50)))
        return
    if args.copy_calibration:
        import ctypes as C
        from kontracts_amd import _native as N
        eng = N.engine(0)
        ms = C.c_double()
        nbytes = args.copy_calibration << 20
        N.check(N.lib().kt_debug_copy(eng.handle, nbytes, 5, C.byref(ms)), eng.handle)
        print(json.dumps({"copy_bytes": nbytes, "reps": 5, "avg_ms": ms.value,
                          "GBps": 2 * nbytes / (ms.value / 1e3) / 1e9}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (paths would shard by index with one "
              "f64 statistics reduction per contract); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import kontracts_amd as K
    from kontracts_amd import _native as N
    from kontracts_amd import roofline as R

    eng = N.engine(0)
    contract = K.up_and_out_call("X", 100.0, 200.0, 1.0)
    model = K.GBM(180.0, 0.25, 0.05, asset="X")
    bumps = (1.8, 0.01, 1e-4)
    n, steps = args.paths, 252

    def one_step():
        return contract.greeks(model, n_paths=n, seed=42, steps_per_year=steps, bumps=bumps)

    for _ in range(args.warmup):
        one_step()
    eng.set_profiling(True)
    eng.kernel_time(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g = one_step()
    t1 = time.perf_counter()
    k_ms, k_launches, k_path_steps = eng.kernel_time(reset=True)
    eng.set_profiling(False)
    fma = valu_probe(N, eng)  # right after the timed steps, at the same clocks
    gen = rng_probe(N, eng, True)

    wall = t1 - t0
    ms_per_step = 1e3 * wall / args.steps
    scen_path_steps = 7.0 * n * steps
    value = scen_path_steps * args.steps / wall
    log(f"greeks: price {g.price:.6f} delta {g.delta:.6f} gamma {g.gamma:.6e} vega {g.vega:.6f} rho {g.rho:.6f}")
    log(f"{ms_per_step:.3f} ms/step, {value:.4e} scenario path-steps/s; kernels {k_launches} launches "
        f"{k_ms / max(k_launches, 1):.3f} ms avg")

    # One launch per step: the 7-scenario CRN kernel (5 distinct log-spot paths).
    avg_launch_s = (k_ms / max(k_launches, 1)) / 1e3
    path_steps_per_launch = float(n) * steps
    lane_ops = R.gbm_barrier(scenarios=7, states=5)
    achieved = path_steps_per_launch * lane_ops / avg_launch_s / 1e12 if avg_launch_s > 0 else 0.0
    peak = R.PEAK_LANE_OPS / 1e12
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 3), "unit": "Tlane-op/s",
            "frac": round(achieved / peak, 4), "traffic": traffic_from_profiles(),
            "kernel": "path_kernel<GBM, 7 CRN scenarios>", "avg_launch_ms": round(avg_launch_s * 1e3, 4),
            "launches": k_launches, "lane_ops_per_path_step": lane_ops,
            "path_steps_per_launch": path_steps_per_launch,
            "peak_basis": "FP32 vector 157.3 TFLOP/s = 78.6e12 FMA lane-ops/s; lane-ops are the minimal gfx950 "
                          "instruction sequence per path-step (kontracts_amd/roofline.py); no MFMA: no step is a "
                          "dense contraction, the work is Philox integer ops, inverse normals and FMAs",
            "fma_issue_measured": round(fma["fma_lane_ops_per_s"] / 1e12, 3),
            "frac_of_measured_fma_issue": round(achieved * 1e12 / fma["fma_lane_ops_per_s"], 4)}
    # Mix-limited ceiling: one normal at the generator's measured rate plus the
    # rest of the path-step at the measured FMA rate (Philox's 64-bit multiplies
    # and the transcendentals take several issue cycles, which lane-ops do not
    # weight).
    rate = path_steps_per_launch / avg_launch_s if avg_launch_s > 0 else 0.0
    ceiling = 1.0 / (1.0 / gen["normals_per_s"] + (lane_ops - R.NORMAL) / fma["fma_lane_ops_per_s"])
    roof.update({"normals_measured": round(gen["normals_per_s"] / 1e12, 4),
                 "mix_ceiling_path_steps_per_s": round(ceiling, 1),
                 "frac_of_mix_ceiling": round(rate / ceiling, 4)})

    cpu = None
    if not args.no_cpu:
        threads = min(len(os.sched_getaffinity(0)), 16)
        cv, dt = cpu_baseline(args.cpu_paths, threads)
        cpu = {"value": round(cv, 1), "unit": "path-steps/s", "cores": threads, "kind": "port",
               "sample": f"oracle greeks_gbm restatement (7 CRN pricings, f64, recursive AST evaluation) on "
                         f"{args.cpu_paths} paths x 252 steps of the same contract, {dt:.1f} s"}
        log(f"cpu baseline {cv:.4e} path-steps/s on {threads} threads ({dt:.1f} s)")

    extras = {}
    if args.extras:

        def timed(fn, reps=5):
            """Median wall time and median path-kernel time of reps calls (after one warm call)."""
            out = fn()
            walls, kms = [], []
            for _ in range(reps):
                eng.set_profiling(True)
                eng.kernel_time(reset=True)
                t = time.perf_counter()
                out = fn()
                walls.append(time.perf_counter() - t)
                kms.append(eng.kernel_time(reset=True)[0])
                eng.set_profiling(False)
            return out, statistics.median(walls), statistics.median(kms)

        extras["timing"] = "single pricings and Greeks: median of 5 calls each, after a warm call"
        grid = [i / 252 for i in range(253)]
        c1 = K.european_call("X", 100.0, 1.0)
        m1 = K.GBM(100.0, 0.2, 0.05, asset="X")
        r1, t1, k1 = timed(lambda: c1.price(m1, n_paths=1_000_000, grid=grid))
        extras["config1_european_gbm_path_steps_per_s"] = 1_000_000 * 252 / t1
        extras["config1_european_gbm_kernel_path_steps_per_s"] = 1_000_000 * 252 / (k1 / 1e3)
        extras["config1_price"] = r1.price
        h = K.heston(100, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X")
        r3, t3, k3 = timed(lambda: c1.price(h, n_paths=2_097_152, steps_per_year=365))
        extras["config3_heston_forward_path_steps_per_s"] = 2_097_152 * 365 / t3
        extras["config3_heston_forward_kernel_ms"] = k3
        extras["config3_price"] = r3.price
        fs = [K.GbmFactor(f"A{i}", 100.0, 0.05, 0.15 + 0.05 * i) for i in range(5)]
        b5 = K.correlated_gbm(fs, [[1.0 if i == j else 0.3 for j in range(5)] for i in range(5)], 0.05)
        avg5 = (K.spot("A0") + K.spot("A1") + K.spot("A2") + K.spot("A3") + K.spot("A4")) / 5.0
        cb = K.scale((avg5 - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)
        rb, tb, kb = timed(lambda: cb.price(b5, n_paths=1 << 20, grid=grid))
        extras["basket5_path_steps_per_s"] = (1 << 20) * 252 / tb
        extras["basket5_kernel_ms"] = kb
        extras["basket5_price"] = rb.price
        ca = K.scale((K.average(K.spot("X")) - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)
        ra, ta, ka = timed(lambda: ca.price(m1, n_paths=1 << 20, steps_per_year=252))
        extras["asian_daily_path_steps_per_s"] = (1 << 20) * 252 / ta
        extras["asian_daily_kernel_ms"] = ka
        extras["asian_price"] = ra.price
        # American put by LSM: s0=36, K=40, sigma=0.2, r=0.06, T=1, 50 dates,
        # 2^20 paths stored in HBM.  Algorithmic HBM bytes: the paths kernel
        # writes payoff + spot (8 B per path and date); each backward pass reads
        # payoff/spot of two dates (16 B) and reads + writes the f64 value (16 B).
        mL = K.GBM(36.0, 0.2, 0.06, asset="X")
        put = K.scale((K.const_(40.0) - K.spot("X")).max(0.0), K.one("USD"))
        dates = [i / 50 for i in range(1, 51)]
        nL = 1 << 20
        put.price_american(mL, dates, n_paths=nL, seed=42)
        eng.set_profiling(True)
        eng.kernel_time(reset=True)
        t = time.perf_counter()
        rL = put.price_american(mL, dates, n_paths=nL, seed=42)
        wallL = time.perf_counter() - t
        kL_ms, _, _ = eng.kernel_time(reset=True)
        eng.set_profiling(False)
        bytesL = nL * len(dates) * (8 + 32)
        extras["config5_lsm_ms"] = wallL * 1e3
        extras["config5_lsm_device_ms"] = kL_ms
        extras["config5_lsm_path_steps_per_s"] = nL * len(dates) / wallL
        extras["config5_lsm_algorithmic_GBps"] = bytesL / (kL_ms / 1e3) / 1e9 if kL_ms > 0 else None
        extras["config5_price"] = rL.price
        extras["config5_std_error"] = rL.std_error
        # Config 3 Greeks: Heston 2^21 paths x 365 steps, d price / d {s0, v0,
        # kappa, theta, sigma_v, rho}: pathwise (one forward + adjoint launch)
        # vs CRN bump-and-reprice (13 pricings); then at equal standard error:
        # paths needed scale with the variance, so the speedup per Greek is
        # (t_bump / t_pathwise) * (se_bump / se_pathwise)^2.
        hm = K.heston(100, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X")
        n3 = 2_097_152
        hb = [1.0, 1e-3, 1e-2, 1e-3, 1e-2, 1e-2]

        gA, tA, kA_ms = timed(lambda: c1.pathwise_greeks(hm, n_paths=n3, steps_per_year=365))
        _, tB, kB_ms = timed(lambda: c1.bump_greeks(hm, bumps=hb, n_paths=n3, steps_per_year=365, std_errors=False))
        gB = c1.bump_greeks(hm, bumps=hb, n_paths=n3, steps_per_year=365)
        eq = {nm: (tB / tA) * (sb / sa) ** 2 for nm, sa, sb in zip(gA.names, gA.std_errors, gB.std_errors)}
        extras["config3_pathwise_ms"] = tA * 1e3
        extras["config3_pathwise_kernel_ms"] = kA_ms
        extras["config3_bump13_ms"] = tB * 1e3
        extras["config3_bump13_kernel_ms"] = kB_ms
        extras["config3_speedup_raw"] = tB / tA
        extras["config3_speedup_equal_stderr"] = eq
        extras["config3_pathwise"] = dict(zip(gA.names, gA.values))
        extras["config3_pathwise_se"] = dict(zip(gA.names, gA.std_errors))
        extras["config3_bump"] = dict(zip(gB.names, gB.values))
        extras["config3_bump_se"] = dict(zip(gB.names, gB.std_errors))
        # The same six Greeks with the variance-model derivatives carried
        # forward as tangents (no checkpoints, no reverse sweep).
        eng.set_option("pathwise_mode", 1)
        try:
            gT, tT, kT_ms = timed(lambda: c1.pathwise_greeks(hm, n_paths=n3, steps_per_year=365))
        finally:
            eng.set_option("pathwise_mode", 0)
        extras["config3_tangent_ms"] = tT * 1e3
        extras["config3_tangent_kernel_ms"] = kT_ms
        extras["config3_tangent_speedup_raw"] = tB / tT
        extras["config3_tangent_speedup_equal_stderr"] = {
            nm: (tB / tT) * (sb / sa) ** 2 for nm, sa, sb in zip(gT.names, gT.std_errors, gB.std_errors)}
        extras["config3_tangent_max_rel_diff"] = max(
            abs(a - b) / max(abs(b), 1e-12) for a, b in zip(gT.values, gA.values))
        # Config 4: the 1,000-contract book in one persistent launch vs one
        # launch per contract (the reference's loop of price_gbm calls).
        from kontracts_amd import book as B
        from kontracts_amd import tune as T
        jobs = B.book(1000)
        tuned = T.autotune_portfolio(B.book(200), repeats=1)
        extras["config4_autotune"] = {"best": list(tuned["best"]),
                                      "device_ms": {f"{k[0]}x{k[1]}": v for k, v in tuned["device_ms"].items()}}
        K.price_portfolio(jobs)
        t = time.perf_counter()
        rp = K.price_portfolio(jobs)
        tp = time.perf_counter() - t
        t = time.perf_counter()
        rs = [job[0].price(job[1], n_paths=job[2], seed=job[3], steps_per_year=job[4], grid=job[5]) for job in jobs]
        ts = time.perf_counter() - t
        extras["config4_path_steps"] = B.path_steps(jobs)
        extras["config4_portfolio_s"] = tp
        extras["config4_per_contract_s"] = ts
        extras["config4_speedup"] = ts / tp
        extras["config4_portfolio_path_steps_per_s"] = B.path_steps(jobs) / tp
        extras["config4_identical_to_per_contract"] = all(a.price == b.price and a.std_error == b.std_error
                                                          for a, b in zip(rp, rs))
        # the book with pathwise Greeks on its Lipschitz payoffs (European,
        # Asian, basket): persistent Greeks launches + the price-only launch
        # for the barriers, vs per contract pathwise Greeks / a price
        K.price_portfolio(jobs, greeks=True)
        t = time.perf_counter()
        rg = K.price_portfolio(jobs, greeks=True)
        tpg = time.perf_counter() - t
        t = time.perf_counter()
        sg = []
        for i, job in enumerate(jobs):
            kw = dict(n_paths=job[2], seed=job[3], steps_per_year=job[4], grid=job[5])
            if B.KINDS[i % 4] == "barrier":
                sg.append((job[0].price(job[1], **kw), None))
            else:
                sg.append((None, job[0].pathwise_greeks(job[1], **kw)))
        tsg = time.perf_counter() - t
        extras["config4_greeks_portfolio_s"] = tpg
        extras["config4_greeks_per_contract_s"] = tsg
        extras["config4_greeks_speedup"] = tsg / tpg
        extras["config4_greeks_identical_to_per_contract"] = all(
            (a[1] is None and a[0].price == b[0].price) or (a[1] is not None and a[1].values == b[1].values)
            for a, b in zip(rg, sg))
        # launch-bound regime: a book of the same kinds at 2^14 paths and 3-month maturities (63 steps)
        small = B.book(1000, paths_exp=(14, 14), maturity=(0.25, 0.25))
        K.price_portfolio(small)
        t = time.perf_counter()
        K.price_portfolio(small)
        tps = time.perf_counter() - t
        t = time.perf_counter()
        for job in small:
            job[0].price(job[1], n_paths=job[2], seed=job[3], steps_per_year=job[4], grid=job[5])
        tss = time.perf_counter() - t
        extras["config4_small_book_speedup"] = tss / tps
        extras["config4_small_book_portfolio_ms"] = tps * 1e3
        K.price_portfolio(small, greeks=True)
        t = time.perf_counter()
        K.price_portfolio(small, greeks=True)
        tpsg = time.perf_counter() - t
        t = time.perf_counter()
        for i, job in enumerate(small):
            kw = dict(n_paths=job[2], seed=job[3], steps_per_year=job[4], grid=job[5])
            if B.KINDS[i % 4] == "barrier":
                job[0].price(job[1], **kw)
            else:
                job[0].pathwise_greeks(job[1], **kw)
        tssg = time.perf_counter() - t
        extras["config4_small_book_greeks_speedup"] = tssg / tpsg
        extras["config4_small_book_greeks_portfolio_ms"] = tpsg * 1e3
        # Calibration-sized calls (1,000 paths x 50 steps): per-call latency of
        # a single price, and one objective evaluation over 10 quotes as one
        # portfolio call (direct launches, then graph replay) vs a loop.
        ko = K.up_and_out_call("X", 100.0, 130.0, 1.0)
        for _ in range(20):
            ko.price(K.GBM(100.0, 0.2, 0.05, asset="X"), n_paths=1000, seed=1)
        t = time.perf_counter()
        for i in range(200):
            ko.price(K.GBM(100.0, 0.2 + 1e-4 * i, 0.05, asset="X"), n_paths=1000, seed=1)
        extras["small_call_price_us"] = (time.perf_counter() - t) / 200 * 1e6
        quotes = [(ko, K.GBM(s, 0.2, 0.05, asset="X"), 1000, 1, 50) for s in range(90, 110, 2)]
        from kontracts_amd import _native as NN
        eng = NN.engine(0)
        for graph in (0, 1):
            eng.set_option("portfolio_graph", graph)
            for _ in range(10):
                K.price_portfolio(quotes)
            t = time.perf_counter()
            for _ in range(100):
                K.price_portfolio(quotes)
            extras[f"objective_10_quotes_portfolio_{'graph' if graph else 'direct'}_us"] = \
                (time.perf_counter() - t) / 100 * 1e6
        eng.set_option("portfolio_graph", 0)
        t = time.perf_counter()
        for _ in range(20):
            for q in quotes:
                q[0].price(q[1], n_paths=q[2], seed=q[3], steps_per_year=q[4])
        extras["objective_10_quotes_loop_us"] = (time.perf_counter() - t) / 20 * 1e6
        # Greek surface: 20 x 20 (spot, vol) cells of the config-2 barrier at
        # 2^16 paths x 252 steps, one launch vs a loop of greeks calls.
        spots = [150.0 + 2.5 * i for i in range(20)]
        vols = [0.15 + 0.01 * j for j in range(20)]
        bar = K.up_and_out_call("X", 100, 200, 1.0)
        K.greek_surface(bar, "X", spots[:2], vols[:2], n_paths=1 << 16, seed=42, steps_per_year=252)
        eng.set_profiling(True)
        eng.kernel_time(reset=True)
        t = time.perf_counter()
        K.greek_surface(bar, "X", spots, vols, n_paths=1 << 16, seed=42, steps_per_year=252)
        tsurf = time.perf_counter() - t
        ksurf_ms = eng.kernel_time(reset=True)[0]
        eng.set_profiling(False)
        extras["surface_20x20_kernel_ms"] = ksurf_ms
        t = time.perf_counter()
        for s_ in spots[:5]:
            for v_ in vols:
                bar.greeks(K.GBM(s_, v_, 0.05, asset="X"), n_paths=1 << 16, seed=42, steps_per_year=252)
        tloop = (time.perf_counter() - t) * 4  # 100 of the 400 cells, scaled
        extras["surface_20x20_s"] = tsurf
        extras["surface_20x20_loop_s"] = tloop
        extras["surface_scenario_path_steps_per_s"] = 400 * 7 * (1 << 16) * 252 / tsurf
        # Variance reduction on a daily-average Asian call (1e6 paths x 252 steps)
        call = K.scale((K.average(K.spot("X")) - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)  # Asian call
        g252 = [i / 252 for i in range(253)]
        call.price(m1, n_paths=1_000_000, seed=42, grid=g252)
        t = time.perf_counter()
        rv = call.price(m1, n_paths=1_000_000, seed=42, grid=g252)
        extras["vr_none_ms"] = (time.perf_counter() - t) * 1e3
        extras["vr_none_std_error"] = rv.std_error
        for label, anti, cv in (("antithetic", True, False), ("control_variate", False, True), ("both", True, True)):
            call.price(m1, n_paths=1_000_000, seed=42, grid=g252, antithetic=anti, control_variate=cv)
            t = time.perf_counter()
            rv = call.price(m1, n_paths=1_000_000, seed=42, grid=g252, antithetic=anti, control_variate=cv)
            extras[f"vr_{label}_ms"] = (time.perf_counter() - t) * 1e3
            extras[f"vr_{label}_std_error"] = rv.std_error

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 paths, f64 payoff/stderr reduction",
        "data": "synthetic (contract + model + seed; counter-based Philox normals generated on the GPU)",
        "config": {"workload": "up-and-out call (S-100)^+ KO at S>=200, GBM s0=180 sigma=0.25 r=0.05, "
                               f"{n} paths x 252 steps, CRN bump Greeks (7 scenario pricings)",
                   "paths": n, "steps": steps, "scenarios": 7, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "greeks": {"price": g.price, "delta": g.delta, "gamma": g.gamma, "vega": g.vega, "rho": g.rho,
                   "price_std_error": g.std_errors[0]},
    }
    if extras:
        line["extras"] = extras
    print(json.dumps(line))


if __name__ == "__main__":
    main()
