"""Autotuner for the persistent portfolio launches.

The tile knobs of a book's launch:

* the grid: resident blocks per compute unit (0 = the kernel's occupancy);
* the queue chunk: work items taken per fetch;
* the register budget of the persistent kernels, compiled as variants:
  the price kernel with the compiler's budget or capped for 6 or 8 waves per
  SIMD (``portfolio_waves``), the Greeks kernels with each variant's default
  budget or capped for 5 waves (``portfolio_adjoint_waves``).

``autotune_portfolio`` times the candidates on the caller's own book with the
engine's event timer (device time of the launch sequence) in two rounds: the
(grid, chunk) pairs under the default budgets, then the register budgets
under the best pair; it leaves the fastest set on the engine.  Results never
depend on the choice: each job's reduction order is fixed.
"""
from __future__ import annotations

from typing import Sequence

from . import _native as N
from . import dsl

DEFAULT_CANDIDATES = ((0, 1), (0, 4), (0, 16), (8, 1), (4, 1), (4, 4))
PRICE_WAVES = (0, 6, 8)
ADJOINT_WAVES = (0, 5)


def _time(eng, jobs, device, greeks, repeats):
    dsl.price_portfolio(jobs, device=device, greeks=greeks)
    eng.set_profiling(True)
    eng.kernel_time(reset=True)
    for _ in range(max(1, repeats)):
        dsl.price_portfolio(jobs, device=device, greeks=greeks)
    ms = eng.kernel_time(reset=True)[0] / max(1, repeats)
    eng.set_profiling(False)
    return ms


def autotune_portfolio(jobs: Sequence[tuple], candidates=DEFAULT_CANDIDATES, repeats: int = 2,
                       device: int | None = None, greeks: bool = False, budgets: bool = True) -> dict:
    """Returns {"best": (blocks_per_cu, chunk, price_waves, adjoint_waves),
    "device_ms": {"bpc x chunk / waves": ms}} and sets the best tile on the
    engine.  greeks: time the book's call with pathwise Greeks
    (price_portfolio(..., greeks=True)), whose persistent Greeks launches
    take the same knobs.  budgets: also try the register-budget variants."""
    eng = N.engine(device)
    jobs = list(jobs)
    timings = {}
    eng.set_option("portfolio_waves", 0)
    eng.set_option("portfolio_adjoint_waves", 0)
    for bpc, chunk in candidates:
        eng.set_option("portfolio_blocks_per_cu", bpc)
        eng.set_option("portfolio_chunk", chunk)
        timings[(bpc, chunk, 0, 0)] = _time(eng, jobs, device, greeks, repeats)
    best = min(timings, key=timings.get)
    eng.set_option("portfolio_blocks_per_cu", best[0])
    eng.set_option("portfolio_chunk", best[1])
    if budgets:
        for pw in PRICE_WAVES:
            for aw in (ADJOINT_WAVES if greeks else (0,)):
                if (pw, aw) == (0, 0):
                    continue
                eng.set_option("portfolio_waves", pw)
                eng.set_option("portfolio_adjoint_waves", aw)
                timings[(best[0], best[1], pw, aw)] = _time(eng, jobs, device, greeks, repeats)
        best = min(timings, key=timings.get)
    eng.set_option("portfolio_blocks_per_cu", best[0])
    eng.set_option("portfolio_chunk", best[1])
    eng.set_option("portfolio_waves", best[2])
    eng.set_option("portfolio_adjoint_waves", best[3])
    return {"best": best,
            "device_ms": {f"{k[0]}x{k[1]}/w{k[2]}/aw{k[3]}": v for k, v in timings.items()}}
