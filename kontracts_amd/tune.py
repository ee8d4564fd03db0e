"""Autotuner for the persistent portfolio launch.

The launch's runtime tile knobs are the resident blocks per compute unit
(the grid) and the work items taken per queue fetch.  ``autotune_portfolio``
times each candidate on the caller's own book with the engine's event timer
(device time of the launch sequence, after one call that captures its graph)
and leaves the fastest one set on the engine.  Results never depend on the
choice: each job's reduction order is fixed.
"""
from __future__ import annotations

from typing import Sequence

from . import _native as N
from . import dsl

DEFAULT_CANDIDATES = ((0, 1), (0, 4), (0, 16), (8, 1), (4, 1), (4, 4))


def autotune_portfolio(jobs: Sequence[tuple], candidates=DEFAULT_CANDIDATES, repeats: int = 2,
                       device: int | None = None, greeks: bool = False) -> dict:
    """Returns {"best": (blocks_per_cu, chunk), "device_ms": {"bpcxchunk": ms}}
    and sets the best candidate on the engine (blocks_per_cu 0 = occupancy).
    greeks: time the book's call with pathwise Greeks (price_portfolio(...,
    greeks=True)), whose persistent Greeks launches take the same knobs."""
    eng = N.engine(device)
    jobs = list(jobs)
    timings = {}
    for bpc, chunk in candidates:
        eng.set_option("portfolio_blocks_per_cu", bpc)
        eng.set_option("portfolio_chunk", chunk)
        dsl.price_portfolio(jobs, device=device, greeks=greeks)
        eng.set_profiling(True)
        eng.kernel_time(reset=True)
        for _ in range(max(1, repeats)):
            dsl.price_portfolio(jobs, device=device, greeks=greeks)
        ms = eng.kernel_time(reset=True)[0] / max(1, repeats)
        eng.set_profiling(False)
        timings[(bpc, chunk)] = ms
    best = min(timings, key=timings.get)
    eng.set_option("portfolio_blocks_per_cu", best[0])
    eng.set_option("portfolio_chunk", best[1])
    return {"best": best, "device_ms": {f"{k[0]}x{k[1]}": v for k, v in timings.items()}}
