"""Autotuner for the persistent portfolio launches.

The tile knobs of a book's launch:

* the grid: resident workgroups per compute unit (0 = the kernel's
  occupancy);
* the queue chunk: work items taken per fetch;
* the register budget of the persistent kernels, compiled as variants:
  the price kernel with the compiler's budget or capped for 6 or 8 waves per
  SIMD (``portfolio_waves``), the Greeks kernels with each variant's default
  budget or capped for 5 waves (``portfolio_adjoint_waves``);
* the step unroll of the price kernel (compiled variants): one or two
  4-point groups per iteration of its uniform-run loops
  (``portfolio_unroll``: 4 or 8 grid points per iteration);
* the block size and paths per lane of the price kernel (compiled
  variants): 256 lanes with one path each, or 128 lanes with two paths each
  of the same 256-path block (``portfolio_lanes``).

``autotune_portfolio`` times the candidates on the caller's own book with the
engine's event timer (device time of the launch sequence) in three rounds:
the (grid, chunk) pairs under the default tile, then the register budgets
under the best pair, then the unroll and lanes variants under the best of
those; it leaves the fastest set on the engine.  Results never depend on the
choice: each job's reduction order is fixed.
"""
from __future__ import annotations

from typing import Sequence

from . import _native as N
from . import dsl

DEFAULT_CANDIDATES = ((0, 1), (0, 4), (0, 16), (8, 1), (4, 1), (4, 4))
PRICE_WAVES = (0, 6, 8)
ADJOINT_WAVES = (0, 5)
UNROLL_LANES = ((1, 256), (2, 256), (1, 128), (2, 128))

# a tile: (blocks_per_cu, chunk, price_waves, adjoint_waves, unroll, lanes)
_OPTIONS = ("portfolio_blocks_per_cu", "portfolio_chunk", "portfolio_waves", "portfolio_adjoint_waves",
            "portfolio_unroll", "portfolio_lanes")
DEFAULT_TILE = (0, 1, 0, 0, 1, 256)


def set_tile(eng, tile) -> None:
    """Sets every knob of a tile tuple on the engine."""
    for name, value in zip(_OPTIONS, tile):
        eng.set_option(name, int(value))


def tile_name(tile) -> str:
    bpc, chunk, pw, aw, unroll, lanes = tile
    return f"{bpc}x{chunk}/w{pw}/aw{aw}/u{unroll}/l{lanes}"


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
                       device: int | None = None, greeks: bool = False, budgets: bool = True,
                       unroll_lanes=UNROLL_LANES) -> dict:
    """Returns {"best": (blocks_per_cu, chunk, price_waves, adjoint_waves,
    unroll, lanes), "device_ms": {tile_name: ms}} and sets the best tile on
    the engine.  greeks: time the book's call with pathwise Greeks
    (price_portfolio(..., greeks=True)), whose persistent Greeks launches
    take the grid, chunk and their own register budget.  budgets: also try
    the register-budget variants.  unroll_lanes: the (unroll, lanes)
    variants of the price kernel tried in the last round."""
    eng = N.engine(device)
    jobs = list(jobs)
    timings = {}

    def trial(tile):
        set_tile(eng, tile)
        timings[tuple(tile)] = _time(eng, jobs, device, greeks, repeats)

    for bpc, chunk in candidates:
        trial((bpc, chunk) + DEFAULT_TILE[2:])
    best = min(timings, key=timings.get)
    if budgets:
        for pw in PRICE_WAVES:
            for aw in (ADJOINT_WAVES if greeks else (0,)):
                if (pw, aw) != (best[2], best[3]):
                    trial(best[:2] + (pw, aw) + best[4:])
        best = min(timings, key=timings.get)
    for unroll, lanes in unroll_lanes:
        if (unroll, lanes) != (best[4], best[5]):
            trial(best[:4] + (unroll, lanes))
    best = min(timings, key=timings.get)
    set_tile(eng, best)
    return {"best": best, "device_ms": {tile_name(k): v for k, v in timings.items()}}
