"""Deterministic synthetic book of uneven contracts for the batched-portfolio
workload: 1,000 contracts from one seed, 250 each of

* European calls,
* arithmetic-Asian calls with monthly fixings (``average_at``),
* up-and-out barrier calls,
* 5-asset basket calls (equal weights, pairwise correlation 0.3),

with strike K ~ U[80, 120] % of spot 100, volatility ~ U[0.15, 0.45],
maturity T ~ U[0.25, 3] years on a daily grid of n = ceil(252 T) steps
(63 .. 756), paths 2^e with e ~ U{16..20}, and the single-asset contracts
split 50/50 between GBM and Heston.  Knobs the workload leaves open are fixed
here:

* rate 0.05; every GBM drift equals the rate;
* Heston: v0 = theta = sigma^2, kappa = 2, sigma_v = 0.3, rho = -0.5;
* barrier H ~ U[1.15, 1.6] x spot (K < H), monitored at every grid point;
* monthly fixings: m = max(1, round(12 T)) dates, fixing i on the daily grid
  point nearest to i T / m (the last one at T);
* basket asset vols drawn independently from the same range;
* job seed = 42 + job index.

Every job carries its explicit daily grid.  ``book(n)`` returns the first n
jobs of the same sequence, so smaller books are prefixes of the full one.
"""
from __future__ import annotations

import math

import numpy as np

from . import dsl

SPOT = 100.0
RATE = 0.05
KINDS = ("european", "asian", "barrier", "basket")


def _daily_grid(t: float) -> tuple[np.ndarray, int]:
    """Points t k / n, k = 0 .. n, as float64 (passed to the engine without a
    copy)."""
    n = int(math.ceil(252.0 * t - 1e-9))
    n = max(n, 1)
    return np.array([t * k / n for k in range(n + 1)], dtype=np.float64), n


def book(n_contracts: int = 1000, seed: int = 42, paths_exp: tuple = (16, 20),
         maturity: tuple = (0.25, 3.0)) -> list[tuple]:
    """Jobs (contract, model, n_paths, seed, steps_per_year, grid) for
    ``price_portfolio``; the kinds cycle European, Asian, barrier, basket.
    ``paths_exp`` and ``maturity`` bound the path-count exponent and the
    maturity (the defaults are the workload's)."""
    rng = np.random.default_rng(seed)
    jobs = []
    corr5 = [[1.0 if i == j else 0.3 for j in range(5)] for i in range(5)]
    for j in range(n_contracts):
        kind = KINDS[j % 4]
        strike = SPOT * rng.uniform(0.8, 1.2)
        vol = rng.uniform(0.15, 0.45)
        t = rng.uniform(maturity[0], maturity[1])
        n_paths = 1 << int(rng.integers(paths_exp[0], paths_exp[1] + 1))
        heston = bool(rng.integers(0, 2)) and kind != "basket"
        h_level = SPOT * rng.uniform(1.15, 1.6)
        basket_vols = rng.uniform(0.15, 0.45, size=5)
        grid, n = _daily_grid(t)
        asset = "X"
        if kind == "basket":
            names = [f"B{i}" for i in range(5)]
            factors = [dsl.GbmFactor(nm, SPOT, RATE, float(v)) for nm, v in zip(names, basket_vols)]
            model = dsl.correlated_gbm(factors, corr5, RATE)
            avg = sum((dsl.spot(nm) for nm in names[1:]), dsl.spot(names[0])) / 5.0
            contract = dsl.scale((avg - strike).clip(0.0), dsl.one("USD")) @ dsl.at(t)
        else:
            if heston:
                v2 = vol * vol
                model = dsl.heston(SPOT, v2, 2.0, v2, 0.3, -0.5, RATE, asset=asset)
            else:
                model = dsl.GBM(SPOT, vol, RATE, asset=asset)
            if kind == "european":
                contract = dsl.european_call(asset, strike, t)
            elif kind == "barrier":
                contract = dsl.up_and_out_call(asset, min(strike, 0.95 * h_level), h_level, t)
            else:
                m = max(1, int(round(12.0 * t)))
                fix = sorted({float(grid[min(n, max(1, int(round(i * n / m))))]) for i in range(1, m + 1)})
                contract = dsl.scale((dsl.average_at(dsl.spot(asset), fix) - strike).clip(0.0),
                                     dsl.one("USD")) @ dsl.at(t)
        jobs.append((contract, model, n_paths, seed + j, 252, grid))
    return jobs


def path_steps(jobs) -> float:
    """Total paths x grid steps of a book."""
    return float(sum(job[2] * (len(job[5]) - 1) for job in jobs))


def job_profile(index: int, job) -> tuple:
    """(kind, model, grid steps, fixings) of job `index` of a book: kind as in
    KINDS, model "gbm", "heston" or "basket5", fixings the number of
    AverageAt dates (0 for the other kinds)."""
    import json

    con, model = job[0], job[1]
    kind = KINDS[index % 4]
    if isinstance(model, dsl.GBM):
        name = "gbm"
    elif model.kind == 1:
        name = "heston"
    else:
        name = "basket5"

    def fixings(node):
        if isinstance(node, dict):
            if "AverageAt" in node:
                return len(node["AverageAt"]["dates"])
            return sum(fixings(v) for v in node.values())
        if isinstance(node, list):
            return sum(fixings(v) for v in node)
        return 0

    return kind, name, len(job[5]) - 1, fixings(json.loads(con.to_json()))
