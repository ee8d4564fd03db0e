"""Greek surfaces over a (spot x vol) grid (src/surface.rs).

``greek_surface`` evaluates greeks_gbm at every cell in one GPU launch
(kt_greek_surface); ``GreekSurface`` holds the [spot][vol] matrices and
exports them as the reference does: CSV with a ``spot\\vol`` header
(src/surface.rs:38-61) and an ASCII PGM (P2) grayscale image normalised
min -> max (src/surface.rs:63-95).  Numbers are written the way Rust's
``Display`` writes an f64 (shortest round-trip digits, never an exponent,
no ``.0`` on integral values), so the text matches the reference's.
"""
from __future__ import annotations

import ctypes as C
import math
from decimal import Decimal
from typing import Sequence

import numpy as np

from . import _native as N
from .dsl import Contract, _cfg


def rust_display(x: float) -> str:
    """f64 as Rust's ``format!("{}", x)`` writes it."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    r = repr(x)  # shortest round-trip digits, as Rust's Display
    if "e" in r or "E" in r:
        r = format(Decimal(r), "f")
    if r.endswith(".0"):
        r = r[:-2]
    return r


def _round_half_away(x: float) -> int:
    """f64::round: halves go away from zero (Python's round goes to even)."""
    a = abs(x)
    r = math.floor(a)
    if a - r >= 0.5:
        r += 1
    return int(r) if x >= 0 else -int(r)


class GreekSurface:
    """src/surface.rs:17-35: matrices of shape [len(spots), len(vols)], row i
    for spots[i], column j for vols[j].  ``rho`` and ``std_errors`` (a dict of
    matrices per quantity) go beyond the reference's fields."""

    def __init__(self, spots, vols, price, delta, gamma, vega, rho=None, std_errors=None):
        self.spots = [float(s) for s in spots]
        self.vols = [float(v) for v in vols]
        self.price = np.asarray(price, dtype=np.float64)
        self.delta = np.asarray(delta, dtype=np.float64)
        self.gamma = np.asarray(gamma, dtype=np.float64)
        self.vega = np.asarray(vega, dtype=np.float64)
        self.rho = None if rho is None else np.asarray(rho, dtype=np.float64)
        self.std_errors = std_errors or {}

    def to_csv(self, quantity) -> str:
        """Header ``spot\\vol,<vols>``, then one line per spot."""
        q = np.asarray(quantity, dtype=np.float64)
        head = "spot\\vol" + "".join("," + rust_display(v) for v in self.vols) + "\n"
        rows = "".join(rust_display(s) + "".join("," + rust_display(q[i, j]) for j in range(len(self.vols))) + "\n"
                       for i, s in enumerate(self.spots))
        return head + rows

    def to_pgm(self, quantity) -> str:
        """P2 grayscale, cols x rows, values scaled to 0..255 from min to max
        (NaN ignored by the range, drawn as 0; a flat matrix spans 1)."""
        q = np.asarray(quantity, dtype=np.float64)
        rows, cols = q.shape
        lo, hi = math.inf, -math.inf
        for x in q.flat:  # f64::min / f64::max ignore NaN
            if not math.isnan(x):
                lo, hi = min(lo, x), max(hi, x)
        span = hi - lo
        if abs(span) < 1e-300:
            span = 1.0
        out = [f"P2\n{cols} {rows}\n255\n"]
        for i in range(rows):
            line = []
            for j in range(cols):
                g = (q[i, j] - lo) / span * 255.0
                level = 0 if math.isnan(g) else _round_half_away(g)  # `as i32` of NaN is 0
                line.append(f"{min(max(level, 0), 255)} ")
            out.append("".join(line) + "\n")
        return "".join(out)

    def __repr__(self):
        return f"GreekSurface({len(self.spots)} spots x {len(self.vols)} vols)"


def greek_surface(contract: Contract, asset: str, spots: Sequence[float], vols: Sequence[float],
                  n_paths: int = 200_000, seed: int = 42, steps_per_year: int = 50, rate: float = 0.05,
                  bumps: tuple[float, float, float] | None = None, antithetic: bool = False,
                  control_variate: bool = False, device: int | None = None) -> GreekSurface:
    """greek_surface(contract, asset, spots, vols, cfg, bumps)
    (src/surface.rs:104-138): greeks_gbm at every (spot, vol), drift = rate,
    every cell at ``seed``."""
    spots = [float(s) for s in spots]
    vols = [float(v) for v in vols]
    ns, nv = len(spots), len(vols)
    cfg = _cfg(n_paths, seed, steps_per_year, rate, None, antithetic, control_variate)
    b = N.KtBumps(*(bumps or (1e-2, 1e-2, 1e-4)))
    sp = (C.c_double * max(ns, 1))(*spots)
    vo = (C.c_double * max(nv, 1))(*vols)
    out = (N.KtGreeks * max(ns * nv, 1))()
    eng = N.engine(device)
    N.check(N.lib().kt_greek_surface(eng.handle, contract.to_json().encode(), asset.encode(), sp, ns, vo, nv,
                                     C.byref(cfg.s), C.byref(b), out), eng.handle)
    cells = out[: ns * nv]

    def mat(f):
        return np.array([f(g) for g in cells], dtype=np.float64).reshape(ns, nv)

    names = ("price", "delta", "gamma", "vega", "rho")
    se = {n: mat(lambda g, k=k: g.std_error[k]) for k, n in enumerate(names)}
    return GreekSurface(spots, vols, mat(lambda g: g.price), mat(lambda g: g.delta), mat(lambda g: g.gamma),
                        mat(lambda g: g.vega), mat(lambda g: g.rho), se)

