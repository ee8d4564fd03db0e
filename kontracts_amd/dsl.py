"""Contract algebra DSL and the pricing API, mirroring the reference's Python
package (python/kontract/_kontract.pyi:11-240, src/bindings.rs:55-472).

Contracts are plain trees; ``to_json`` writes serde's externally tagged form
(src/ast.rs:210-218), which is what crosses the C ABI.  Pricing goes to the
GPU engine through ``kontracts_amd._native``.
"""
from __future__ import annotations

import array
import ctypes as C
import json
import math
import threading
from typing import Iterable, Sequence

import numpy as np

from . import _native as N

USD = "USD"
EUR = "EUR"
GBP = "GBP"
JPY = "JPY"

__version__ = "0.1.0"


# ---------------------------------------------------------------------------
# JSON (serde_json writes floats with ryu: shortest round-trip digits,
# exponents without '+' or leading zeros)
# ---------------------------------------------------------------------------
def _fmt_float(x: float) -> str:
    x = float(x)
    if not math.isfinite(x):
        raise ValueError("non-finite numbers have no JSON form")
    s = repr(x)
    if "e" in s:
        mant, exp = s.split("e")
        sign = "-" if exp.startswith("-") else ""
        digits = exp.lstrip("+-").lstrip("0") or "0"
        if "." in mant and mant.endswith(".0"):
            mant = mant[:-2]
        return f"{mant}e{sign}{digits}"
    return s


def _dump(node) -> str:
    if isinstance(node, dict):
        return "{" + ",".join(json.dumps(k) + ":" + _dump(v) for k, v in node.items()) + "}"
    if isinstance(node, list):
        return "[" + ",".join(_dump(v) for v in node) + "]"
    if isinstance(node, bool):
        return "true" if node else "false"
    if isinstance(node, float):
        return _fmt_float(node)
    if node is None:
        return "null"
    if isinstance(node, str):
        return json.dumps(node, ensure_ascii=False)
    if isinstance(node, int):
        return _fmt_float(float(node))
    raise TypeError(f"cannot serialise {type(node)}")


def _floats_as_float(node):
    """json.loads gives ints for integral numbers; the AST holds f64."""
    if isinstance(node, dict):
        return {k: _floats_as_float(v) for k, v in node.items()}
    if isinstance(node, list):
        return [_floats_as_float(v) for v in node]
    if isinstance(node, int) and not isinstance(node, bool):
        return float(node)
    return node


# ---------------------------------------------------------------------------
# Observable / Condition / Contract
# ---------------------------------------------------------------------------
def _obs(x) -> "Observable":
    if isinstance(x, Observable):
        return x
    if isinstance(x, (int, float)) and not isinstance(x, bool):
        return Observable({"Const": float(x)})
    raise TypeError("expected a number or an Observable")


class Observable:
    """A real-valued process (src/ast.rs:34-71)."""

    __slots__ = ("node",)
    __array_priority__ = 1000

    def __init__(self, node):
        self.node = node

    # payoff helpers
    def clip(self, floor: float) -> "Observable":
        return self.max(float(floor))

    def max(self, other) -> "Observable":
        return Observable({"Max": [self.node, _obs(other).node]})

    def min(self, other) -> "Observable":
        return Observable({"Min": [self.node, _obs(other).node]})

    # temporal observables (src/ast.rs:57-70)
    def average(self) -> "Observable":
        return average(self)

    def average_over(self, from_year: float, to_year: float) -> "Observable":
        return average_over(self, from_year, to_year)

    def running_max(self) -> "Observable":
        return running_max(self)

    def running_min(self) -> "Observable":
        return running_min(self)

    # pipeline style
    def add(self, o):
        return self + o

    def sub(self, o):
        return self - o

    def mul(self, o):
        return Observable({"Mul": [self.node, _obs(o).node]})

    def div(self, o):
        return self / o

    def scale(self, contract: "Contract") -> "Contract":
        return Contract({"Scale": [self.node, contract.node]})

    def __add__(self, o):
        return Observable({"Add": [self.node, _obs(o).node]})

    def __radd__(self, o):
        return Observable({"Add": [_obs(o).node, self.node]})

    def __sub__(self, o):
        return Observable({"Sub": [self.node, _obs(o).node]})

    def __rsub__(self, o):
        return Observable({"Sub": [_obs(o).node, self.node]})

    def __truediv__(self, o):
        return Observable({"Div": [self.node, _obs(o).node]})

    def __rtruediv__(self, o):
        return Observable({"Div": [_obs(o).node, self.node]})

    def __neg__(self):
        return Observable({"Neg": self.node})

    def __mul__(self, o):
        if isinstance(o, Contract):
            return Contract({"Scale": [self.node, o.node]})
        return Observable({"Mul": [self.node, _obs(o).node]})

    def __rmul__(self, o):
        return Observable({"Mul": [_obs(o).node, self.node]})

    def __ge__(self, o):
        return Condition({"Ge": [self.node, _obs(o).node]})

    def __gt__(self, o):
        return Condition({"Gt": [self.node, _obs(o).node]})

    def __le__(self, o):
        return Condition({"Le": [self.node, _obs(o).node]})

    def __lt__(self, o):
        return Condition({"Lt": [self.node, _obs(o).node]})

    def __eq__(self, other):
        return isinstance(other, Observable) and other.node == self.node

    __hash__ = None

    def __repr__(self):
        return f"Observable({_dump(self.node)})"


class Condition:
    """A boolean process (src/ast.rs:142-161)."""

    __slots__ = ("node",)

    def __init__(self, node):
        self.node = node

    def __and__(self, o: "Condition"):
        return Condition({"And": [self.node, o.node]})

    def __or__(self, o: "Condition"):
        return Condition({"Or": [self.node, o.node]})

    def __invert__(self):
        return Condition({"Not": self.node})

    def __eq__(self, other):
        return isinstance(other, Condition) and other.node == self.node

    __hash__ = None

    def __repr__(self):
        return f"Condition({_dump(self.node)})"


class PriceResult:
    """src/pricer.rs:58-71."""

    __slots__ = ("price", "std_error", "sample_std", "ci95_low", "ci95_high", "n_paths")

    def __init__(self, r: N.KtPriceResult):
        self.price = r.price
        self.std_error = r.std_error
        self.sample_std = r.sample_std
        self.ci95_low = r.ci95_low
        self.ci95_high = r.ci95_high
        self.n_paths = int(r.n_paths)

    @classmethod
    def _from_fields(cls, t) -> "PriceResult":
        """From a kt_price_result's fields in declaration order (price,
        sample_std, std_error, ci95_low, ci95_high, n_paths)."""
        r = cls.__new__(cls)
        r.price, r.sample_std, r.std_error, r.ci95_low, r.ci95_high = t[0], t[1], t[2], t[3], t[4]
        r.n_paths = int(t[5])
        return r

    def paths_for_tolerance(self, tol: float) -> int:
        """src/pricer.rs:81-86."""
        if tol <= 0.0 or self.sample_std == 0.0:
            return 0
        return int(math.ceil((1.959963984540054 * self.sample_std / tol) ** 2))

    def __repr__(self):
        return (f"PriceResult(price={self.price:.6f}, std_error={self.std_error:.6f}, "
                f"ci95=[{self.ci95_low:.6f}, {self.ci95_high:.6f}], n_paths={self.n_paths})")


class Greeks:
    """src/greeks.rs:43-54, plus the Monte Carlo standard error of each value."""

    __slots__ = ("price", "delta", "gamma", "vega", "rho", "std_errors")

    def __init__(self, g: N.KtGreeks):
        self.price, self.delta, self.gamma, self.vega, self.rho = g.price, g.delta, g.gamma, g.vega, g.rho
        self.std_errors = list(g.std_error)

    def __repr__(self):
        return (f"Greeks(price={self.price:.6f}, delta={self.delta:.6f}, gamma={self.gamma:.6f}, "
                f"vega={self.vega:.6f}, rho={self.rho:.6f})")


class ModelGreeks:
    """Price and its derivatives with respect to the model's parameters, each
    with its Monte Carlo standard error.  ``names`` lists the parameters:
    GBM (s0, sigma, mu), Heston (s0, v0, kappa, theta, sigma_v, rho),
    correlated GBM (s0, sigma, mu of each asset)."""

    __slots__ = ("price", "price_std_error", "names", "values", "std_errors", "method")

    def __init__(self, names, out, se, method):
        self.price, self.price_std_error = out[0], se[0]
        self.names = list(names)
        self.values = list(out[1:1 + len(names)])
        self.std_errors = list(se[1:1 + len(names)])
        self.method = method

    @classmethod
    def _from_rows(cls, names, price, price_std_error, values, std_errors, method) -> "ModelGreeks":
        """From the price, its error and fresh lists of the derivatives and
        their errors (taken as they are, not copied)."""
        g = cls.__new__(cls)
        g.price, g.price_std_error = price, price_std_error
        g.names = list(names)
        g.values, g.std_errors, g.method = values, std_errors, method
        return g

    def __getitem__(self, name: str) -> float:
        return self.values[self.names.index(name)]

    def as_dict(self) -> dict:
        return dict(zip(self.names, self.values))

    def __repr__(self):
        body = ", ".join(f"{n}={v:.6f}" for n, v in zip(self.names, self.values))
        return f"ModelGreeks({self.method}: price={self.price:.6f}, {body})"


_GBM_NAMES = ("s0", "sigma", "mu")
_HESTON_NAMES = ("s0", "v0", "kappa", "theta", "sigma_v", "rho")
_BASKET_NAMES: dict = {}  # factor names -> parameter names


def _param_names(model) -> tuple:
    if isinstance(model, GBM):
        return _GBM_NAMES
    if isinstance(model, Model) and model.kind == N.KT_MODEL_HESTON:
        return _HESTON_NAMES
    if isinstance(model, Model) and model.kind == N.KT_MODEL_CORRELATED_GBM:
        fs = tuple(f.name for f in model.params["factors"])
        names = _BASKET_NAMES.get(fs)
        if names is None:
            names = tuple([f"s0[{n}]" for n in fs] + [f"sigma[{n}]" for n in fs] + [f"mu[{n}]" for n in fs])
            if len(_BASKET_NAMES) < 4096:
                _BASKET_NAMES[fs] = names
        return names
    raise TypeError("expected a GBM or Model")


class Contract:
    """Peyton-Jones combinators (src/ast.rs:187-206).  A contract is an
    immutable value: its serde JSON is computed once and kept."""

    __slots__ = ("node", "_json")

    def __init__(self, node):
        self.node = node
        self._json = None

    def __matmul__(self, cond: Condition) -> "Contract":
        return Contract({"When": [cond.node, self.node]})

    def when(self, cond: Condition) -> "Contract":
        return self @ cond

    def __add__(self, other: "Contract") -> "Contract":
        return Contract({"And": [self.node, other.node]})

    def __neg__(self) -> "Contract":
        return Contract({"Give": self.node})

    def until(self, cond: Condition) -> "Contract":
        return Contract({"Until": [cond.node, self.node]})

    def anytime(self, cond: Condition) -> "Contract":
        return Contract({"Anytime": [cond.node, self.node]})

    def or_(self, other: "Contract") -> "Contract":
        return Contract({"Or": [self.node, other.node]})

    def give(self) -> "Contract":
        return -self

    def to_json(self) -> str:
        if self._json is None:
            self._json = _dump(self.node)
        return self._json

    @staticmethod
    def from_json(s: str) -> "Contract":
        try:
            node = json.loads(s)
        except json.JSONDecodeError as e:
            raise N.KontractError(5, str(e)) from None
        node = _floats_as_float(node)
        # Validate the shape with the engine's own reader (host only); only
        # deserialisation errors matter here, as in Contract::from_json.
        try:
            plan(Contract(node))
        except N.KontractError as e:
            if e.code == 5:
                raise
        return Contract(node)

    def __eq__(self, other):
        return isinstance(other, Contract) and other.node == self.node

    __hash__ = None

    def __repr__(self):
        return f"Contract({self.to_json()})"

    # --- pricing (src/bindings.rs:350-467) --------------------------------
    def price(self, model, n_paths: int = 100_000, seed: int = 42, steps_per_year: int = 50,
              antithetic: bool = False, control_variate: bool = False, grid: Sequence[float] | None = None,
              device: int | None = None) -> PriceResult:
        m = _model_struct(model)
        cfg = _cfg(n_paths, seed, steps_per_year, model.rate, grid, antithetic, control_variate)
        eng = N.engine(device)
        out = N.KtPriceResult()
        N.check(N.lib().kt_price(eng.handle, self.to_json().encode(), C.byref(m.s), C.byref(cfg.s),
                                 C.byref(out)), eng.handle)
        return PriceResult(out)

    def greeks(self, model: "GBM", n_paths: int = 200_000, seed: int = 42, steps_per_year: int = 50,
               bumps: tuple[float, float, float] | None = None, device: int | None = None,
               antithetic: bool = False, control_variate: bool = False) -> Greeks:
        """greeks_gbm (src/greeks.rs:60-98).  antithetic / control_variate
        make every repricing variance-reduced, as a McConfig carrying them
        does in the reference."""
        if not isinstance(model, GBM):
            raise TypeError("greeks expects a GBM model")
        cfg = _cfg(n_paths, seed, steps_per_year, model.r, None, antithetic, control_variate)
        b = N.KtBumps(*(bumps or (1e-2, 1e-2, 1e-4)))
        eng = N.engine(device)
        out = N.KtGreeks()
        N.check(N.lib().kt_greeks_gbm(eng.handle, self.to_json().encode(), model.asset.encode(), model.s0,
                                      model.sigma, C.byref(cfg.s), C.byref(b), C.byref(out)), eng.handle)
        return Greeks(out)

    def pathwise_greeks(self, model, n_paths: int = 200_000, seed: int = 42, steps_per_year: int = 50,
                        grid: Sequence[float] | None = None, device: int | None = None) -> ModelGreeks:
        """Pathwise (adjoint) derivatives of the price with respect to every
        model parameter, from one GPU launch.  Lipschitz payoffs paid on one
        date only (European, Asian, lookback, basket); others raise
        KontractError(Unsupported): use ``bump_greeks`` for them."""
        names = _param_names(model)
        m = _model_struct(model)
        cfg = _cfg(n_paths, seed, steps_per_year, model.rate, grid)
        out = (C.c_double * (1 + len(names)))()
        se = (C.c_double * (1 + len(names)))()
        eng = N.engine(device)
        N.check(N.lib().kt_greeks_adjoint(eng.handle, self.to_json().encode(), C.byref(m.s), C.byref(cfg.s), out,
                                          se, len(out)), eng.handle)
        return ModelGreeks(names, list(out), list(se), "pathwise")

    def bump_greeks(self, model, bumps: Sequence[float] | None = None, n_paths: int = 200_000, seed: int = 42,
                    steps_per_year: int = 50, grid: Sequence[float] | None = None, device: int | None = None,
                    std_errors: bool = True) -> ModelGreeks:
        """Central finite differences over the model parameters with common
        random numbers (the src/greeks.rs pattern), GBM or Heston.  Default
        bumps: 1% of spot, then 0.01 (GBM sigma), 1e-4 (mu); Heston 1e-3 (v0),
        1e-2 (kappa), 1e-3 (theta), 1e-2 (sigma_v), 1e-2 (rho)."""
        names = _param_names(model)
        if bumps is None:
            s0 = model.s0 if isinstance(model, GBM) else model.params.get("spot", 1.0)
            bumps = [0.01 * s0, 0.01, 1e-4] if isinstance(model, GBM) else [0.01 * s0, 1e-3, 1e-2, 1e-3, 1e-2, 1e-2]
        m = _model_struct(model)
        cfg = _cfg(n_paths, seed, steps_per_year, model.rate, grid)
        h = _dbl(bumps)
        out = (C.c_double * (1 + len(names)))()
        se = (C.c_double * (1 + len(names)))()
        eng = N.engine(device)
        N.check(N.lib().kt_greeks_bump(eng.handle, self.to_json().encode(), C.byref(m.s), C.byref(cfg.s), h,
                                       len(bumps), out, se if std_errors else None), eng.handle)
        return ModelGreeks(names, list(out), list(se) if std_errors else [math.nan] * len(out), "bump")

    def price_american(self, model, exercise_dates: Sequence[float], n_paths: int = 100_000, seed: int = 42,
                       n_basis: int = 3, device: int | None = None) -> PriceResult:
        m = _model_struct(model)
        cfg = _cfg(n_paths, seed, 50, model.rate, None)
        dates = (C.c_double * max(len(exercise_dates), 1))(*[float(d) for d in exercise_dates])
        eng = N.engine(device)
        out = N.KtPriceResult()
        N.check(N.lib().kt_price_american(eng.handle, self.to_json().encode(), dates, len(exercise_dates),
                                          C.byref(m.s), C.byref(cfg.s), int(n_basis), C.byref(out)), eng.handle)
        return PriceResult(out)


# ---------------------------------------------------------------------------
# Constructors (src/ast.rs:226-448, src/bindings.rs:696-790)
# ---------------------------------------------------------------------------
def zero() -> Contract:
    return Contract("Zero")


def one(ccy: str) -> Contract:
    return Contract({"One": ccy})


def give(c: Contract) -> Contract:
    return -c


def spot(name: str) -> Observable:
    return Observable({"Spot": name})


S = spot


def const_(x: float) -> Observable:
    return Observable({"Const": float(x)})


konst = const_


def at(t: float) -> Condition:
    return Condition({"At": float(t)})


def bool_(b: bool) -> Condition:
    return Condition({"Bool": bool(b)})


def average(obs: Observable) -> Observable:
    return Observable({"Average": {"obs": _obs(obs).node, "from_year": None, "to_year": None}})


def average_over(obs: Observable, from_year: float, to_year: float) -> Observable:
    return Observable({"Average": {"obs": _obs(obs).node, "from_year": float(from_year),
                                   "to_year": float(to_year)}})


def average_at(obs: Observable, dates: Sequence[float]) -> Observable:
    """Arithmetic mean of ``obs`` over fixing dates (extension, no reference
    counterpart): at evaluation time t, the mean over the fixings on or before
    t.  The fixings join the pricing grid as fixed dates; a monthly-fixing
    Asian on a daily grid is ``average_at(S, [i / 12 for i in 1..12])``."""
    return Observable({"AverageAt": {"obs": _obs(obs).node, "dates": [float(d) for d in dates]}})


def running_max(obs: Observable) -> Observable:
    return Observable({"RunningMax": _obs(obs).node})


def running_min(obs: Observable) -> Observable:
    return Observable({"RunningMin": _obs(obs).node})


def scale(obs, c: Contract) -> Contract:
    return Contract({"Scale": [_obs(obs).node, c.node]})


def when(cond: Condition, c: Contract) -> Contract:
    return Contract({"When": [cond.node, c.node]})


def anytime(cond: Condition, c: Contract) -> Contract:
    return Contract({"Anytime": [cond.node, c.node]})


def until(cond: Condition, c: Contract) -> Contract:
    return Contract({"Until": [cond.node, c.node]})


def and_(a: Contract, b: Contract) -> Contract:
    return a + b


def or_(a: Contract, b: Contract) -> Contract:
    return a.or_(b)


# Product catalogue (src/products.rs:16-84)
def zero_coupon_bond(ccy: str, t: float) -> Contract:
    return one(ccy) @ at(t)


def european_call(asset: str, k: float, t: float, ccy: str = USD) -> Contract:
    return ((spot(asset) - const_(k)).max(const_(0.0)) * one(ccy)) @ at(t)


def european_put(asset: str, k: float, t: float, ccy: str = USD) -> Contract:
    return ((const_(k) - spot(asset)).max(const_(0.0)) * one(ccy)) @ at(t)


def forward(asset: str, k: float, t: float, ccy: str = USD) -> Contract:
    return ((spot(asset) - const_(k)) * one(ccy)) @ at(t)


def straddle(asset: str, k: float, t: float, ccy: str = USD) -> Contract:
    return european_call(asset, k, t, ccy) + european_put(asset, k, t, ccy)


def bull_call_spread(asset: str, k_low: float, k_high: float, t: float, ccy: str = USD) -> Contract:
    return european_call(asset, k_low, t, ccy) + (-european_call(asset, k_high, t, ccy))


def cash_or_nothing_call(asset: str, k: float, payout: float, t: float, ccy: str = USD) -> Contract:
    return when(at(t), when(spot(asset) >= const_(k), scale(const_(payout), one(ccy))))


def up_and_out_call(asset: str, k: float, barrier: float, t: float, ccy: str = USD) -> Contract:
    return until(spot(asset) >= const_(barrier), european_call(asset, k, t, ccy))


def down_and_out_call(asset: str, k: float, barrier: float, t: float, ccy: str = USD) -> Contract:
    return until(spot(asset) <= const_(barrier), european_call(asset, k, t, ccy))


# ---------------------------------------------------------------------------
# Models (src/bindings.rs:478-567, 793-836)
# ---------------------------------------------------------------------------
class GBM:
    """``GBM(s0, sigma, r, q=0.0, asset="ASSET")``: drift r - q, discount r."""

    def __init__(self, s0: float, sigma: float, r: float, q: float = 0.0, asset: str = "ASSET"):
        self.s0, self.sigma, self.r, self.q, self.asset = float(s0), float(sigma), float(r), float(q), asset

    @property
    def rate(self) -> float:
        return self.r

    def __repr__(self):
        return f"GBM(asset={self.asset!r}, s0={self.s0}, sigma={self.sigma}, r={self.r}, q={self.q})"


class Model:
    """A simulator built by a factory (``heston``, ``correlated_gbm``)."""

    def __init__(self, kind: int, rate: float, label: str, **params):
        self.kind, self.rate, self.label = kind, float(rate), label
        self.params = params

    def __repr__(self):
        return f"Model({self.label}, rate={self.rate})"


class GbmFactor:
    """``GbmFactor(name, s0, mu, sigma)`` (src/fx.rs:149-170)."""

    def __init__(self, name: str, s0: float, mu: float, sigma: float):
        self.name, self.s0, self.mu, self.sigma = name, float(s0), float(mu), float(sigma)

    def __repr__(self):
        return f"GbmFactor(name={self.name!r}, s0={self.s0}, mu={self.mu}, sigma={self.sigma})"


def heston(spot: float, v0: float, kappa: float, theta: float, sigma_v: float, rho: float, r: float,
           asset: str = "ASSET", scheme: str = "floor") -> Model:
    """Heston model (src/bindings.rs:547-567).  ``scheme="floor"`` is the
    reference's variance floor after each step; ``"full_truncation"`` is the
    alternative scheme."""
    if scheme not in ("floor", "full_truncation"):
        raise ValueError("scheme must be 'floor' or 'full_truncation'")
    return Model(N.KT_MODEL_HESTON, r, "Heston", asset=asset, spot=float(spot), v0=float(v0), kappa=float(kappa),
                 theta=float(theta), sigma_v=float(sigma_v), rho=float(rho), r=float(r),
                 scheme=0 if scheme == "floor" else 1)


def correlated_gbm(factors: Sequence[GbmFactor], corr: Sequence[Sequence[float]], r: float) -> Model:
    """N correlated GBMs (src/bindings.rs:826-836, src/multi_asset.rs:62-83)."""
    factors = list(factors)
    k = len(factors)
    if k < 2:
        raise N.KontractError(6, "CorrelatedGbmN: at least 2 assets required")
    if len(corr) != k or any(len(row) != k for row in corr):
        raise N.KontractError(6, f"CorrelatedGbmN: correlation matrix must be {k}x{k}")
    if any(f.sigma < 0 for f in factors):
        raise N.KontractError(6, "CorrelatedGbmN: all volatilities must be >= 0")
    return Model(N.KT_MODEL_CORRELATED_GBM, r, "CorrelatedGbmN", factors=factors,
                 corr=[[float(x) for x in row] for row in corr])


class _Struct:
    """Keeps a ctypes struct and the buffers it points to alive together."""

    def __init__(self, s, keep):
        self.s = s
        self.keep = keep


def _shared_f64(xs) -> bool:
    """True when _dbl shares xs's buffer instead of copying it."""
    return isinstance(xs, np.ndarray) and xs.dtype == np.float64 and xs.ndim == 1 and bool(xs.size) \
        and xs.flags.c_contiguous and xs.flags.writeable


def _dbl(xs) -> C.Array:
    """A C double array of xs: a float64 numpy array is shared without a
    copy; anything else is copied (array.array converts a list in one C
    loop).  The ctypes array keeps the buffer alive."""
    if _shared_f64(xs):
        return (C.c_double * xs.size).from_buffer(xs)
    if isinstance(xs, (list, tuple)) and len(xs) <= 32:
        return (C.c_double * max(len(xs), 1))(*xs)
    a = array.array("d", xs)
    if not a:
        a.append(0.0)
    return (C.c_double * len(a)).from_buffer(a)


def _model_key(model):
    """The values _model_struct reads from a model (its cache key)."""
    if isinstance(model, GBM):
        return (model.s0, model.sigma, model.r, model.q, model.asset)
    p = model.params
    if model.kind == N.KT_MODEL_CORRELATED_GBM:
        return (tuple((f.name, f.s0, f.mu, f.sigma) for f in p["factors"]),
                tuple(tuple(row) for row in p["corr"]))
    return tuple(p.items())


def _model_struct(model) -> _Struct:
    """The model's kt_model struct, kept on the model object and rebuilt
    when any value it was built from has changed (a book priced again
    marshals each model once)."""
    if not isinstance(model, (GBM, Model)):
        raise TypeError("expected a GBM or Model")
    key = _model_key(model)
    cached = getattr(model, "_kt_struct", None)
    if cached is not None and cached[0] == key:
        return cached[1]
    st = _build_model_struct(model)
    model._kt_struct = (key, st)
    return st


def _build_model_struct(model) -> _Struct:
    if isinstance(model, GBM):
        names = (C.c_char_p * 1)(model.asset.encode())
        s0, mu, sg = _dbl([model.s0]), _dbl([model.r - model.q]), _dbl([model.sigma])
        s = N.KtModel(N.KT_MODEL_GBM, 1, names, s0, mu, sg, None, 0, 0, 0, 0, 0, 0, 0)
        return _Struct(s, (names, s0, mu, sg))
    if isinstance(model, Model):
        p = model.params
        if model.kind == N.KT_MODEL_HESTON:
            names = (C.c_char_p * 1)(p["asset"].encode())
            s0, mu, sg = _dbl([p["spot"]]), _dbl([p["r"]]), _dbl([0.0])
            s = N.KtModel(N.KT_MODEL_HESTON, 1, names, s0, mu, sg, None, p["v0"], p["kappa"], p["theta"],
                          p["sigma_v"], p["rho"], p["scheme"], 0)
            return _Struct(s, (names, s0, mu, sg))
        if model.kind == N.KT_MODEL_CORRELATED_GBM:
            fs = p["factors"]
            k = len(fs)
            names = (C.c_char_p * k)(*[f.name.encode() for f in fs])
            s0, mu, sg = _dbl([f.s0 for f in fs]), _dbl([f.mu for f in fs]), _dbl([f.sigma for f in fs])
            corr = _dbl([x for row in p["corr"] for x in row])
            s = N.KtModel(N.KT_MODEL_CORRELATED_GBM, k, names, s0, mu, sg, corr, 0, 0, 0, 0, 0, 0, 0)
            return _Struct(s, (names, s0, mu, sg, corr))
    raise TypeError("expected a GBM or Model")


def _cfg(n_paths, seed, steps_per_year, rate, grid, antithetic=False, control_variate=False) -> _Struct:
    if n_paths < 0 or seed < 0 or steps_per_year < 0:
        raise OverflowError("can't convert negative int to unsigned")
    g = _dbl(grid) if grid is not None else None
    s = N.KtMcConfig(int(n_paths), int(seed) & 0xFFFFFFFFFFFFFFFF, int(steps_per_year), float(rate),
                     g, len(grid) if grid is not None else 0, 1 if antithetic else 0, 1 if control_variate else 0)
    return _Struct(s, (g,))


# ---------------------------------------------------------------------------
# Batch, portfolio, host-only plan
# ---------------------------------------------------------------------------
def price_batch(contracts: Iterable[Contract], model, n_paths: int = 100_000, seed: int = 42,
                steps_per_year: int = 50, device: int | None = None) -> list[PriceResult]:
    """price_batch_gbm (src/pricer.rs:250-294): one shared simulation."""
    contracts = list(contracts)
    if not contracts:
        return []
    m = _model_struct(model)
    cfg = _cfg(n_paths, seed, steps_per_year, model.rate, None)
    arr = (C.c_char_p * len(contracts))(*[c.to_json().encode() for c in contracts])
    out = (N.KtPriceResult * len(contracts))()
    eng = N.engine(device)
    N.check(N.lib().kt_price_batch(eng.handle, arr, len(contracts), C.byref(m.s), C.byref(cfg.s), out), eng.handle)
    return [PriceResult(r) for r in out]


class _JobsMemo:
    """The last book's kt_job array.  A book priced again with the same job
    tuples (tuples are immutable, and so are contracts) reuses it when every
    model's struct is unchanged (models are mutable: _model_struct rebuilds
    the struct of a changed model, and the array is then built afresh).  Only
    books whose grids are absent, tuples, or float64 arrays the engine reads
    in place are kept (a copied list could change behind the copy).  The
    array is only read by the engine."""

    def __init__(self):
        self.jobs = None
        self.arr = None
        self.models = None
        self.keep = None


_jobs_memo = _JobsMemo()
_jobs_lock = threading.Lock()


def _jobs_array(jobs):
    with _jobs_lock:
        memo = _jobs_memo
        if memo.jobs is not None and len(memo.jobs) == len(jobs) and all(a is b for a, b in zip(jobs, memo.jobs)) \
                and all(_model_struct(job[1]) is m for job, m in zip(jobs, memo.models)):
            return memo.arr, memo.keep
    keep = []
    models = []
    arr = (N.KtJob * len(jobs))()
    for i, job in enumerate(jobs):
        con, model, n_paths, seed, spy = job[:5]
        grid = job[5] if len(job) > 5 else None
        anti = bool(job[6]) if len(job) > 6 else False
        cv = bool(job[7]) if len(job) > 7 else False
        m = _model_struct(model)
        cfg = _cfg(n_paths, seed, spy, model.rate, grid, anti, cv)
        js = con.to_json().encode()
        keep += [m, cfg, js]
        models.append(m)
        arr[i].contract_json = js
        arr[i].model = m.s
        arr[i].cfg = cfg.s
    if all(len(job) <= 5 or job[5] is None or isinstance(job[5], tuple) or _shared_f64(job[5]) for job in jobs):
        with _jobs_lock:
            memo.jobs, memo.arr, memo.models, memo.keep = list(jobs), arr, models, keep
    return arr, keep


def price_portfolio(jobs: Sequence[tuple], device: int | None = None, greeks: bool = False):
    """Heterogeneous portfolio: jobs are (contract, model, n_paths, seed,
    steps_per_year[, grid[, antithetic[, control_variate]]]) tuples, priced in
    one persistent launch; each result equals ``contract.price(model, n_paths,
    seed, steps_per_year, antithetic, control_variate, grid=grid)``
    (variance-reduced jobs run their own launches).

    With ``greeks=True`` returns (PriceResult, ModelGreeks | None) pairs: the
    jobs whose payoff admits pathwise Greeks (``Contract.pathwise_greeks``)
    get their derivatives from persistent Greeks launches, equal to
    ``pathwise_greeks`` on the job alone; the others (barriers, digitals,
    variance-reduced jobs) are priced without Greeks (None)."""
    jobs = list(jobs)
    if not jobs:
        return []
    arr, keep = _jobs_array(jobs)
    out = (N.KtPriceResult * len(jobs))()
    eng = N.engine(device)
    if not greeks:
        N.check(N.lib().kt_price_portfolio(eng.handle, arr, len(jobs), out), eng.handle)
        return [PriceResult._from_fields(t) for t in np.ctypeslib.as_array(out).tolist()]
    stride = 24
    g = (C.c_double * (stride * len(jobs)))()
    se = (C.c_double * (stride * len(jobs)))()
    N.check(N.lib().kt_price_portfolio_greeks(eng.handle, arr, len(jobs), out, g, se, stride), eng.handle)
    # the outputs as Python lists in one conversion each
    prices = np.ctypeslib.as_array(out).tolist()
    grows = np.ctypeslib.as_array(g).reshape(len(jobs), stride).tolist()
    srows = np.ctypeslib.as_array(se).reshape(len(jobs), stride).tolist()
    res = []
    for j, job in enumerate(jobs):
        pr = PriceResult._from_fields(prices[j])
        names = _param_names(job[1])
        n = len(names)
        row = grows[j][:n]
        if any(x != x for x in row):  # NaN: no pathwise Greeks for this job
            res.append((pr, None))
            continue
        res.append((pr, ModelGreeks._from_rows(names, pr.price, pr.std_error, row, srows[j][:n], "pathwise")))
    return res


class Plan:
    """compile() result (src/compiler.rs:25-34) plus its grid."""

    def __init__(self, assets, fixed_dates, horizon, needs_fine_grid):
        self.assets, self.fixed_dates, self.horizon, self.needs_fine_grid = assets, fixed_dates, horizon, \
            needs_fine_grid

    def time_grid(self, steps_per_year: int) -> list[float]:
        return self.time_grid_forced(steps_per_year, False)

    def time_grid_forced(self, steps_per_year: int, force_fine: bool) -> list[float]:
        return _plan_raw(self._json, steps_per_year, force_fine)[0]


def _plan_raw(js: str, spy: int, force: bool):
    L = N.lib()
    cap = 1 << 16
    grid = (C.c_double * cap)()
    dates = (C.c_double * cap)()
    ng, nd = C.c_uint64(), C.c_uint64()
    fine, hz = C.c_int32(), C.c_double()
    assets = C.create_string_buffer(1 << 16)
    N.check(L.kt_plan(js.encode(), int(spy), 1 if force else 0, grid, C.byref(ng), dates, C.byref(nd), cap,
                      C.byref(fine), C.byref(hz), assets, len(assets)))
    if ng.value > cap or nd.value > cap:
        raise N.KontractError(7, "grid too large for this helper")
    names = assets.value.decode()
    return (list(grid[: ng.value]), list(dates[: nd.value]), hz.value, bool(fine.value),
            names.split("\n") if names else [])


def plan(contract: Contract) -> Plan:
    """compile (src/compiler.rs:69-85), host only."""
    js = contract.to_json()
    _, dates, hz, fine, assets = _plan_raw(js, 0, False)
    p = Plan(assets, dates, hz, fine)
    p._json = js
    return p


compile = plan  # the reference's name


def describe_program(contract: Contract, model, n_paths: int = 1, seed: int = 42, steps_per_year: int = 50,
                     grid: Sequence[float] | None = None) -> str:
    """Text listing of the op program the GPU engine runs for a contract."""
    m = _model_struct(model)
    cfg = _cfg(n_paths, seed, steps_per_year, model.rate, grid)
    buf = C.create_string_buffer(1 << 20)
    N.check(N.lib().kt_describe_program(contract.to_json().encode(), C.byref(m.s), C.byref(cfg.s), buf, len(buf)))
    return buf.value.decode()
