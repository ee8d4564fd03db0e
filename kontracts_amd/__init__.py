"""kontracts_amd — MI355X Monte Carlo pricing engine for kontract contracts.

Drop-in for the path-simulation hot path of the reference's Python package
(python/kontract/__init__.py): the same contract algebra and the same
``Contract.price(model, n_paths, seed, steps_per_year)``, ``Contract.greeks``
and ``Contract.price_american`` calls, executed by hand-written gfx950 kernels
behind the C ABI in include/kontracts.h.  Beyond the reference:
``Contract.pathwise_greeks`` / ``Contract.bump_greeks`` (derivatives with
respect to every model parameter) and ``kontracts_amd.autograd`` (a
torch.autograd.Function over the pathwise Greeks).  ``greek_surface``
follows the reference's Rust ``surface::greek_surface``.
"""
from ._native import DeviceError, KontractError
from .surface import GreekSurface, greek_surface
from .dsl import (
    EUR,
    GBM,
    GBP,
    JPY,
    USD,
    Condition,
    Contract,
    GbmFactor,
    Greeks,
    Model,
    ModelGreeks,
    Observable,
    Plan,
    PriceResult,
    S,
    __version__,
    and_,
    anytime,
    at,
    average,
    average_at,
    average_over,
    bool_,
    bull_call_spread,
    cash_or_nothing_call,
    compile,
    const_,
    correlated_gbm,
    describe_program,
    down_and_out_call,
    european_call,
    european_put,
    forward,
    give,
    heston,
    konst,
    one,
    or_,
    plan,
    price_batch,
    price_portfolio,
    running_max,
    running_min,
    scale,
    spot,
    straddle,
    until,
    up_and_out_call,
    when,
    zero,
    zero_coupon_bond,
)

__all__ = [
    "DeviceError", "KontractError", "EUR", "GBM", "GBP", "JPY", "USD", "Condition", "Contract", "GbmFactor",
    "Greeks", "Model", "ModelGreeks", "Observable", "Plan", "PriceResult", "S", "__version__", "and_", "anytime", "at",
    "average", "average_at", "average_over", "bool_", "bull_call_spread", "cash_or_nothing_call", "compile", "const_",
    "correlated_gbm", "describe_program", "down_and_out_call", "european_call", "european_put", "forward", "give",
    "heston", "konst", "one", "or_", "plan", "price_batch", "price_portfolio", "running_max", "running_min",
    "scale", "spot", "straddle", "until", "up_and_out_call", "when", "zero", "zero_coupon_bond",
    "GreekSurface", "greek_surface",
]
