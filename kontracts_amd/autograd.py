"""torch.autograd bridge for the engine's pathwise Greeks.

``MonteCarloPrice.apply(params, contract, build_model, n_paths, seed,
steps_per_year)`` prices ``contract`` under ``build_model(*params)`` and, on
backward, returns the upstream gradient times the pathwise derivatives of the
price with respect to ``params`` (one ``kt_greeks_adjoint`` launch computes
both).  The derivative is the Monte Carlo estimate at the same seed as the
price, so the pair is consistent: a finite difference of prices at one seed
converges to it as the bump shrinks.

``gbm_price`` and ``heston_price`` bind the parameter order of the GBM
(s0, sigma, mu) and Heston (s0, v0, kappa, theta, sigma_v, rho) models.
"""
from __future__ import annotations

import torch

from . import dsl


class MonteCarloPrice(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, contract, build_model, n_paths, seed, steps_per_year):
        values = [float(v) for v in params.detach().cpu().reshape(-1).tolist()]
        g = contract.pathwise_greeks(build_model(*values), n_paths=n_paths, seed=seed,
                                     steps_per_year=steps_per_year)
        if len(g.values) != len(values):
            raise ValueError(f"model has {len(g.values)} parameters, got {len(values)}")
        ctx.save_for_backward(torch.tensor(g.values, dtype=params.dtype, device=params.device).reshape(params.shape))
        return torch.tensor(g.price, dtype=params.dtype, device=params.device)

    @staticmethod
    def backward(ctx, grad_out):
        (d,) = ctx.saved_tensors
        return grad_out * d, None, None, None, None, None


def gbm_price(contract, params, r: float, asset: str = "ASSET", n_paths: int = 200_000, seed: int = 42,
              steps_per_year: int = 50):
    """Differentiable price under GBM with params = (s0, sigma, mu) and
    discount rate r (the model's drift is mu)."""
    def build(s0, sigma, mu):
        return dsl.GBM(s0, sigma, r, q=r - mu, asset=asset)
    return MonteCarloPrice.apply(params, contract, build, n_paths, seed, steps_per_year)


def heston_price(contract, params, r: float, asset: str = "ASSET", scheme: str = "floor", n_paths: int = 200_000,
                 seed: int = 42, steps_per_year: int = 50):
    """Differentiable price under Heston with params = (s0, v0, kappa, theta,
    sigma_v, rho) and rate r."""
    def build(s0, v0, kappa, theta, sigma_v, rho):
        return dsl.heston(s0, v0, kappa, theta, sigma_v, rho, r, asset=asset, scheme=scheme)
    return MonteCarloPrice.apply(params, contract, build, n_paths, seed, steps_per_year)
