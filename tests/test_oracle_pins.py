"""CPU pins of the oracle functions that the reference's closed forms and its
own statistical tests fix: the correlated GBM below perfect correlation
(tests/j27_basket.rs), the Heston CRN bump Greeks (against central
differences of the Heston characteristic-function price, itself pinned to
Fang-Oosterlee in test_oracle.py) and the reference's discrete-barrier
pattern (tests/barriers.rs:60-84).

Each check is chosen so that a plausible mistake fails it: drawing
w = C z instead of w = L z changes the correlation (0.7 -> 0.94) and the
spread volatility; a swapped, mis-signed or mis-scaled bump moves a Greek by
far more than the tolerance; a dropped BGK shift moves the barrier price by
several percent.
"""
import math

import numpy as np
import pytest

import kontracts_amd as K
import oracle_lib as O


def bs(s, k, t, r, sg):
    return O.bs_call(s, k, t, r, sg)


def spread_margrabe(s1, s2, t, sg1, sg2, rho):
    """Margrabe (1978) exchange option max(S1 - S2, 0): exact under GBM with
    both assets drifting at the discount rate (the rate cancels)."""
    sg = math.sqrt(sg1 * sg1 + sg2 * sg2 - 2.0 * rho * sg1 * sg2)
    return bs(s1, s2, t, 0.0, sg)


# ---------------------------------------------------------------------------
# CorrelatedGbmN below perfect correlation (src/multi_asset.rs:62-179)
# ---------------------------------------------------------------------------
def test_correlated_gbm_recovers_rho():
    # tests/j27_basket.rs:27-66: log-returns over [0, 1] of two assets with
    # rho = 0.7 have empirical correlation within 0.02 of it (100,000 paths)
    m = O.basket(["S1", "S2"], [100.0, 100.0], [0.0, 0.0], [0.25, 0.20], [[1.0, 0.7], [0.7, 1.0]])
    x = O.simulate(m, [0.0, 1.0], 100_000, 42)
    la, lb = np.log(x[:, 0, 1] / 100.0), np.log(x[:, 1, 1] / 100.0)
    corr = np.corrcoef(la, lb)[0, 1]
    assert abs(corr - 0.7) < 0.02, corr
    # the marginal log-return moments of the exact log step
    for i, (sg, l) in enumerate(((0.25, la), (0.20, lb))):
        assert abs(l.mean() - (-0.5 * sg * sg)) < 4 * sg / math.sqrt(len(l))
        assert abs(l.std() / sg - 1.0) < 0.01


def test_correlated_gbm_per_step_innovations_config4_model():
    # The portfolio's 5-asset basket model (pairwise rho = 0.3): standardised
    # per-step log-returns w_i = (dlog S_i - (mu_i - sigma_i^2 / 2) dt) /
    # (sigma_i sqrt(dt)) have unit variance, correlation matrix rho, and no
    # correlation across steps (each grid point draws fresh normals).
    k, n_steps, n_paths = 5, 8, 40_000
    sig = [0.15, 0.25, 0.35, 0.45, 0.30]
    corr = [[1.0 if i == j else 0.3 for j in range(k)] for i in range(k)]
    m = O.basket([f"B{i}" for i in range(k)], [100.0] * k, [0.05] * k, sig, corr)
    dt = 1.0 / 252
    grid = [i * dt for i in range(n_steps + 1)]
    x = O.simulate(m, grid, n_paths, 7)
    lr = np.diff(np.log(x), axis=2)  # [path][asset][step]
    w = (lr - ((0.05 - 0.5 * np.array(sig) ** 2) * dt)[None, :, None]) / (np.array(sig) * math.sqrt(dt))[None, :, None]
    flat = w.transpose(1, 0, 2).reshape(k, -1)
    c = np.corrcoef(flat)
    se = (1 - 0.3 ** 2) / math.sqrt(flat.shape[1])
    off = c[~np.eye(k, dtype=bool)]
    assert np.all(np.abs(off - 0.3) < 5 * se), off
    assert np.all(np.abs(flat.var(axis=1) - 1.0) < 0.02)
    lag = np.corrcoef(w[:, 0, :-1].reshape(-1), w[:, 0, 1:].reshape(-1))[0, 1]
    assert abs(lag) < 5 / math.sqrt(n_paths * (n_steps - 1))
    # the factor w = L z itself: L L^T == rho
    L = np.linalg.cholesky(np.array(corr))
    np.testing.assert_allclose(L @ L.T, corr, atol=1e-12)


def test_basket_three_assets_vs_black_scholes_sigma_basket():
    # tests/j27_basket.rs:73-107: equal weights, equal vols, uniform rho = 0.5:
    # the lognormal approximation BS(sigma sqrt((1 + 2 rho) / 3)) within 3 %
    s0, k, t, r, sg, rho = 100.0, 100.0, 1.0, 0.03, 0.20, 0.5
    m = O.basket(["S1", "S2", "S3"], [s0] * 3, [r] * 3, [sg] * 3,
                 [[1.0, rho, rho], [rho, 1.0, rho], [rho, rho, 1.0]])
    c = K.scale(((K.spot("S1") + K.spot("S2") + K.spot("S3")) / 3.0 - k).clip(0.0), K.one("USD")) @ K.at(t)
    res, _, _ = O.price(c.to_json(), m, O.cfg(200_000, 7, 50, r))
    ana = bs(s0, k, t, r, sg * math.sqrt((1.0 + 2.0 * rho) / 3.0))
    assert abs(res["price"] - ana) / ana < 0.03
    # and clearly not the single-asset value (the basket diversifies)
    assert res["price"] < bs(s0, k, t, r, sg) - 10 * res["std_error"]


def test_spread_option_vs_margrabe():
    # tests/j27_basket.rs:112-140 (3 %), and since Margrabe is exact here,
    # within 4 standard errors too
    s0, t, r, sg1, sg2, rho = 100.0, 1.0, 0.03, 0.20, 0.20, 0.5
    m = O.basket(["S1", "S2"], [s0, s0], [r, r], [sg1, sg2], [[1.0, rho], [rho, 1.0]])
    c = K.scale((K.spot("S1") - K.spot("S2")).clip(0.0), K.one("USD")) @ K.at(t)
    res, _, _ = O.price(c.to_json(), m, O.cfg(300_000, 11, 50, r))
    ana = spread_margrabe(s0, s0, t, sg1, sg2, rho)
    assert abs(res["price"] - ana) / ana < 0.03
    assert abs(res["price"] - ana) < 4 * res["std_error"]
    # unequal vols and a negative correlation
    m2 = O.basket(["S1", "S2"], [110.0, 100.0], [r, r], [0.3, 0.15], [[1.0, -0.4], [-0.4, 1.0]])
    res2, _, _ = O.price(c.to_json(), m2, O.cfg(200_000, 5, 12, r))
    ana2 = spread_margrabe(110.0, 100.0, t, 0.3, 0.15, -0.4)
    assert abs(res2["price"] - ana2) < 4 * res2["std_error"]


# ---------------------------------------------------------------------------
# Heston bump Greeks against the characteristic-function price
# ---------------------------------------------------------------------------
HESTON_P = dict(s0=100.0, k=100.0, t=1.0, r=0.05, v0=0.04, kappa=2.0, theta=0.04, sv=0.3, rho=-0.5)
HESTON_KEYS = ("s0", "v0", "kappa", "theta", "sv", "rho")


def heston_cf(p):
    return O.heston_call(p["s0"], p["k"], p["t"], p["r"], p["v0"], p["kappa"], p["theta"], p["sv"], p["rho"])


def heston_cf_greeks(p=HESTON_P):
    """Central differences of the characteristic-function price, step 1e-4
    relative: the quadrature is a smooth function of the parameters, so the
    differences carry no Monte Carlo noise and an O(h^2) error far below any
    test tolerance."""
    out = []
    for key in HESTON_KEYS:
        h = 1e-4 * max(abs(p[key]), 1e-2)
        up, dn = dict(p), dict(p)
        up[key] += h
        dn[key] -= h
        out.append((heston_cf(up) - heston_cf(dn)) / (2.0 * h))
    return out


def test_heston_cf_greeks_signs_and_delta_bound():
    g = heston_cf_greeks()
    d_s0, d_v0, d_kappa, d_theta, d_sv, d_rho = g
    assert 0.5 < d_s0 < 1.0  # an ATM call's delta
    assert d_v0 > 0 and d_theta > 0  # more variance, dearer call
    # theta = v0 keeps the expected variance flat; faster reversion narrows the
    # spread of the integrated variance, which lifts an ATM call under skew
    assert d_kappa > 0
    assert d_sv < 0 and d_rho < 0  # ATM call under negative skew: vol of vol and rho cheapen it
    # sigma_v -> 0 limit: the v0 derivative tends to the BS vega w.r.t. the
    # average variance, vega / (2 sigma) weighted by (1 - e^{-kappa T}) / (kappa T)
    p = dict(HESTON_P, sv=1e-4, rho=0.0)
    g0 = heston_cf_greeks(p)
    vega = O.bs_call_greeks(100, 100, 1.0, 0.05, 0.2)["vega"]
    w = (1 - math.exp(-2.0)) / 2.0
    assert g0[1] == pytest.approx(vega / (2 * 0.2) * w, rel=2e-3)
    assert g0[0] == pytest.approx(O.bs_call_greeks(100, 100, 1.0, 0.05, 0.2)["delta"], rel=1e-4)


def test_heston_bump_greeks_vs_characteristic_function():
    # 8 independent batches of 8,192 paths x 50 steps a year; the batch means'
    # spread gives each Greek's standard error.  Tolerance 4 SE plus 3 % of
    # the value for the Euler bias of 50 steps (the price itself sits within
    # 1.2 SE of the CF at this size).
    c = K.european_call("X", 100.0, 1.0).to_json()
    h = [1.0, 1e-3, 1e-2, 1e-3, 1e-2, 1e-2]
    m = O.heston("X", 100.0, 0.05, 0.04, 2.0, 0.04, 0.3, -0.5)
    vals = np.array([O.greeks_heston_bump(c, m, O.cfg(1 << 13, seed, 50, 0.05), h) for seed in range(1, 9)])
    mean = vals.mean(axis=0)
    se = vals.std(axis=0, ddof=1) / math.sqrt(len(vals))
    assert abs(mean[0] - heston_cf(HESTON_P)) < 4 * se[0] + 0.01 * mean[0]
    for name, got, s, want in zip(HESTON_KEYS, mean[1:], se[1:], heston_cf_greeks()):
        assert abs(got - want) < 4 * s + 0.03 * abs(want), (name, got, want, s)


# ---------------------------------------------------------------------------
# Discrete-barrier pattern of the reference (tests/barriers.rs:60-84)
# ---------------------------------------------------------------------------
def down_and_out_bgk(s, k, h, r, sg, t, steps):
    """Down-and-out call (h <= k) = BS - down-and-in (Reiner-Rubinstein),
    with the Broadie-Glasserman-Kou shift h e^{-0.5826 sigma sqrt(dt)} for
    monitoring at the grid points."""
    h = h * math.exp(-0.5826 * sg * math.sqrt(t / steps))
    lam = (r + 0.5 * sg * sg) / (sg * sg)
    st = sg * math.sqrt(t)
    y = math.log(h * h / (s * k)) / st + lam * st
    nc = lambda x: 0.5 * math.erfc(-x / math.sqrt(2.0))
    di = s * (h / s) ** (2 * lam) * nc(y) - k * math.exp(-r * t) * (h / s) ** (2 * lam - 2) * nc(y - st)
    return bs(s, k, t, r, sg) - di


def test_down_and_out_call_reference_pattern():
    s0, k, h, r, sg, t, steps = 100.0, 100.0, 90.0, 0.05, 0.20, 1.0, 250
    ko = K.down_and_out_call("AAPL", k, h, t)
    res, _, _ = O.price(ko.to_json(), O.gbm("AAPL", s0, r, sg), O.cfg(300_000, 2024, steps, r))
    ana = down_and_out_bgk(s0, k, h, r, sg, t, steps)
    assert abs(res["price"] - ana) / ana < 0.02
    # without the discrete-monitoring shift the formula is off by more
    ana_cont = down_and_out_bgk(s0, k, h, r, sg, t, 10 ** 12)
    assert abs(res["price"] - ana_cont) > abs(res["price"] - ana)


def test_config2_up_and_out_reference_pattern():
    # the bench contract (s0 180, K 100, H 200, sigma 0.25, 252 daily points):
    # Reiner-Rubinstein with H e^{+0.5826 sigma sqrt(dt)} within 2 %
    s0, k, h, r, sg = 180.0, 100.0, 200.0, 0.05, 0.25
    c = K.up_and_out_call("X", k, h, 1.0)
    res, _, _ = O.price(c.to_json(), O.gbm("X", s0, r, sg), O.cfg(200_000, 42, 252, r))
    ana = O.uo_call(s0, k, h * math.exp(0.5826 * sg * math.sqrt(1 / 252)), 1.0, r, sg)
    assert abs(res["price"] - ana) / ana < 0.02
    assert abs(res["price"] - ana) < 4 * res["std_error"] + 0.005 * ana
