"""CPU tests that pin the oracle's variance reduction (src/pricer.rs:117-228,
src/variance_reduction.rs) and the Black-Scholes value the control variate
corrects with (src/numerics.rs:22-71).

* The Abramowitz-Stegun normal CDF stays inside its published error bound of
  the exact CDF, and is not the exact CDF; the engine's host copy
  (kt_black_scholes_call) equals the oracle's.
* Antithetic pairs are paths 0..n/2-1 of the stream and their negations: the
  oracle's pair statistics equal those built from the stream's own normals
  fed in with both signs.
* The control variate's price is mean(target) - mean(control) + BS on the
  same paths, its spread that of the per-path difference; the (s0, sigma)
  defaults for non-GBM models and the first-asset control follow the
  reference; antithetic pairs fall back to plain paths off GBM.
* The reference's own acceptance criteria (tests/j15_variance_reduction.rs
  VR-1..VR-4) hold on the oracle.
"""
import math

import numpy as np
import pytest
from scipy.special import ndtr

import oracle_lib as O
import kontracts_amd as K
from kontracts_amd import _native as N

GBM_S = O.gbm("S", 100.0, 0.05, 0.2)


def uniform_grid(t, n):
    return [i * t / n for i in range(n + 1)]


# ---------------------------------------------------------------------------
# Black-Scholes with the reference's normal CDF
# ---------------------------------------------------------------------------
def test_as_normal_cdf_within_published_bound():
    xs = np.linspace(-8.0, 8.0, 4001)
    got = np.array([O.norm_cdf_as(x) for x in xs])
    err = np.abs(got - ndtr(xs))
    # A&S 7.1.26: |erf error| <= 1.5e-7, so |N error| <= 7.5e-8
    assert err.max() <= 7.5e-8
    assert err.max() > 1e-9  # the approximation, not the exact CDF
    assert abs(O.norm_cdf_as(0.0) - 0.5) < 1e-9
    # symmetry N(-x) = 1 - N(x) holds exactly by construction
    for x in (0.3, 1.7, 4.2):
        assert O.norm_cdf_as(-x) == pytest.approx(1.0 - O.norm_cdf_as(x), abs=1e-15)


def test_as_black_scholes_and_engine_copy():
    L = N.lib()
    for s, k, t, r, sg in [(100, 100, 1, 0.05, 0.2), (80, 100, 0.5, 0.01, 0.4), (150, 100, 3, 0.07, 0.15),
                           (100, 100, 1e-4, 0.05, 0.2), (42.5, 41.0, 2.2, 0.0, 0.33)]:
        a = O.bs_call_as(s, k, t, r, sg)
        assert abs(a - O.bs_call(s, k, t, r, sg)) <= 7.5e-8 * (s + k) + 1e-12
        assert L.kt_black_scholes_call(s, k, t, r, sg) == pytest.approx(a, rel=1e-14, abs=1e-14)
    # degenerate branch: max(s - k e^{-rt}, 0)
    assert O.bs_call_as(100, 90, 0.0, 0.05, 0.2) == pytest.approx(10.0)
    assert O.bs_call_as(100, 110, 1.0, 0.05, 0.0) == pytest.approx(max(100 - 110 * math.exp(-0.05), 0.0))
    assert L.kt_black_scholes_call(100, 110, 1.0, 0.05, 0.0) == O.bs_call_as(100, 110, 1.0, 0.05, 0.0)


# ---------------------------------------------------------------------------
# Antithetic pairs
# ---------------------------------------------------------------------------
def fed_normals(seed, n_paths, n_grid, m=1):
    z = np.zeros((n_paths, n_grid, m), dtype=np.float32)
    for p in range(n_paths):
        z[p] = O.normals(seed, p, 0, n_grid * m).reshape(n_grid, m)
    return z


@pytest.mark.parametrize("kind", ["european", "asian", "barrier"])
def test_antithetic_pairs_are_the_negated_stream(kind):
    t = 1.0
    c = {
        "european": K.european_call("S", 100.0, t),
        "asian": K.scale((K.average(K.spot("S")) - 100.0).clip(0.0), K.one("USD")) @ K.at(t),
        "barrier": K.up_and_out_call("S", 100.0, 125.0, t),
    }[kind].to_json()
    grid = uniform_grid(t, 24)
    n = 301  # odd: 150 pairs
    cfg = O.cfg(n, 9, 50, 0.05, grid=grid)
    got = O.price_vr(c, GBM_S, cfg, antithetic=True)
    z = fed_normals(9, n // 2, len(grid))
    cfg_h = O.cfg(n // 2, 9, 50, 0.05, grid=grid)
    _, pv_b, _ = O.price(c, GBM_S, cfg_h, normals=z, want_pv=True)
    _, pv_a, _ = O.price(c, GBM_S, cfg_h, normals=-z, want_pv=True)
    pair = 0.5 * (pv_b + pv_a)
    assert got["n_paths"] == n // 2
    assert got["price"] == pytest.approx(pair.mean(), rel=2e-6, abs=1e-9)
    assert got["sample_std"] == pytest.approx(pair.std(ddof=1), rel=2e-5)
    assert got["std_error"] == pytest.approx(got["sample_std"] / math.sqrt(n // 2), rel=1e-12)
    # a pairing with the wrong partner moves the price by far more than that
    other = 0.5 * (pv_b + np.roll(pv_a, 1))
    assert abs(other.std(ddof=1) - pair.std(ddof=1)) > 1e-3 * pair.std(ddof=1)


def test_antithetic_falls_back_to_plain_paths_off_gbm():
    h = O.heston("S", 100.0, 0.05, 0.04, 2.0, 0.04, 0.3, -0.5)
    c = K.european_call("S", 100.0, 1.0).to_json()
    cfg = O.cfg(2000, 5, 24, 0.05)
    plain, _, _ = O.price(c, h, cfg)
    assert O.price_vr(c, h, cfg, antithetic=True) == plain
    assert O.price_vr(c, h, cfg, antithetic=True, control_variate=True) == \
        O.price_vr(c, h, cfg, control_variate=True)


# ---------------------------------------------------------------------------
# Control variate
# ---------------------------------------------------------------------------
def cv_by_hand(target, control, model, cfg, bs):
    """Both contracts on the target's paths: its grid, the same stream."""
    _, pt, grid = O.price(target, model, cfg, want_pv=True)
    on_grid = O.cfg(cfg.n_paths, cfg.seed, cfg.steps_per_year, cfg.rate, grid=grid)
    _, pc, _ = O.price(control, model, on_grid, want_pv=True)
    d = pt - pc
    return d.mean() + bs, d.std(ddof=1)


def test_control_variate_gbm():
    t = 1.0
    target = K.up_and_out_call("S", 95.0, 130.0, t).to_json()
    control = K.european_call("S", 100.0, t).to_json()
    cfg = O.cfg(4000, 3, 100, 0.05)
    got = O.price_vr(target, GBM_S, cfg, control_variate=True)
    bs = O.bs_call_as(100.0, 100.0, t, 0.05, 0.2)
    price, sd = cv_by_hand(target, control, GBM_S, cfg, bs)
    assert got["price"] == pytest.approx(price, rel=1e-10)
    assert got["sample_std"] == pytest.approx(sd, rel=1e-9)
    assert got["n_paths"] == 4000


def test_control_variate_defaults_off_gbm():
    # Heston: the control is the call struck at 100 on the model's asset,
    # corrected with BS(100, 100, T, r, 0.2), whatever the spot.
    h = O.heston("Y", 120.0, 0.03, 0.09, 1.5, 0.06, 0.4, -0.6)
    t = 0.75
    target = K.european_call("Y", 110.0, t).to_json()
    cfg = O.cfg(3000, 11, 40, 0.03)
    got = O.price_vr(target, h, cfg, control_variate=True)
    grid = O.compile_plan(target, 40, True)["grid"]
    assert grid[-1] == t
    price, sd = cv_by_hand(target, K.european_call("Y", 100.0, t).to_json(), h, cfg,
                           O.bs_call_as(100.0, 100.0, t, 0.03, 0.2))
    assert got["price"] == pytest.approx(price, rel=1e-10)
    assert got["sample_std"] == pytest.approx(sd, rel=1e-9)
    # correlated GBM: control on the first asset, defaults (100, 0.2)
    b = O.basket(["A", "B"], [90.0, 110.0], [0.02, 0.02], [0.3, 0.25], [[1.0, 0.4], [0.4, 1.0]])
    tgt = (K.scale((K.spot("B") - K.spot("A")).clip(0.0), K.one("USD")) @ K.at(1.0)).to_json()
    cfg = O.cfg(3000, 12, 12, 0.02)
    got = O.price_vr(tgt, b, cfg, control_variate=True)
    price, sd = cv_by_hand(tgt, K.european_call("A", 100.0, 1.0).to_json(), b, cfg,
                           O.bs_call_as(100.0, 100.0, 1.0, 0.02, 0.2))
    assert got["price"] == pytest.approx(price, rel=1e-10)


def test_antithetic_with_control_variate():
    t = 1.0
    target = (K.scale((K.average(K.spot("S")) - 98.0).clip(0.0), K.one("USD")) @ K.at(t)).to_json()
    control = K.european_call("S", 100.0, t).to_json()
    # the control rides on the target's grid
    cfg = O.cfg(5001, 21, 50, 0.05, grid=O.compile_plan(target, 50)["grid"])
    got = O.price_vr(target, GBM_S, cfg, antithetic=True, control_variate=True)
    at = O.price_vr(target, GBM_S, cfg, antithetic=True)
    ac = O.price_vr(control, GBM_S, cfg, antithetic=True)
    bs = O.bs_call_as(100.0, 100.0, t, 0.05, 0.2)
    assert got["price"] == pytest.approx(at["price"] - (ac["price"] - bs), rel=1e-12)
    assert got["sample_std"] == at["sample_std"]
    assert got["n_paths"] == 5000
    assert got["std_error"] == pytest.approx(at["sample_std"] / math.sqrt(5000), rel=1e-12)


def test_vr_edge_counts():
    c = K.european_call("S", 100.0, 1.0).to_json()
    bs = O.bs_call_as(100.0, 100.0, 1.0, 0.05, 0.2)
    r = O.price_vr(c, GBM_S, O.cfg(1, 1, 10, 0.05), antithetic=True)
    assert r["n_paths"] == 0 and r["price"] == 0.0 and r["std_error"] == 0.0
    r = O.price_vr(c, GBM_S, O.cfg(1, 1, 10, 0.05), antithetic=True, control_variate=True)
    assert r["n_paths"] == 0 and r["price"] == pytest.approx(bs, rel=1e-15) and math.isnan(r["std_error"])
    r = O.price_vr(c, GBM_S, O.cfg(0, 1, 10, 0.05), control_variate=True)
    assert r["n_paths"] == 0 and r["price"] == 0.0


# ---------------------------------------------------------------------------
# tests/j15_variance_reduction.rs acceptance criteria, on the oracle
# ---------------------------------------------------------------------------
def test_vr1_antithetic_reduces_spread():
    c = K.european_call("S", 100.0, 1.0).to_json()
    g = uniform_grid(1.0, 50)
    base, _, _ = O.price(c, GBM_S, O.cfg(20000, 42, 50, 0.05, grid=g))
    anti = O.price_vr(c, GBM_S, O.cfg(40000, 42, 50, 0.05, grid=g), antithetic=True)
    assert anti["sample_std"] < 0.75 * base["sample_std"]


def test_vr2_antithetic_price_matches_bs():
    c = K.european_call("S", 100.0, 1.0).to_json()
    r = O.price_vr(c, GBM_S, O.cfg(100000, 99, 50, 0.05, grid=uniform_grid(1.0, 50)), antithetic=True)
    bs = O.bs_call(100, 100, 1, 0.05, 0.2)
    assert abs(r["price"] - bs) / bs < 0.01


def test_vr3_optimal_beta_reduces_spread_on_knock_out():
    ko = K.up_and_out_call("S", 100.0, 130.0, 1.0).to_json()
    call = K.european_call("S", 100.0, 1.0).to_json()
    cfg = O.cfg(60000, 77, 100, 0.05, grid=uniform_grid(1.0, 100))
    _, y, _ = O.price(ko, GBM_S, cfg, want_pv=True)
    _, x, _ = O.price(call, GBM_S, cfg, want_pv=True)
    beta = np.cov(y, x, ddof=1)[0, 1] / np.var(x, ddof=1)
    assert np.std(y - beta * (x - x.mean()), ddof=1) < np.std(y, ddof=1)


def test_vr4_control_variate_unbiased_and_identity():
    c = K.european_call("S", 100.0, 1.0).to_json()
    cfg = O.cfg(100000, 2024, 50, 0.05, grid=uniform_grid(1.0, 50))
    r = O.price_vr(c, GBM_S, cfg, control_variate=True)
    bs_as = O.bs_call_as(100, 100, 1, 0.05, 0.2)
    # target == control: every per-path correction cancels
    assert r["price"] == pytest.approx(bs_as, abs=1e-9)
    assert r["sample_std"] < 1e-9
    assert abs(r["price"] - O.bs_call(100, 100, 1, 0.05, 0.2)) / r["price"] < 0.01
