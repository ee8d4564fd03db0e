"""GPU pathwise Greeks (kt_greeks_adjoint) against independent references.

* Closed forms: Black-Scholes delta and vega, and d/dmu = T s0 N(d1) at
  mu = r, for a European call under GBM.
* The oracle's FP64 central finite differences with common random numbers on
  the same Philox stream (the pattern of src/greeks.rs:60-98, extended to
  each model parameter): GBM European / Asian / lookback, correlated-GBM
  basket, Heston (floor and full-truncation schemes).  Pathwise and CRN
  finite differences evaluate the same paths, so they agree far inside the
  Monte Carlo error; the tolerance is a fraction of the standard error plus
  the finite-difference bias.
* The GPU's own bump Greeks, and the engine's price of the same contract.
"""
import ctypes as C
import math

import numpy as np
import pytest
import torch

import kontracts_amd as K
from kontracts_amd import _native as N
from kontracts_amd import autograd as AG
import oracle_lib as O

pytestmark = pytest.mark.gpu


def call(asset, k, t=1.0):
    return K.scale((K.spot(asset) - k).clip(0.0), K.one("USD")) @ K.at(t)


def put(asset, k, t=1.0):
    return K.scale((K.const_(k) - K.spot(asset)).clip(0.0), K.one("USD")) @ K.at(t)


def asian_call(asset, k, t=1.0):
    return K.scale((K.average(K.spot(asset)) - k).clip(0.0), K.one("USD")) @ K.at(t)


def lookback(asset, t=1.0):
    return K.scale(K.running_max(K.spot(asset)) - K.spot(asset), K.one("USD")) @ K.at(t)


def oracle_fd(json_s, make, params, h, n_paths, seed, spy, rate, grid=None):
    """Central CRN differences of oracle prices: d price / d params[i]."""
    out = []
    for i, hi in enumerate(h):
        up = list(params)
        dn = list(params)
        up[i] += hi
        dn[i] -= hi
        c = O.cfg(n_paths, seed, spy, rate, grid)
        pu = O.price(json_s, make(up), c)[0]["price"]
        pd = O.price(json_s, make(dn), c)[0]["price"]
        out.append((pu - pd) / (2 * hi))
    return out


def check_close(g, ref, fd_bias, label, frac=0.25):
    for name, v, se, r, b in zip(g.names, g.values, g.std_errors, ref, fd_bias):
        tol = frac * se + b
        print(f"{label} {name}: pathwise {v:.6f} (se {se:.2e})  reference {r:.6f}  |diff| {abs(v - r):.2e}  tol {tol:.2e}")
        assert abs(v - r) <= tol, (label, name)


def test_gbm_european_vs_black_scholes():
    s0, k, t, r, sg = 100.0, 100.0, 1.0, 0.05, 0.2
    m = K.GBM(s0, sg, r, asset="X")
    g = call("X", k).pathwise_greeks(m, n_paths=1 << 20, seed=42, grid=[0.0, 1.0])
    bs = O.bs_call_greeks(s0, k, t, r, sg)
    ref = [bs["delta"], bs["vega"], t * s0 * bs["delta"]]
    for name, v, se, rf in zip(g.names, g.values, g.std_errors, ref):
        print(f"{name}: {v:.6f} +- {se:.6f}  BS {rf:.6f}")
        assert abs(v - rf) <= 4 * se
    assert abs(g.price - O.bs_call(s0, k, t, r, sg)) <= 4 * g.price_std_error


def test_pathwise_price_equals_engine_price():
    m = K.GBM(100.0, 0.25, 0.03, asset="X")
    for c in (call("X", 95.0), asian_call("X", 100.0), lookback("X")):
        g = c.pathwise_greeks(m, n_paths=100_000, seed=5, steps_per_year=52)
        p = c.price(m, n_paths=100_000, seed=5, steps_per_year=52)
        # same paths and payoff; the pricer's straight-line groups and the
        # adjoint's interpreter may round the FP32 path state differently
        assert g.price == pytest.approx(p.price, rel=1e-7)
        assert g.price_std_error == pytest.approx(p.std_error, rel=1e-6)


@pytest.mark.parametrize("kind", ["european", "asian", "lookback"])
def test_gbm_vs_oracle_crn_fd(kind):
    s0, sg, r = 100.0, 0.3, 0.04
    mu = r - 0.01  # drift differs from the discount rate
    c = {"european": put("X", 105.0), "asian": asian_call("X", 100.0), "lookback": lookback("X")}[kind]
    m = K.GBM(s0, sg, r, q=r - mu, asset="X")
    n, seed, spy = 65536, 11, 52
    g = c.pathwise_greeks(m, n_paths=n, seed=seed, steps_per_year=spy)
    make = lambda p: O.gbm("X", p[0], p[2], p[1])
    h = [1e-3 * s0, 1e-4, 1e-4]
    ref = oracle_fd(c.to_json(), make, [s0, sg, mu], h, n, seed, spy, r)
    check_close(g, ref, [2e-3 * abs(x) + 1e-4 for x in ref], kind)


@pytest.mark.parametrize("grid_kind", ["equal", "unequal"])
def test_gbm_fixings_vs_oracle_crn_fd(grid_kind):
    """An average_at over fixings 12 points apart: the op-free groups between
    them run as loops of their own (one read of their run's end), from VGPR
    step rows on the equal grid and from the table on the unequal one, and
    each fixing goes through the op interpreter; against central CRN
    differences of the oracle on the same Philox stream."""
    s0, sg, r = 100.0, 0.3, 0.04
    mu = r - 0.01
    equal = [k / 96 for k in range(97)]
    fix = sorted([equal[k] for k in range(12, 97, 12)] + [equal[43]])
    grid = equal if grid_kind == "equal" else sorted(set(equal + [0.013, 0.377, 0.5011, 0.9]))
    c = K.scale((K.average_at(K.spot("X"), fix) - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)
    m = K.GBM(s0, sg, r, q=r - mu, asset="X")
    n, seed = 65536, 17
    g = c.pathwise_greeks(m, n_paths=n, seed=seed, grid=grid)
    make = lambda p: O.gbm("X", p[0], p[2], p[1])
    h = [1e-3 * s0, 1e-4, 1e-4]
    ref = oracle_fd(c.to_json(), make, [s0, sg, mu], h, n, seed, 50, r, grid=grid)
    check_close(g, ref, [2e-3 * abs(x) + 1e-4 for x in ref], f"fixings-{grid_kind}")


def test_basket_vs_oracle_crn_fd():
    names = ["A", "B", "C"]
    s0 = [100.0, 90.0, 110.0]
    sig = [0.2, 0.3, 0.25]
    mu = [0.03, 0.04, 0.05]
    corr = [[1.0, 0.3, 0.5], [0.3, 1.0, 0.2], [0.5, 0.2, 1.0]]
    r = 0.05
    fac = [K.GbmFactor(nm, s, m_, sg) for nm, s, m_, sg in zip(names, s0, mu, sig)]
    m = K.correlated_gbm(fac, corr, r)
    bask = (K.spot("A") + K.spot("B") * 2.0 - K.spot("C")) / 2.0
    c = K.scale((bask - 50.0).clip(0.0), K.one("USD")) @ K.at(1.0)
    n, seed = 65536, 3
    g = c.pathwise_greeks(m, n_paths=n, seed=seed, grid=[0.0, 0.5, 1.0])
    assert g.names == ["s0[A]", "s0[B]", "s0[C]", "sigma[A]", "sigma[B]", "sigma[C]", "mu[A]", "mu[B]", "mu[C]"]
    params = s0 + sig + mu

    def make(p):
        return O.basket(names, p[0:3], p[6:9], p[3:6], corr)
    h = [1e-3 * x for x in s0] + [1e-4] * 6
    ref = oracle_fd(c.to_json(), make, params, h, n, seed, 50, r, grid=[0.0, 0.5, 1.0])
    check_close(g, ref, [2e-3 * abs(x) + 1e-4 for x in ref], "basket")


@pytest.mark.parametrize("scheme", ["floor", "full_truncation"])
def test_heston_vs_oracle_crn_fd(scheme):
    p = dict(s0=100.0, v0=0.04, kappa=2.0, theta=0.04, sigma_v=0.3, rho=-0.5)
    r = 0.05
    m = K.heston(p["s0"], p["v0"], p["kappa"], p["theta"], p["sigma_v"], p["rho"], r, asset="X", scheme=scheme)
    c = call("X", 100.0)
    n, seed, spy = 32768, 42, 365
    g = c.pathwise_greeks(m, n_paths=n, seed=seed, steps_per_year=spy)
    h = [1e-3 * p["s0"], 1e-4, 1e-3, 1e-4, 1e-3, 1e-3]
    om = O.heston("X", p["s0"], r, p["v0"], p["kappa"], p["theta"], p["sigma_v"], p["rho"],
                  scheme=0 if scheme == "floor" else 1)
    ref = O.greeks_heston_bump(c.to_json(), om, O.cfg(n, seed, spy, r), h)
    assert g.price == pytest.approx(ref[0], rel=2e-3)
    check_close(g, ref[1:], [3e-3 * abs(x) + 2e-3 for x in ref[1:]], f"heston-{scheme}")


@pytest.mark.parametrize("grid_kind", ["equal", "unequal"])
def test_heston_fixings_vs_oracle_bump(grid_kind):
    """Heston pathwise Greeks of an average_at (fixings 12 points apart):
    the forward pass's op-free stretches, the checkpointed log-spot walk-back
    and the sweep with running accumulators (rows hoisted on the equal grid,
    read per point on the unequal one), against the oracle's CRN bump
    Greeks on the same Philox stream."""
    p = dict(s0=100.0, v0=0.04, kappa=2.0, theta=0.04, sigma_v=0.3, rho=-0.5)
    r = 0.05
    m = K.heston(p["s0"], p["v0"], p["kappa"], p["theta"], p["sigma_v"], p["rho"], r, asset="X")
    equal = [k / 96 for k in range(97)]
    fix = sorted([equal[k] for k in range(12, 97, 12)] + [equal[43]])
    grid = equal if grid_kind == "equal" else sorted(set(equal + [0.013, 0.377, 0.5011, 0.9]))
    c = K.scale((K.average_at(K.spot("X"), fix) - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)
    n, seed = 32768, 23
    g = c.pathwise_greeks(m, n_paths=n, seed=seed, grid=grid)
    h = [1e-3 * p["s0"], 1e-4, 1e-3, 1e-4, 1e-3, 1e-3]
    om = O.heston("X", p["s0"], r, p["v0"], p["kappa"], p["theta"], p["sigma_v"], p["rho"], scheme=0)
    ref = O.greeks_heston_bump(c.to_json(), om, O.cfg(n, seed, 50, r, grid=grid), h)
    assert g.price == pytest.approx(ref[0], rel=2e-3)
    check_close(g, ref[1:], [3e-3 * abs(x) + 2e-3 for x in ref[1:]], f"heston-fixings-{grid_kind}")


def test_heston_asian_and_lookback_vs_gpu_bump():
    # Feller holds (2 kappa theta = 0.12 > sigma_v^2 = 0.09).  Where it fails
    # the variance sits at the floor often, sqrt(v) is not Lipschitz there,
    # and the pathwise estimator is biased (about one standard error on
    # v0 / theta / sigma_v for a lookback at sigma_v = 0.4): bump Greeks apply.
    m = K.heston(100.0, 0.05, 1.5, 0.04, 0.3, -0.7, 0.03, asset="X")
    for c in (asian_call("X", 100.0), lookback("X"), put("X", 95.0)):
        g = c.pathwise_greeks(m, n_paths=1 << 18, seed=8, steps_per_year=100)
        b = c.bump_greeks(m, bumps=[0.1, 1e-4, 1e-3, 1e-4, 1e-3, 1e-3], n_paths=1 << 18, seed=8, steps_per_year=100)
        # the GPU's finite differences carry FP32 path rounding divided by 2h
        check_close(g, b.values, [3e-3 * abs(x) + 2e-3 for x in b.values], "heston-bump", frac=0.6)
        # same paths: the two estimators' spreads agree closely
        for se_b, se_g in zip(b.std_errors, g.std_errors):
            assert 0.8 <= se_b / se_g <= 1.25


def test_unsupported_payoffs_fall_back():
    m = K.GBM(100.0, 0.2, 0.05, asset="X")
    for c in (K.up_and_out_call("X", 100.0, 150.0, 1.0),
              K.cash_or_nothing_call("X", 100.0, 10.0, 1.0),
              call("X", 100.0, 0.5) + call("X", 100.0, 1.0)):
        with pytest.raises(N.KontractError) as ei:
            c.pathwise_greeks(m, n_paths=1024)
        assert ei.value.kind == "Unsupported"
    with pytest.raises(N.KontractError) as ei:
        call("Y", 100.0).pathwise_greeks(m, n_paths=1024)
    assert ei.value.kind == "UnknownAsset"
    g = call("X", 100.0).pathwise_greeks(m, n_paths=0)
    assert g.price == 0.0 and all(v == 0.0 for v in g.values)


def test_torch_autograd_function():
    c = call("X", 100.0)
    params = torch.tensor([100.0, 0.2, 0.05], dtype=torch.float64, requires_grad=True)
    price = AG.gbm_price(c, params, r=0.05, asset="X", n_paths=1 << 18, seed=1)
    price.backward()
    g = c.pathwise_greeks(K.GBM(100.0, 0.2, 0.05, asset="X"), n_paths=1 << 18, seed=1)
    assert price.item() == pytest.approx(g.price, rel=1e-12)
    np.testing.assert_allclose(params.grad.numpy(), g.values, rtol=1e-12)
    hp = torch.tensor([100.0, 0.04, 2.0, 0.04, 0.3, -0.5], dtype=torch.float64, requires_grad=True)
    (2.0 * AG.heston_price(c, hp, r=0.05, asset="X", n_paths=1 << 16, seed=2, steps_per_year=100)).backward()
    gh = c.pathwise_greeks(K.heston(100.0, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X"), n_paths=1 << 16, seed=2,
                           steps_per_year=100)
    np.testing.assert_allclose(hp.grad.numpy(), 2.0 * np.array(gh.values), rtol=1e-12)


def test_heston_checkpoint_interval_option():
    eng = N.engine(0)
    m = K.heston(100.0, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X")
    res = {}
    try:
        for seg in (16, 32):
            eng.set_option("adjoint_segment", seg)
            res[seg] = [c.pathwise_greeks(m, n_paths=1 << 16, seed=4, steps_per_year=365)
                        for c in (call("X", 100.0), asian_call("X", 100.0))]
    finally:
        eng.set_option("adjoint_segment", 16)
    # same recomputed variance path either way; the log-spot walked back from a
    # checkpoint differs only by FP32 rounding over a different stretch
    np.testing.assert_allclose(res[16][0].values, res[32][0].values, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(res[16][1].values, res[32][1].values, rtol=1e-4, atol=1e-6)
    for bad in (("adjoint_segment", 8), ("no_such_option", 1)):
        with pytest.raises(N.KontractError) as ei:
            eng.set_option(*bad)
        assert ei.value.kind == "InvalidArgument"


def test_heston_tangent_mode_equals_reverse_sweep():
    """The forward-tangent mode computes the same pathwise derivatives as the
    reverse sweep (same paths; only the order of FP32 operations differs),
    for both schemes, with and without running averages / extremes."""
    eng = N.engine(0)
    models = [K.heston(100.0, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X"),
              K.heston(100.0, 0.05, 1.5, 0.04, 0.3, -0.7, 0.03, asset="X", scheme="full_truncation"),
              K.heston(100.0, 0.04, 2.0, 0.04, 0.9, -0.7, 0.05, asset="X")]  # Feller violated: floor active
    cons = [call("X", 100.0), asian_call("X", 100.0), lookback("X"), put("X", 95.0)]
    try:
        for m in models:
            for c in cons:
                eng.set_option("pathwise_mode", 0)
                a = c.pathwise_greeks(m, n_paths=1 << 15, seed=5, steps_per_year=120)
                eng.set_option("pathwise_mode", 1)
                b = c.pathwise_greeks(m, n_paths=1 << 15, seed=5, steps_per_year=120)
                assert b.price == a.price  # the forward pass is the same
                for name, x, y, se in zip(a.names, a.values, b.values, a.std_errors):
                    assert abs(x - y) <= 1e-3 * abs(x) + 0.02 * se, (name, x, y, se)
        # the portfolio's tangent variants equal per-contract tangent Greeks
        jobs = [(c, models[0], 3000 + 500 * i, i, 50) for i, c in enumerate(cons)]
        got = K.price_portfolio(jobs, greeks=True)
        for (pr, g), (c, m, n, seed, spy) in zip(got, jobs):
            w = c.pathwise_greeks(m, n_paths=n, seed=seed, steps_per_year=spy)
            assert g.values == w.values and pr.price == w.price
    finally:
        eng.set_option("pathwise_mode", 0)
    with pytest.raises(N.KontractError):
        eng.set_option("pathwise_mode", 2)


def test_heston_sweep_on_equal_and_unequal_steps():
    """The reverse sweep reads its step rows from registers (or rows hoisted
    out of its loops, with running averages) on a grid of equal steps and
    from the tables otherwise, and the forward pass runs the op-free stretches
    between an average's fixings from registers on the equal grid; on either
    grid it must agree with the forward-tangent mode (same paths, FP32
    rounding only), and the price with the plain pricer's."""
    eng = N.engine(0)
    m = K.heston(100.0, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X")
    equal = [i / 96 for i in range(97)]
    unequal = sorted(set([i / 96 for i in range(97)] + [0.013, 0.377, 0.5011, 0.9]))
    # fixings every 8th point of the equal grid (runs of op-free groups
    # between them), and one off the 4-point groups' boundaries
    fix = [equal[k] for k in range(8, 97, 8)] + [equal[43]]
    monthly = K.scale((K.average_at(K.spot("X"), sorted(fix)) - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)
    # (contract, relative and standard-error tolerances): with running
    # averages the log-spot is walked back from checkpoints, so the two modes
    # differ by FP32 rounding over a checkpoint interval
    cases = [(call("X", 100.0), 1e-5, 1e-3), (monthly, 1e-4, 2e-3), (asian_call("X", 100.0), 1e-4, 2e-3)]
    try:
        for c, rel, sef in cases:
            for grid in (equal, unequal):
                eng.set_option("pathwise_mode", 0)
                a = c.pathwise_greeks(m, n_paths=1 << 15, seed=13, grid=grid)
                eng.set_option("pathwise_mode", 1)
                b = c.pathwise_greeks(m, n_paths=1 << 15, seed=13, grid=grid)
                assert a.price == b.price
                for name, x, y, se in zip(a.names, a.values, b.values, a.std_errors):
                    assert abs(x - y) <= rel * abs(x) + sef * se, (len(grid), name, x, y)
                p = c.price(m, n_paths=1 << 15, seed=13, grid=grid)
                assert a.price == pytest.approx(p.price, rel=1e-6)
    finally:
        eng.set_option("pathwise_mode", 0)


@pytest.mark.parametrize("n", [1, 257, 1000])
def test_ragged_path_counts(n):
    """Path counts that leave a block partly empty: the pathwise launch's
    price is the pricer's on the same paths, one path gives zero standard
    errors, and the Heston reverse sweep (checkpoints indexed by global path)
    agrees with the forward-tangent mode."""
    eng = N.engine(0)
    gbm = K.GBM(100.0, 0.25, 0.03, asset="X")
    hes = K.heston(100.0, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X")
    fac = [K.GbmFactor(nm, 100.0, 0.03, 0.2 + 0.05 * i) for i, nm in enumerate("ABC")]
    bsk = K.correlated_gbm(fac, [[1.0 if i == j else 0.3 for j in range(3)] for i in range(3)], 0.03)
    bc = K.scale(((K.spot("A") + K.spot("B") + K.spot("C")) / 3.0 - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)
    for c, m, spy in ((call("X", 95.0), gbm, 52), (asian_call("X", 100.0), hes, 120), (bc, bsk, 50)):
        g = c.pathwise_greeks(m, n_paths=n, seed=3, steps_per_year=spy)
        p = c.price(m, n_paths=n, seed=3, steps_per_year=spy)
        assert g.price == pytest.approx(p.price, rel=1e-6, abs=1e-9)
        assert all(np.isfinite(g.values))
        if n == 1:
            assert g.price_std_error == 0.0 and all(se == 0.0 for se in g.std_errors)
    try:
        eng.set_option("pathwise_mode", 0)
        a = asian_call("X", 100.0).pathwise_greeks(hes, n_paths=n, seed=3, steps_per_year=120)
        eng.set_option("pathwise_mode", 1)
        b = asian_call("X", 100.0).pathwise_greeks(hes, n_paths=n, seed=3, steps_per_year=120)
        assert a.price == b.price
        for name, x, y in zip(a.names, a.values, b.values):
            assert abs(x - y) <= 1e-4 * max(1.0, abs(x)), (n, name, x, y)
    finally:
        eng.set_option("pathwise_mode", 0)
