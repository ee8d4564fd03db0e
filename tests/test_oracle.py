"""CPU tests that pin the oracle (oracle/oracle.cpp) before it is trusted.

Each group checks the oracle against something it cannot simply restate:
the Random123 known-answer vectors, scipy's ndtri, the reference's own golden
values and deterministic known-answer tests (kontract tests/*.rs), closed forms,
limits and parity relations.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy.special import ndtri

import oracle_lib as O
import kontracts_amd as K

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------------------
# RNG
# ---------------------------------------------------------------------------
def _philox_py(c, k):
    """Independent restatement from the SC'11 paper, in Python integers."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c, k = list(c), list(k)
    for r in range(10):
        if r:
            k = [(k[0] + W0) & 0xFFFFFFFF, (k[1] + W1) & 0xFFFFFFFF]
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [(p1 >> 32) ^ c[1] ^ k[0], p1 & 0xFFFFFFFF, (p0 >> 32) ^ c[3] ^ k[1], p0 & 0xFFFFFFFF]
    return c


def test_philox_kat_vectors():
    for v in golden("philox_kat.json")["vectors"]:
        ctr = [int(x, 16) for x in v["ctr"]]
        key = [int(x, 16) for x in v["key"]]
        assert O.philox(ctr, key) == [int(x, 16) for x in v["out"]]


def test_philox_python_independent():
    rng = np.random.default_rng(1)
    for v in golden("philox_kat.json")["vectors"]:
        assert _philox_py([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]]) == \
            [int(x, 16) for x in v["out"]]
    for _ in range(200):
        ctr = [int(x) for x in rng.integers(0, 2**32, 4)]
        key = [int(x) for x in rng.integers(0, 2**32, 2)]
        assert O.philox(ctr, key) == _philox_py(ctr, key)


def test_uniform_mapping_is_cell_midpoint():
    L = O.load()
    for w in [0, 1, 255, 256, 2**31, 2**32 - 1]:
        u = L.orc_uniform_from_word(w)
        assert u == (2 * (w >> 8) + 1) / 2**25
    assert 0.0 < L.orc_uniform_from_word(0) < L.orc_uniform_from_word(2**32 - 1) < 1.0


def test_inverse_normal_f64_matches_ndtri():
    L = O.load()
    us = np.concatenate([np.linspace(1e-8, 1 - 1e-8, 2001), [2**-25, 1 - 2**-25, 0.5, 0.02425, 0.97575]])
    got = np.array([L.orc_inv_norm(float(u)) for u in us])
    np.testing.assert_allclose(got, ndtri(us), rtol=0, atol=1e-12)


def test_oracle_normals_use_philox_words():
    # normal j of a path = Phi^-1(u(word j%4 of Philox block j/4))
    seed, path = 42, 7
    z = O.normals(seed, path, 0, 8)
    L = O.load()
    for j in range(8):
        words = O.philox([j // 4, path, 0, 0], [seed, 0])
        assert z[j] == pytest.approx(float(ndtri(L.orc_uniform_from_word(words[j % 4]))), abs=1e-12)


# ---------------------------------------------------------------------------
# Closed forms
# ---------------------------------------------------------------------------
def test_black_scholes_golden_value():
    g = golden("reference_values.json")["bs_call_s100_k100_r005_sig02_t1"]["value"]
    assert abs(O.bs_call(100, 100, 1, 0.05, 0.2) - g) < 1e-12


def test_black_scholes_parity_and_greeks():
    for s, k, t, r, sg in [(100, 90, 0.5, 0.03, 0.3), (36, 40, 1, 0.06, 0.2), (180, 100, 2, 0.01, 0.45)]:
        c, p = O.bs_call(s, k, t, r, sg), O.bs_put(s, k, t, r, sg)
        assert c - p == pytest.approx(s - k * math.exp(-r * t), abs=1e-10)
        g = O.bs_call_greeks(s, k, t, r, sg)
        h = 1e-4
        assert g["delta"] == pytest.approx((O.bs_call(s + h, k, t, r, sg) - O.bs_call(s - h, k, t, r, sg)) / (2 * h),
                                           rel=1e-6)
        assert g["vega"] == pytest.approx((O.bs_call(s, k, t, r, sg + h) - O.bs_call(s, k, t, r, sg - h)) / (2 * h),
                                          rel=1e-6)
        assert g["rho"] == pytest.approx((O.bs_call(s, k, t, r + h, sg) - O.bs_call(s, k, t, r - h, sg)) / (2 * h),
                                         rel=1e-6)
        hh = 1e-2
        assert g["gamma"] == pytest.approx(
            (O.bs_call(s + hh, k, t, r, sg) - 2 * c + O.bs_call(s - hh, k, t, r, sg)) / hh**2, rel=1e-4)


def test_heston_literature_value():
    ref = golden("reference_values.json")["heston_fang_oosterlee"]
    p = ref["params"]
    v = O.heston_call(p["s0"], p["k"], p["t"], p["r"], p["v0"], p["kappa"], p["theta"], p["sigma_v"], p["rho"])
    assert v == pytest.approx(ref["value"], abs=2e-6)


def test_heston_deterministic_variance_limit():
    # sigma_v -> 0: v(t) = theta + (v0-theta) e^{-kappa t}; price = BS with the integrated variance.
    s0, k, t, r, v0, kappa, theta = 100.0, 95.0, 1.5, 0.03, 0.09, 1.3, 0.04
    iv = theta * t + (v0 - theta) * (1 - math.exp(-kappa * t)) / kappa
    bs = O.bs_call(s0, k, t, r, math.sqrt(iv / t))
    assert O.heston_call(s0, k, t, r, v0, kappa, theta, 1e-5, 0.3) == pytest.approx(bs, abs=1e-5)


def test_heston_skew_direction_and_bounds():
    # For an out-of-the-money call, positive spot-vol correlation fattens the
    # right tail and raises the price; both stay within no-arbitrage bounds.
    args = (100.0, 130.0, 1.0, 0.05, 0.04, 2.0, 0.04, 0.3)
    c_neg = O.heston_call(*args, -0.5)
    c_pos = O.heston_call(*args, 0.5)
    assert c_neg < c_pos
    for c in (c_neg, c_pos):
        assert max(0.0, 100 - 130 * math.exp(-0.05)) < c < 100


def test_up_and_out_formula_limits():
    s, k, t, r, sg = 100.0, 100.0, 1.0, 0.05, 0.2
    assert O.uo_call(s, k, 1e9, t, r, sg) == pytest.approx(O.bs_call(s, k, t, r, sg), abs=1e-9)
    assert O.uo_call(s, k, 100.0, t, r, sg) == pytest.approx(0.0, abs=1e-9)
    vals = [O.uo_call(s, k, h, t, r, sg) for h in (110, 130, 160, 250)]
    assert all(a < b for a, b in zip(vals, vals[1:]))


def test_up_and_out_formula_vs_fine_monitoring_mc():
    # Discrete monitoring at 2000 steps, shifted by Broadie-Glasserman-Kou,
    # agrees with the continuous formula.
    s, k, h, t, r, sg = 100.0, 100.0, 125.0, 1.0, 0.05, 0.2
    steps = 2000
    c = K.up_and_out_call("X", k, h, t)
    res, _, _ = O.price(c.to_json(), O.gbm("X", s, r, sg), O.cfg(20000, 3, steps, r))
    h_eff = h * math.exp(-0.5826 * sg * math.sqrt(t / steps))
    ana = O.uo_call(s, k, h_eff, t, r, sg)
    assert abs(res["price"] - ana) < 4 * res["std_error"] + 0.01 * ana


def test_crr_tree():
    ref = golden("reference_values.json")["american_put_ls2001"]
    p = ref["params"]
    put = O.crr_american(p["s0"], p["k"], p["t"], p["r"], p["sigma"], 3000, False)
    assert put == pytest.approx(ref["american"], abs=2e-3)
    call = O.crr_american(100, 100, 1, 0.05, 0.2, 3000, True)
    assert call == pytest.approx(O.bs_call(100, 100, 1, 0.05, 0.2), abs=5e-3)


# ---------------------------------------------------------------------------
# Reference known-answer tests on hand-built paths (tests/observable.rs,
# tests/j26_temporal_obs.rs)
# ---------------------------------------------------------------------------
TIMES3 = [0.0, 0.5, 1.0]
SPOTS2 = {"AAPL": [100.0, 110.0, 150.0], "MSFT": [200.0, 190.0, 210.0]}


def ev(obs, t, times=TIMES3, spots=SPOTS2):
    return O.eval_observable(json.dumps(obs.node), times, spots, t)


def test_observable_known_answers():
    S = K.spot
    for t in range(3):
        assert ev(K.const_(42.0), t) == 42.0
    assert ev(S("AAPL"), 0) == 100.0 and ev(S("AAPL"), 2) == 150.0 and ev(S("MSFT"), 1) == 190.0
    assert ev((S("AAPL") + S("MSFT")) / K.const_(2.0), 2) == 180.0
    assert ev(S("AAPL") * K.const_(2.0) - K.const_(50.0), 1) == 170.0
    assert ev(-S("AAPL"), 0) == -100.0
    call = (S("AAPL") - 120.0).max(0.0)
    assert ev(call, 2) == 30.0 and ev(call, 0) == 0.0
    assert ev((K.const_(120.0) - S("AAPL")).max(0.0), 0) == 20.0
    assert ev(S("AAPL").min(130.0), 2) == 130.0
    assert ev((S("AAPL") - S("MSFT")).max(0.0), 2) == 0.0


def test_observable_errors():
    with pytest.raises(O.OracleError) as e:
        ev(K.spot("TSLA"), 0)
    assert e.value.code == 1  # UnknownAsset
    with pytest.raises(O.OracleError) as e:
        ev(K.spot("AAPL"), 99)
    assert e.value.code == 3  # StepOutOfRange


def test_temporal_known_answers():
    t5 = [0.0, 0.25, 0.5, 0.75, 1.0]
    for t in range(5):
        assert abs(ev(K.average(K.const_(7.0)), t, t5, {"X": [100, 105, 98, 102, 110]}) - 7.0) < 1e-14
        assert abs(ev(K.running_max(K.spot("S")), t, t5, {"S": [1, 2, 3, 4, 5]}) - (t + 1)) < 1e-14
        assert abs(ev(K.running_min(K.spot("S")), t, t5, {"S": [5, 4, 3, 2, 1]}) - (5 - t)) < 1e-14
    assert abs(ev(K.average_over(K.spot("S"), 0.25, 0.75), 4, t5, {"S": [1, 2, 3, 4, 5]}) - 3.0) < 1e-12
    vals = [90.0, 95.0, 100.0, 105.0, 110.0]
    arith = ev(K.average(K.spot("S")), 4, t5, {"S": vals})
    assert arith >= math.exp(np.mean(np.log(vals))) - 1e-10
    # Average includes t=0 and averages up to the evaluation step.
    assert ev(K.average(K.spot("S")), 2, t5, {"S": [1, 2, 3, 4, 5]}) == pytest.approx(2.0)
    with pytest.raises(O.OracleError) as e:  # window from 0.8 evaluated at t=0.5: from > to
        ev(K.Observable({"Average": {"obs": {"Spot": "S"}, "from_year": 0.8, "to_year": None}}), 2, t5,
           {"S": [1, 2, 3, 4, 5]})
    assert e.value.code == 6


def test_average_at_extension_known_answers():
    """AverageAt (extension, no reference counterpart): mean over the fixings
    on or before the evaluation step.  Pinned by hand-computed means, by its
    reduction to the reference Average when every grid point is a fixing, and
    by its error cases."""
    t5 = [0.0, 0.25, 0.5, 0.75, 1.0]
    sp = {"S": [1.0, 2.0, 4.0, 8.0, 16.0]}
    avg_at = lambda dates: K.average_at(K.spot("S"), dates)
    assert ev(avg_at([0.25, 0.75]), 4, t5, sp) == (2.0 + 8.0) / 2
    assert ev(avg_at([0.25, 0.75]), 2, t5, sp) == 2.0  # only the fixing at 0.25 is past
    assert ev(avg_at([1.0]), 4, t5, sp) == 16.0
    for t in range(5):  # every grid point up to now is a fixing: the reference Average
        assert ev(avg_at(t5), t, t5, sp) == ev(K.average(K.spot("S")), t, t5, sp)
    cases = [([0.25, 0.75], 0, 6),      # no fixing on or before t = 0
             ([], 4, 6),                # no fixings at all
             ([0.75, 0.25], 4, 6),      # not increasing
             ([0.3], 4, 2)]             # fixing date not on the grid: TimeOutOfRange
    for dates, t, code in cases:
        with pytest.raises(O.OracleError) as e:
            ev(avg_at(dates), t, t5, sp)
        assert e.value.code == code, dates


def test_cashflow_semantics_on_hand_built_path():
    times = [0.0, 0.5, 1.0, 1.5]
    spots = {"X": [100.0, 120.0, 90.0, 130.0]}
    # scale samples at the flow date (src/pricer.rs:389-393)
    pv, flows = O.eval_path(K.scale(K.spot("X"), K.one("USD") @ K.at(1.0)).to_json(), times, spots)
    assert flows == [(90.0, 2)]
    # until keeps flows strictly before the first activation
    c = K.until(K.spot("X") >= 120.0, (K.one("USD") @ K.at(0.5)) + (K.one("USD") @ K.at(1.0)))
    assert O.eval_path(c.to_json(), times, spots)[1] == []  # flow at 0.5 coincides with the hit
    c = K.until(K.spot("X") >= 125.0, (K.one("USD") @ K.at(0.5)) + (K.one("USD") @ K.at(1.0)))
    assert O.eval_path(c.to_json(), times, spots)[1] == [(1.0, 1), (1.0, 2)]
    # when nested under when starts scanning at the outer activation
    c = K.when(K.spot("X") >= 110.0, K.when(K.spot("X") <= 95.0, K.one("USD")))
    assert O.eval_path(c.to_json(), times, spots)[1] == [(1.0, 2)]
    # give negates, and concatenates
    c = -(K.one("USD") @ K.at(1.5)) + K.one("USD")
    assert O.eval_path(c.to_json(), times, spots)[1] == [(-1.0, 3), (1.0, 0)]
    # discounting uses the flow's grid time
    pv, _ = O.eval_path((K.one("USD") @ K.at(1.5)).to_json(), times, spots, rate=0.1)
    assert pv == pytest.approx(math.exp(-0.15), abs=1e-15)


# ---------------------------------------------------------------------------
# Compiler (tests/compiler.rs)
# ---------------------------------------------------------------------------
def ecall(asset, k, t):
    return K.european_call(asset, k, t)


def test_compiler_reference_cases():
    p = O.compile_plan((K.one("USD") @ K.at(1.0)).to_json(), 50)
    assert p["assets"] == [] and p["fixed_dates"] == [1.0] and p["horizon"] == 1.0 and not p["needs_fine_grid"]
    assert p["grid"] == [0.0, 1.0]
    p = O.compile_plan(ecall("AAPL", 150, 1.0).to_json())
    assert p["assets"] == ["AAPL"] and not p["needs_fine_grid"]
    ko = K.until(K.spot("AAPL") >= 200.0, ecall("AAPL", 150, 1.0))
    p = O.compile_plan(ko.to_json(), 12)
    assert p["needs_fine_grid"] and len(p["grid"]) == 13 and p["grid"][0] == 0.0 and p["grid"][-1] == 1.0
    basket = K.scale(K.spot("MSFT") + K.spot("AAPL"), K.one("USD")) @ K.at(1.0)
    assert O.compile_plan(basket.to_json())["assets"] == ["AAPL", "MSFT"]
    comp = (K.one("USD") @ K.at(0.5)) + K.until(K.spot("AAPL") >= 150.0, ecall("AAPL", 120, 1.0))
    p = O.compile_plan(comp.to_json(), 10)
    assert p["fixed_dates"] == [0.5, 1.0] and p["needs_fine_grid"]
    assert any(abs(t - 0.5) < 1e-12 for t in p["grid"])
    p = O.compile_plan(K.zero().to_json(), 12)
    assert p["assets"] == [] and p["fixed_dates"] == [] and p["horizon"] == 0.0 and p["grid"] == [0.0]
    amer = K.anytime(K.spot("AAPL") <= 80.0, K.one("USD") @ K.at(2.0))
    p = O.compile_plan(amer.to_json())
    assert p["needs_fine_grid"] and p["horizon"] == 2.0
    with pytest.raises(O.OracleError) as e:
        O.compile_plan((K.one("USD") @ K.at(-1.0)).to_json())
    assert e.value.code == 6
    dup = (K.one("USD") @ K.at(1.0)) + (K.one("EUR") @ K.at(1.0))
    assert O.compile_plan(dup.to_json())["fixed_dates"] == [1.0]


# ---------------------------------------------------------------------------
# Pricer semantics with deterministic models (tests/pricer.rs, barriers.rs,
# diagnostics.rs, portfolio.rs)
# ---------------------------------------------------------------------------
def flat():
    return O.gbm("AAPL", 100.0, 0.05, 0.0)


def det_cfg(rate):
    return O.cfg(16, 1, 4, rate)


def oprice(c, model, cfg):
    return O.price(c.to_json(), model, cfg)[0]


def test_pricer_known_answers():
    assert oprice(K.zero(), flat(), det_cfg(0.05))["price"] == 0.0
    assert abs(oprice(K.one("USD"), flat(), det_cfg(0.05))["price"] - 1.0) < 1e-12
    assert abs(oprice(-K.one("USD"), flat(), det_cfg(0.05))["price"] + 1.0) < 1e-12
    assert abs(oprice(K.one("USD") + (K.one("USD") + K.one("USD")), flat(), det_cfg(0.0))["price"] - 3.0) < 1e-12
    assert abs(oprice(K.scale(K.const_(7.5), K.one("USD")), flat(), det_cfg(0.0))["price"] - 7.5) < 1e-12
    assert abs(oprice(K.one("USD") @ K.at(1.0), flat(), det_cfg(0.05))["price"] - math.exp(-0.05)) < 1e-12
    c = K.scale(K.spot("AAPL"), K.one("USD") @ K.at(1.0))
    assert abs(oprice(c, flat(), det_cfg(0.05))["price"] - 100.0) < 1e-9
    with pytest.raises(O.OracleError) as e:
        oprice(K.one("USD").or_(K.one("USD")), flat(), det_cfg(0.0))
    assert e.value.code == 7


def test_barrier_known_answers():
    amer = K.anytime(K.spot("AAPL") >= 105.0, K.one("USD") @ K.at(2.0))
    assert abs(oprice(amer, flat(), O.cfg(8, 1, 12, 0.05))["price"] - math.exp(-0.1)) < 1e-9
    never = K.anytime(K.spot("AAPL") >= 200.0, K.one("USD") @ K.at(2.0))
    assert oprice(never, flat(), O.cfg(8, 1, 12, 0.05))["price"] == 0.0
    ko = K.until(K.spot("AAPL") <= 200.0, ecall("AAPL", 100.0, 1.0))
    assert oprice(ko, O.gbm("AAPL", 100, 0.05, 0.2), O.cfg(1000, 1, 10, 0.05))["price"] == 0.0


def test_diagnostics_known_answers():
    res = oprice(K.one("USD") @ K.at(1.0), O.gbm("AAPL", 100, 0.05, 0.0), O.cfg(1000, 1, 1, 0.05))
    assert res["sample_std"] < 1e-12 and res["std_error"] < 1e-12 and res["n_paths"] == 1000
    res = oprice(ecall("AAPL", 100, 1.0), O.gbm("AAPL", 100, 0.05, 0.2), O.cfg(20000, 7, 1, 0.05))
    assert res["std_error"] > 0 and res["ci95_low"] < res["price"] < res["ci95_high"]
    assert abs(res["ci95_high"] - res["ci95_low"] - 2 * 1.959963984540054 * res["std_error"]) < 1e-9


def test_summarize_two_pass():
    x = np.random.default_rng(3).normal(5, 2, 1001)
    r = O.summarize(x)
    assert r["price"] == pytest.approx(x.mean(), rel=1e-14)
    assert r["sample_std"] == pytest.approx(x.std(ddof=1), rel=1e-12)
    assert r["std_error"] == pytest.approx(x.std(ddof=1) / math.sqrt(1001), rel=1e-12)
    assert O.summarize([])["price"] == 0.0 and O.summarize([3.0])["sample_std"] == 0.0


def test_batch_equals_individual():
    model = O.gbm("AAPL", 100, 0.05, 0.2)
    cfg = O.cfg(5000, 2024, 1, 0.05)
    cons = [ecall("AAPL", float(k), 1.0) for k in range(80, 121, 8)]
    batch = O.price_batch([c.to_json() for c in cons], model, cfg)
    for c, b in zip(cons, batch):
        assert abs(oprice(c, model, cfg)["price"] - b["price"]) < 1e-12
    assert all(a["price"] >= b["price"] - 1e-9 for a, b in zip(batch, batch[1:]))


def test_european_call_vs_black_scholes():
    res = oprice(ecall("AAPL", 100, 1.0), O.gbm("AAPL", 100, 0.05, 0.2), O.cfg(200000, 2024, 1, 0.05))
    bs = O.bs_call(100, 100, 1, 0.05, 0.2)
    assert abs(res["price"] - bs) < 3.5 * res["std_error"]


def test_gbm_fine_grid_moments():
    # exact log step: E[S_T] = S0 e^{mu T}, independent of the number of steps
    c = K.scale(K.spot("X"), K.one("USD")) @ K.at(1.0)
    grid = [i / 50 for i in range(51)]
    res, _, g = O.price(c.to_json(), O.gbm("X", 100, 0.05, 0.2), O.cfg(100000, 11, 0, 0.0, grid=grid))
    assert g == grid
    assert abs(res["price"] - 100 * math.exp(0.05)) < 3.5 * res["std_error"]


def test_heston_oracle_vs_closed_form():
    # sigma_v = 0: the variance follows its deterministic ODE -> BS with integrated variance (Euler bias small)
    c = ecall("X", 100, 1.0)
    res = oprice(c, O.heston("X", 100, 0.05, 0.04, 2.0, 0.04, 0.0, -0.5), O.cfg(100000, 5, 100, 0.05))
    assert abs(res["price"] - O.bs_call(100, 100, 1, 0.05, 0.2)) < 3.5 * res["std_error"]


def test_basket_oracle_perfect_correlation():
    # rho = 1, equal parameters: (S1+S2)/2 = S1 path by path (tests/j27_basket.rs)
    b = O.basket(["S1", "S2"], [100, 100], [0.03, 0.03], [0.2, 0.2], [[1, 1], [1, 1]])
    c = K.scale(((K.spot("S1") + K.spot("S2")) / 2.0 - 100.0).max(0.0), K.one("USD")) @ K.at(1.0)
    res = oprice(c, b, O.cfg(50000, 3, 50, 0.03))
    single = oprice(ecall("S1", 100, 1.0), b, O.cfg(50000, 3, 50, 0.03))
    assert abs(res["price"] - single["price"]) < 1e-9


# ---------------------------------------------------------------------------
# LSM (tests/j17_lsm.rs)
# ---------------------------------------------------------------------------
def test_solve_linear_matches_numpy_and_singular_gives_zero():
    import ctypes as C
    L = O.load()
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        b = rng.normal(size=3)
        x = (C.c_double * 3)()
        L.orc_solve_linear((C.c_double * 9)(*a.reshape(-1)), (C.c_double * 3)(*b), 3, x)
        np.testing.assert_allclose(list(x), np.linalg.solve(a, b), rtol=1e-12, atol=1e-12)
    x = (C.c_double * 2)()
    L.orc_solve_linear((C.c_double * 4)(1e-13, 0, 0, 1e-13), (C.c_double * 2)(1, 1), 2, x)
    assert list(x) == [0.0, 0.0]


def test_lsm_put_vs_crr():
    s0, k, t, r, sg = 100.0, 100.0, 1.0, 0.06, 0.25
    payoff = K.scale((K.const_(k) - K.spot("AAPL")).max(0.0), K.one("USD"))
    dates = [t * i / 50 for i in range(1, 51)]
    res, _ = O.price_american(payoff.to_json(), dates, O.gbm("AAPL", s0, r, sg), O.cfg(60000, 7, 50, r))
    crr = O.crr_american(s0, k, t, r, sg, 500, False)
    assert abs(res["price"] - crr) / crr < 0.01
    assert res["price"] > O.bs_put(s0, k, t, r, sg)
