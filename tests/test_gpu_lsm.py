"""GPU parity of the Longstaff-Schwartz path (src/lsm.rs) through the C ABI.

Mirrors the reference's own LSM tests (tests/j17_lsm.rs: American call ==
European call, put above the European put, call and put within 1% of a
500-step CRR tree, n_basis 2 vs 3) and adds parity with the oracle on the
same Philox stream: same seed, same paths, same regression.  The GPU stores
spots and payoffs in FP32 and sums the normal equations in a different order
than the oracle, so an exercise decision can flip on a path whose payoff sits
within rounding of the fitted continuation value; everything else is equal
to within FP32 rounding of the payoff.
"""
import ctypes as C
import math

import numpy as np
import pytest

import kontracts_amd as K
from kontracts_amd import _native as N
from kontracts_amd.dsl import _cfg, _model_struct
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return N.engine(0)


def put_payoff(asset, k):
    return K.scale((K.const_(k) - K.spot(asset)).max(0.0), K.one("USD"))


def call_payoff(asset, k):
    return K.scale((K.spot(asset) - K.const_(k)).max(0.0), K.one("USD"))


def dates_uniform(t, n):
    return [t * i / n for i in range(1, n + 1)]


def gpu_american(eng, contract, model, dates, n_paths, seed, n_basis=3):
    m = _model_struct(model)
    cfg = _cfg(n_paths, seed, 50, model.rate, None)
    d = (C.c_double * max(len(dates), 1))(*dates)
    pv = np.zeros(max(n_paths, 1), dtype=np.float64)
    coef = np.zeros(max(len(dates) - 1, 1) * (n_basis + 1), dtype=np.float64)
    res = N.KtPriceResult()
    N.check(N.lib().kt_debug_american(eng.handle, contract.to_json().encode(), d, len(dates), C.byref(m.s),
                                      C.byref(cfg.s), n_basis, pv.ctypes.data_as(C.POINTER(C.c_double)),
                                      coef.ctypes.data_as(C.POINTER(C.c_double)), C.byref(res)), eng.handle)
    return K.PriceResult(res), pv[:n_paths], coef.reshape(-1, n_basis + 1)[: max(len(dates) - 1, 0)]


def compare_with_oracle(eng, contract, model, omodel, dates, n_paths, seed, n_basis=3, flip_budget=0.01):
    g, gpv, _ = gpu_american(eng, contract, model, dates, n_paths, seed, n_basis)
    o, opv = O.price_american(contract.to_json(), dates, omodel, O.cfg(n_paths, seed, 50, model.rate), n_basis,
                              want_pv=True)
    assert g.n_paths == n_paths
    scale = np.abs(opv) + 1.0
    bad = np.abs(gpv - opv) > 1e-4 * scale
    print(f"gpu {g.price:.6f} +- {g.std_error:.6f}  oracle {o['price']:.6f} +- {o['std_error']:.6f}  "
          f"flipped {bad.sum()} of {n_paths}")
    assert bad.mean() <= flip_budget
    # FP32 payoff rounding (1e-5 relative) bounds the gap when no decision flips
    assert abs(g.price - o["price"]) <= 0.5 * o["std_error"] + 1e-5 * (abs(o["price"]) + 1.0)
    assert abs(g.std_error - o["std_error"]) <= 0.05 * o["std_error"] + 1e-6
    return g, o


def test_lsm_put_matches_oracle_same_stream(eng):
    model = K.GBM(36.0, 0.2, 0.06, asset="S")
    compare_with_oracle(eng, put_payoff("S", 40.0), model, O.gbm("S", 36.0, 0.06, 0.2), dates_uniform(1.0, 50),
                        65536, 42)


@pytest.mark.parametrize("n_basis", [2, 4, 6])
def test_lsm_basis_sizes_match_oracle(eng, n_basis):
    model = K.GBM(100.0, 0.3, 0.05, asset="S")
    # n_basis 6 raises S^10 ~ 1e20 into the normal equations; the reference
    # solves them as they are, and so does the GPU (FP64 both ways).
    compare_with_oracle(eng, put_payoff("S", 105.0), model, O.gbm("S", 100.0, 0.05, 0.3), dates_uniform(1.0, 40),
                        32768, 7, n_basis=n_basis, flip_budget=0.03)


def test_lsm_heston_and_basket_match_oracle(eng):
    hm = K.heston(100.0, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="S")
    compare_with_oracle(eng, put_payoff("S", 100.0), hm, O.heston("S", 100.0, 0.05, 0.04, 2.0, 0.04, 0.3, -0.5),
                        dates_uniform(1.0, 12), 16384, 3)
    corr = [[1.0, 0.3, 0.3], [0.3, 1.0, 0.3], [0.3, 0.3, 1.0]]
    fac = [K.GbmFactor(n, 100.0, 0.05, s) for n, s in zip("ABC", (0.2, 0.25, 0.3))]
    bm = K.correlated_gbm(fac, corr, 0.05)
    basket = (K.spot("A") + K.spot("B") + K.spot("C")) / 3.0
    payoff = K.scale((K.const_(100.0) - basket).max(0.0), K.one("USD"))
    compare_with_oracle(eng, payoff, bm, O.basket(list("ABC"), [100.0] * 3, [0.05] * 3, [0.2, 0.25, 0.3], corr),
                        dates_uniform(1.0, 10), 16384, 5)


def test_lsm_single_date_is_the_european_price(eng):
    # One exercise date: LSM is the European price on the grid [0, T], and
    # both paths draw the same normal for that step.
    model = K.GBM(100.0, 0.25, 0.06, asset="S")
    put = put_payoff("S", 110.0)
    g, gpv, _ = gpu_american(eng, put, model, [1.0], 100000, 11)
    eu = K.european_put("S", 110.0, 1.0).price(model, n_paths=100000, seed=11, grid=[0.0, 1.0])
    assert g.price == pytest.approx(eu.price, rel=1e-12, abs=1e-12)
    assert g.std_error == pytest.approx(eu.std_error, rel=1e-9)


def test_config5_american_put_literature(eng):
    # S0=36, K=40, sigma=0.2, r=0.06, T=1, 50 exercise dates, 2^20 paths.
    # Longstaff & Schwartz (2001), Table 1: simulated 4.472 (s.e. 0.010);
    # a 500-step CRR tree gives the continuous-exercise value ~4.487.
    model = K.GBM(36.0, 0.2, 0.06, asset="S")
    dates = dates_uniform(1.0, 50)
    g, _, coef = gpu_american(eng, put_payoff("S", 40.0), model, dates, 1 << 20, 42)
    o, _ = O.price_american(put_payoff("S", 40.0).to_json(), dates, O.gbm("S", 36.0, 0.06, 0.2),
                            O.cfg(1 << 20, 42, 50, 0.06))
    print(f"gpu {g.price:.6f} +- {g.std_error:.6f}  oracle {o['price']:.6f}")
    assert abs(g.price - o["price"]) <= 3 * o["std_error"]
    assert abs(g.price - 4.472) <= 3 * 0.010
    crr = O.crr_american(36.0, 40.0, 1.0, 0.06, 0.2, 500, False)
    assert g.price <= crr + 3 * g.std_error
    assert np.all(coef[:, 3] == 1.0)  # every date had ITM paths and a fit
    assert g.price >= 4.0  # at least the immediate exercise value


def test_american_call_equals_european_call(eng):
    # tests/j17_lsm.rs:110-131
    model = K.GBM(100.0, 0.2, 0.05, asset="AAPL")
    lsm = call_payoff("AAPL", 95.0).price_american(model, dates_uniform(1.0, 20), n_paths=60000, seed=7)
    bs = O.bs_call(100.0, 95.0, 1.0, 0.05, 0.2)
    assert abs(lsm.price - bs) / bs < 0.02


def test_american_put_exceeds_european_put(eng):
    # tests/j17_lsm.rs:134-152
    model = K.GBM(100.0, 0.25, 0.06, asset="AAPL")
    lsm = put_payoff("AAPL", 110.0).price_american(model, dates_uniform(1.0, 20), n_paths=60000, seed=7)
    bs = O.bs_put(100.0, 110.0, 1.0, 0.06, 0.25)
    assert lsm.price > bs
    assert lsm.price - bs > 0.05


@pytest.mark.parametrize("is_call,r,sigma", [(True, 0.05, 0.2), (False, 0.06, 0.25)])
def test_american_matches_crr(eng, is_call, r, sigma):
    # tests/j17_lsm.rs:156-198
    model = K.GBM(100.0, sigma, r, asset="AAPL")
    payoff = call_payoff("AAPL", 100.0) if is_call else put_payoff("AAPL", 100.0)
    lsm = payoff.price_american(model, dates_uniform(1.0, 50), n_paths=60000, seed=7)
    crr = O.crr_american(100.0, 100.0, 1.0, r, sigma, 500, is_call)
    assert abs(lsm.price - crr) / crr < 0.01
    if not is_call:
        assert lsm.price >= 0.0


def test_basis_sensitivity_converges(eng):
    # tests/j17_lsm.rs:203-232
    model = K.GBM(100.0, 0.3, 0.05, asset="AAPL")
    payoff = put_payoff("AAPL", 105.0)
    dates = dates_uniform(1.0, 40)
    p2 = payoff.price_american(model, dates, n_paths=60000, seed=7, n_basis=2).price
    p3 = payoff.price_american(model, dates, n_paths=60000, seed=7, n_basis=3).price
    crr = O.crr_american(100.0, 105.0, 1.0, 0.05, 0.3, 500, False)
    assert abs(p2 - p3) / p3 < 0.02
    assert abs(p3 - crr) / crr < 0.02


def test_lsm_argument_errors(eng):
    model = K.GBM(100.0, 0.2, 0.05, asset="S")
    put = put_payoff("S", 100.0)
    for dates, nb in (([], 3), ([0.5, 0.5], 3), ([0.0, 1.0], 3), ([0.5, 0.25], 3), ([1.0], 1)):
        with pytest.raises(N.KontractError) as ei:
            put.price_american(model, dates, n_paths=1000, n_basis=nb)
        assert ei.value.kind == "MalformedContract"
    with pytest.raises(N.KontractError) as ei:
        put.price_american(model, [1.0], n_paths=1000, n_basis=7)
    assert ei.value.kind == "Unsupported"
    # a payoff whose flows are not paid at the exercise date
    deferred = K.when(K.at(2.0), put)
    with pytest.raises(N.KontractError) as ei:
        deferred.price_american(model, [1.0], n_paths=1000)
    assert ei.value.kind == "Unsupported"
    # unknown asset in the payoff is raised from the paths
    with pytest.raises(N.KontractError) as ei:
        put_payoff("X", 100.0).price_american(model, [0.5, 1.0], n_paths=1000)
    assert ei.value.kind == "UnknownAsset"


def test_lsm_zero_paths_and_tiny_counts(eng):
    model = K.GBM(36.0, 0.2, 0.06, asset="S")
    put = put_payoff("S", 40.0)
    r0 = put.price_american(model, dates_uniform(1.0, 5), n_paths=0)
    assert r0.price == 0.0 and r0.n_paths == 0
    # fewer ITM paths than basis functions: no regression, carry realised values
    for n in (1, 2, 3, 255, 257):
        compare_with_oracle(eng, put, model, O.gbm("S", 36.0, 0.06, 0.2), dates_uniform(1.0, 5), n, 9,
                            flip_budget=0.0)


@pytest.mark.parametrize("kind", ["put", "call"])
def test_lsm_straight_line_payoff_equals_interpreter(eng, kind):
    """A vanilla put / call takes the paths kernel's straight-line payoff; the
    same payoff plus 0.0 (one more op) is interpreted.  The two must store
    the same payoffs, fit the same regressions and give every path the same
    value, to the last bit."""
    k = 40.0
    base = (K.const_(k) - K.spot("S")) if kind == "put" else (K.spot("S") - k)
    vanilla = K.scale(base.max(0.0), K.one("USD"))
    interpreted = K.scale(base.max(0.0) + 0.0, K.one("USD"))
    model = K.GBM(36.0, 0.2, 0.06, asset="S")
    dates = dates_uniform(1.0, 50)
    gv, pv_v, cv = gpu_american(eng, vanilla, model, dates, 100_000, 42)
    gi, pv_i, ci = gpu_american(eng, interpreted, model, dates, 100_000, 42)
    assert np.array_equal(pv_v, pv_i)
    assert np.array_equal(cv, ci)
    assert gv.price == gi.price and gv.std_error == gi.std_error
