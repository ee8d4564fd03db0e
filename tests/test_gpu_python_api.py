"""The drop-in claim at the Python surface: the checks the reference's own
Python tests make of its package (python/tests/test_bindings.py:25-70 and
python/tests/test_temporal_basket.py:26-91), restated here against
``import kontracts_amd as kontract`` and run on the GPU with the reference's
path counts, seeds, grids and tolerances.  The Black-Scholes comparator is
the oracle's (exact erfc).
"""
import math

import pytest

import kontracts_amd as kontract
from kontracts_amd import GBM, S, USD, at, one
import oracle_lib as O

pytestmark = pytest.mark.gpu


def atm_call(asset="AAPL", strike=100.0, t=1.0):
    # the fluent form: observable * contract is a scale, @ a condition a when
    return ((S(asset) - strike).clip(0.0) * one(USD)) @ at(t)


def gbm(asset="AAPL", s0=100.0, sigma=0.20, r=0.05):
    return GBM(s0=s0, sigma=sigma, r=r, asset=asset)


# python/tests/test_bindings.py -------------------------------------------------
def test_package_surface():
    assert isinstance(kontract.__version__, str) and kontract.__version__
    assert kontract.USD == "USD"


def test_fluent_contract_and_json_round_trip():
    c = atm_call()
    assert isinstance(c, kontract.Contract)
    back = kontract.Contract.from_json(c.to_json())
    assert back.to_json() == c.to_json()
    assert back == c


def test_european_call_within_one_percent_of_black_scholes():
    res = atm_call().price(gbm(), n_paths=400_000, seed=2024, steps_per_year=1)
    bs = O.bs_call(100.0, 100.0, 1.0, 0.05, 0.20)
    assert abs(res.price - bs) / bs < 0.01
    assert res.std_error > 0.0
    assert res.ci95_low < res.price < res.ci95_high
    assert res.n_paths == 400_000


def test_greeks_signs_for_an_atm_call():
    g = atm_call().greeks(gbm(), n_paths=300_000, seed=7, steps_per_year=1)
    assert 0.0 < g.delta < 1.0
    assert g.vega > 0.0
    assert g.gamma > 0.0


def test_until_knock_out_below_vanilla():
    call = atm_call()
    knocked = call.until(S("AAPL") >= 150.0)
    m = gbm()
    ko = knocked.price(m, n_paths=100_000, seed=1, steps_per_year=100).price
    vanilla = call.price(m, n_paths=100_000, seed=1, steps_per_year=100).price
    assert 0.0 < ko < vanilla


def test_book_by_addition():
    book = one(USD) + one(USD)
    res = book.price(gbm(r=0.0), n_paths=1000, seed=1, steps_per_year=1)
    assert abs(res.price - 2.0) < 1e-9


# python/tests/test_temporal_basket.py ------------------------------------------
def test_asian_below_vanilla_and_positive():
    m = gbm("X")
    asian = ((kontract.average(S("X")) - 100.0).clip(0.0) * one(USD)) @ at(1.0)
    vanilla = kontract.european_call("X", 100.0, 1.0, USD)
    a = asian.price(m, n_paths=120_000, seed=7, steps_per_year=50).price
    v = vanilla.price(m, n_paths=120_000, seed=7, steps_per_year=50).price
    assert 0.0 < a < v


def test_average_method_equals_function():
    via_fn = ((kontract.average(S("X")) - 100.0).clip(0.0) * one(USD)) @ at(1.0)
    via_method = ((S("X").average() - 100.0).clip(0.0) * one(USD)) @ at(1.0)
    assert via_fn.to_json() == via_method.to_json()
    m = gbm("X")
    assert via_fn.price(m, n_paths=10_000, seed=3).price == via_method.price(m, n_paths=10_000, seed=3).price


def test_lookback_above_vanilla():
    m = gbm("X")
    lookback = ((kontract.running_max(S("X")) - 100.0).clip(0.0) * one(USD)) @ at(1.0)
    vanilla = kontract.european_call("X", 100.0, 1.0, USD)
    lb = lookback.price(m, n_paths=120_000, seed=9, steps_per_year=50).price
    v = vanilla.price(m, n_paths=120_000, seed=9, steps_per_year=50).price
    assert lb > v


def test_call_on_running_min_below_call_on_running_max():
    m = gbm("X")
    on_min = ((kontract.running_min(S("X")) - 80.0).clip(0.0) * one(USD)) @ at(1.0)
    on_max = ((kontract.running_max(S("X")) - 80.0).clip(0.0) * one(USD)) @ at(1.0)
    assert on_min.price(m, n_paths=80_000, seed=4, steps_per_year=50).price < \
        on_max.price(m, n_paths=80_000, seed=4, steps_per_year=50).price


def test_basket_within_four_percent_of_black_scholes_sigma_basket():
    rho = 0.5
    fs = [kontract.GbmFactor(nm, 100.0, 0.05, 0.20) for nm in ("S1", "S2", "S3")]
    model = kontract.correlated_gbm(fs, [[1.0, rho, rho], [rho, 1.0, rho], [rho, rho, 1.0]], r=0.05)
    basket = (S("S1") + S("S2") + S("S3")) / 3.0
    c = ((basket - 100.0).clip(0.0) * one(USD)) @ at(1.0)
    price = c.price(model, n_paths=200_000, seed=7, steps_per_year=50).price
    ana = O.bs_call(100.0, 100.0, 1.0, 0.05, 0.20 * math.sqrt((1.0 + 2.0 * rho) / 3.0))
    assert abs(price - ana) / ana < 0.04


def test_correlated_gbm_rejects_a_matrix_of_the_wrong_shape():
    fs = [kontract.GbmFactor("S1", 100.0, 0.05, 0.2), kontract.GbmFactor("S2", 100.0, 0.05, 0.2)]
    with pytest.raises((ValueError, RuntimeError)):
        kontract.correlated_gbm(fs, [[1.0, 0.5, 0.0]], r=0.05)
