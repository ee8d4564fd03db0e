"""Greek surface on the GPU (kt_greek_surface, src/surface.rs:104-138): every
cell of the one-launch surface equals greeks_gbm at that point, cells agree
with the oracle's surface on the same stream, and the reference's own checks
(tests/surface.rs: shapes, delta 2 % / vega 3 % of Black-Scholes, monotone
delta, empty grid rejected) hold."""
import math

import numpy as np
import pytest
from scipy.stats import norm

import kontracts_amd as K
import oracle_lib as O

pytestmark = pytest.mark.gpu

CALL = K.european_call("AAPL", 100.0, 1.0)


def greeks_at(contract, asset, s, v, n, seed, spy, rate, bumps, **kw):
    return contract.greeks(K.GBM(s, v, rate, asset=asset), n_paths=n, seed=seed, steps_per_year=spy, bumps=bumps, **kw)


def assert_cells_equal_greeks(surf, contract, asset, n, seed, spy, rate, bumps, **kw):
    for i, s in enumerate(surf.spots):
        for j, v in enumerate(surf.vols):
            g = greeks_at(contract, asset, s, v, n, seed, spy, rate, bumps, **kw)
            got = (surf.price[i, j], surf.delta[i, j], surf.gamma[i, j], surf.vega[i, j], surf.rho[i, j])
            assert got == (g.price, g.delta, g.gamma, g.vega, g.rho), (i, j)
            assert surf.std_errors["delta"][i, j] == g.std_errors[1] or math.isnan(g.std_errors[1])


def test_surface_matches_black_scholes_and_shapes():
    spots, vols = [90.0, 100.0, 110.0], [0.15, 0.20, 0.25]
    surf = K.greek_surface(CALL, "AAPL", spots, vols, n_paths=200_000, seed=2024, steps_per_year=1, rate=0.05,
                           bumps=(1.0, 1e-2, 1e-4))
    for m in (surf.price, surf.delta, surf.gamma, surf.vega, surf.rho):
        assert m.shape == (3, 3)
    for i, s in enumerate(spots):
        for j, v in enumerate(vols):
            d1 = (math.log(s / 100.0) + (0.05 + 0.5 * v * v)) / v
            d_bs, v_bs = norm.cdf(d1), s * norm.pdf(d1)
            assert abs(surf.delta[i, j] - d_bs) / d_bs < 0.02
            assert abs(surf.vega[i, j] - v_bs) / v_bs < 0.03
    assert_cells_equal_greeks(surf, CALL, "AAPL", 200_000, 2024, 1, 0.05, (1.0, 1e-2, 1e-4))


def test_surface_delta_monotone_and_exports():
    surf = K.greek_surface(CALL, "AAPL", [80.0, 100.0, 120.0], [0.20], n_paths=200_000, seed=2024, steps_per_year=1)
    assert surf.delta[0, 0] < surf.delta[1, 0] < surf.delta[2, 0]
    s2 = K.greek_surface(CALL, "AAPL", [95.0, 105.0], [0.18, 0.22], n_paths=50_000, seed=2024, steps_per_year=1)
    csv = s2.to_csv(s2.delta)
    assert csv.startswith("spot\\vol,0.18,0.22") and len(csv.splitlines()) == 3
    assert s2.to_pgm(s2.delta).startswith("P2\n2 2\n255")


def test_surface_knock_out_cells_and_oracle():
    ko = K.up_and_out_call("X", 100.0, 140.0, 1.0)
    spots, vols = [95.0, 105.0, 120.0], [0.18, 0.3]
    n = 20_000
    surf = K.greek_surface(ko, "X", spots, vols, n_paths=n, seed=11, steps_per_year=252, rate=0.05,
                           bumps=(1.0, 0.01, 1e-4))
    assert_cells_equal_greeks(surf, ko, "X", n, 11, 252, 0.05, (1.0, 0.01, 1e-4))
    cells = O.greek_surface(ko.to_json(), "X", spots, vols, O.cfg(n, 11, 252, 0.05), 1.0, 0.01, 1e-4)
    for c, o in enumerate(cells):
        i, j = divmod(c, len(vols))
        # same Philox stream: FP32 paths and barrier flips are the only differences
        assert abs(surf.price[i, j] - o["price"]) < 0.5 * surf.std_errors["price"][i, j]
        assert abs(surf.delta[i, j] - o["delta"]) < 0.5 * surf.std_errors["delta"][i, j] + 1e-3


def test_surface_variance_reduction_and_large_programs_go_by_cell():
    surf = K.greek_surface(CALL, "AAPL", [95.0, 105.0], [0.2, 0.3], n_paths=20_000, seed=5, steps_per_year=12,
                           antithetic=True)
    assert_cells_equal_greeks(surf, CALL, "AAPL", 20_000, 5, 12, 0.05, None, antithetic=True)
    many = K.european_call("X", 90.0, 1.0)
    for k in range(95, 160, 5):  # 31 registers: beyond the scenario kernel
        many = many + K.when(K.spot("X") >= float(k), K.one("USD")) + K.european_call("X", float(k), 1.0)
    surf = K.greek_surface(many, "X", [100.0], [0.2, 0.25], n_paths=4000, seed=6, steps_per_year=12)
    assert_cells_equal_greeks(surf, many, "X", 4000, 6, 12, 0.05, None)


def test_surface_errors():
    with pytest.raises(K.KontractError) as e:
        K.greek_surface(CALL, "AAPL", [], [0.2], n_paths=1000)
    assert e.value.kind == "InconsistentPath"
    with pytest.raises(K.KontractError) as e:
        K.greek_surface(K.european_call("Y", 100.0, 1.0), "AAPL", [100.0], [0.2], n_paths=1000)
    assert e.value.kind == "UnknownAsset"
    z = K.greek_surface(CALL, "AAPL", [100.0, 110.0], [0.2], n_paths=0)
    assert np.all(z.price == 0.0)


def test_surface_sixteen_register_programs():
    # 13 registers: the surface kernel's 16-register file (the contracts above use the 8-register one)
    S = K.spot("X")
    e = (K.running_max(S) - 95.0).clip(0.0) + (105.0 - K.running_min(S)).clip(0.0)
    c = K.up_and_out_call("X", 100.0, 200.0, 1.0) + K.until(S >= 230.0, K.scale(e, K.one("USD")) @ K.at(1.0))
    surf = K.greek_surface(c, "X", [95.0, 110.0], [0.2, 0.3], n_paths=8192, seed=9, steps_per_year=52, rate=0.03,
                           bumps=(1.0, 0.01, 1e-4))
    assert_cells_equal_greeks(surf, c, "X", 8192, 9, 52, 0.03, (1.0, 0.01, 1e-4))
