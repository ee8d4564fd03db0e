"""CPU tests of the Greek surface's host side (src/surface.rs): the CSV / PGM
exports and the Rust ``Display`` number format they use, and the oracle's
surface (a loop of greeks_gbm, src/surface.rs:104-138) against
Black-Scholes with the reference's own tolerances (tests/surface.rs)."""
import math

import numpy as np
import pytest
from scipy.stats import norm

import oracle_lib as O
import kontracts_amd as K
from kontracts_amd.surface import GreekSurface, rust_display


@pytest.mark.parametrize("x,text", [
    (0.18, "0.18"), (95.0, "95"), (100.0, "100"), (-3.0, "-3"), (-0.0, "-0"), (0.1 + 0.2, "0.30000000000000004"),
    (1e-7, "0.0000001"), (1.5e-10, "0.00000000015"), (1e21, "1000000000000000000000"), (2.5e16, "25000000000000000"),
    (123456789.125, "123456789.125"), (float("nan"), "NaN"), (float("inf"), "inf"), (float("-inf"), "-inf"),
])
def test_rust_display(x, text):
    assert rust_display(x) == text


def surface_2x2():
    return GreekSurface([95.0, 105.0], [0.18, 0.22], [[5.1, 6.2], [9.3, 10.4]], [[0.4, 0.45], [0.62, 0.6]],
                        [[0.02, 0.018], [0.019, 0.017]], [[37.0, 38.5], [36.1, 37.9]])


def test_csv_export_shape_and_text():
    s = surface_2x2()
    csv = s.to_csv(s.delta)
    assert csv.startswith("spot\\vol,0.18,0.22")  # tests/surface.rs csv check
    assert csv.count("\n") == 3 and len(csv.splitlines()) == 3
    assert csv == "spot\\vol,0.18,0.22\n95,0.4,0.45\n105,0.62,0.6\n"


def test_pgm_export():
    s = surface_2x2()
    pgm = s.to_pgm(s.delta)
    assert pgm.startswith("P2\n2 2\n255")  # tests/surface.rs pgm check
    # min -> 0, max -> 255, others rounded half away from zero
    lo, hi = 0.4, 0.62
    want = [[0, round((0.45 - lo) / (hi - lo) * 255.0)], [255, round((0.6 - lo) / (hi - lo) * 255.0)]]
    assert pgm == "P2\n2 2\n255\n" + "".join("".join(f"{v} " for v in row) + "\n" for row in want)
    # a flat matrix spans 1: every cell is 0; a half rounds up, NaN draws 0
    flat = GreekSurface([1.0], [0.1, 0.2], *[[[2.0, 2.0]]] * 4)
    assert flat.to_pgm(flat.price) == "P2\n2 1\n255\n0 0 \n"
    half = GreekSurface([1.0], [0.1, 0.2, 0.3], *[[[0.0, 0.5 / 255.0, 1.0]]] * 4)
    assert half.to_pgm(half.price) == "P2\n3 1\n255\n0 1 255 \n"
    nan = GreekSurface([1.0], [0.1, 0.2, 0.3], *[[[0.0, float("nan"), 1.0]]] * 4)
    assert nan.to_pgm(nan.price) == "P2\n3 1\n255\n0 0 255 \n"


def bs_delta(s, k, r, sg, t):
    return norm.cdf((math.log(s / k) + (r + 0.5 * sg * sg) * t) / (sg * math.sqrt(t)))


def bs_vega(s, k, r, sg, t):
    d1 = (math.log(s / k) + (r + 0.5 * sg * sg) * t) / (sg * math.sqrt(t))
    return s * norm.pdf(d1) * math.sqrt(t)


def test_oracle_surface_matches_black_scholes():
    spots, vols = [90.0, 100.0, 110.0], [0.15, 0.20, 0.25]
    c = K.european_call("AAPL", 100.0, 1.0).to_json()
    cells = O.greek_surface(c, "AAPL", spots, vols, O.cfg(200_000, 2024, 1, 0.05), h_spot=1.0)
    assert len(cells) == 9
    for i, s in enumerate(spots):
        for j, v in enumerate(vols):
            g = cells[i * 3 + j]
            assert abs(g["delta"] - bs_delta(s, 100, 0.05, v, 1.0)) / bs_delta(s, 100, 0.05, v, 1.0) < 0.02
            assert abs(g["vega"] - bs_vega(s, 100, 0.05, v, 1.0)) / bs_vega(s, 100, 0.05, v, 1.0) < 0.03
    # each cell is greeks_gbm at that point
    one = O.greeks_gbm(c, "AAPL", 110.0, 0.2, O.cfg(200_000, 2024, 1, 0.05), h_spot=1.0)
    assert cells[7] == one


def test_oracle_surface_delta_monotone_and_empty_grid():
    c = K.european_call("AAPL", 100.0, 1.0).to_json()
    cells = O.greek_surface(c, "AAPL", [80.0, 100.0, 120.0], [0.2], O.cfg(200_000, 2024, 1, 0.05))
    assert cells[0]["delta"] < cells[1]["delta"] < cells[2]["delta"]
    with pytest.raises(O.OracleError) as e:
        O.greek_surface(c, "AAPL", [], [0.2], O.cfg(1000, 1, 1, 0.05))
    assert e.value.code == 4
