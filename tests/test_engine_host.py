"""CPU tests of the engine's host side: the C ABI library loads and exports
every symbol the header declares; the contract reader, plan and grid agree
with the oracle's restatement of src/compiler.rs; serde JSON shapes."""
import json
import os
import re

import pytest

import kontracts_amd as K
from kontracts_amd import _native as N
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "kontracts.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(kt_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    declared = header_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
    assert declared == N.exported_symbols()


def test_engine_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(K.DeviceError):
        K.european_call("X", 100, 1.0).price(K.GBM(100, 0.2, 0.05, asset="X"), n_paths=10)


CONTRACTS = [
    K.one("USD") @ K.at(1.0),
    K.european_call("AAPL", 150, 1.0),
    K.up_and_out_call("AAPL", 100, 200, 1.0),
    K.zero(),
    K.anytime(K.spot("AAPL") <= 80.0, K.one("USD") @ K.at(2.0)),
    (K.one("USD") @ K.at(0.5)) + K.until(K.spot("AAPL") >= 150.0, K.european_call("AAPL", 120, 1.0)),
    K.scale(K.spot("MSFT") + K.spot("AAPL"), K.one("USD")) @ K.at(1.0),
    K.scale((K.average(K.spot("S")) - 100.0).clip(0.0), K.one("USD")) @ K.at(0.75),
    K.scale((K.running_max(K.spot("S")) - 90.0).clip(0.0), K.one("USD")) @ K.at(3.0),
    K.cash_or_nothing_call("X", 100, 10, 1.3),
    (K.one("USD") @ K.at(1.0)) + (K.one("EUR") @ K.at(1.0)) + (K.one("USD") @ K.at(1.0 + 1e-13)),
    K.one("USD") @ (K.at(0.3) & ~K.at(0.7)),
]


@pytest.mark.parametrize("i", range(len(CONTRACTS)))
@pytest.mark.parametrize("spy,force", [(0, False), (12, False), (252, False), (365, True), (7, True)])
def test_plan_matches_oracle(i, spy, force):
    js = CONTRACTS[i].to_json()
    ref = O.compile_plan(js, spy, force)
    grid, dates, hz, fine, assets = K.dsl._plan_raw(js, spy, force)
    assert grid == ref["grid"]  # bit-exact f64 grid
    assert dates == ref["fixed_dates"] and hz == ref["horizon"]
    assert fine == ref["needs_fine_grid"] and assets == ref["assets"]


def test_plan_errors_match_reference():
    with pytest.raises(K.KontractError) as e:
        K.plan(K.one("USD") @ K.at(-1.0))
    assert e.value.kind == "MalformedContract"
    with pytest.raises(K.KontractError) as e:
        K.Contract.from_json("{ not valid json")
    assert e.value.kind == "Serde"
    with pytest.raises(K.KontractError) as e:
        K.Contract.from_json('"Unknown"')
    assert e.value.kind == "Serde"


def test_serde_json_shapes():
    c = K.european_call("X", 100.0, 1.0)
    assert c.to_json() == ('{"When":[{"At":1.0},{"Scale":[{"Max":[{"Sub":[{"Spot":"X"},{"Const":100.0}]},'
                           '{"Const":0.0}]},{"One":"USD"}]}]}')
    assert K.zero().to_json() == '"Zero"'
    assert K.average(K.spot("S")).node == {"Average": {"obs": {"Spot": "S"}, "from_year": None, "to_year": None}}
    assert K.Contract({"Scale": [{"Const": 1e-05}, "Zero"]}).to_json() == '{"Scale":[{"Const":1e-5},"Zero"]}'


def test_json_round_trip_through_engine_reader():
    cases = CONTRACTS + [
        -K.one("EUR"),
        K.one("USD").or_(K.one("EUR")),
        K.scale(-K.spot("AAPL"), K.one("USD")),
        K.when(~(K.spot("AAPL") >= 200.0) | (K.spot("MSFT") < 100.0), K.one("USD")),
        K.scale((((K.spot("AAPL") + 1.0) * 2.0 - K.spot("MSFT")) / 3.0), K.one("USD")),
        K.scale(K.average_over(K.spot("X"), 0.25, 0.75).min(K.running_min(K.spot("X"))), K.one("USD")),
    ]
    for c in cases:
        back = K.Contract.from_json(c.to_json())
        assert back == c
        K.plan(back)  # the engine's reader accepts it


def test_reference_serialised_json_accepted():
    # serde output as the reference writes it (whitespace, integral floats)
    js = ('{"Until": [{"Ge": [{"Spot": "AAPL"}, {"Const": 200}]}, {"When": [{"At": 1}, {"Scale": '
          '[{"Max": [{"Sub": [{"Spot": "AAPL"}, {"Const": 150.0}]}, {"Const": 0.0}]}, {"One": "USD"}]}]}]}')
    c = K.Contract.from_json(js)
    assert c == K.up_and_out_call("AAPL", 150, 200, 1.0)
    # Average with a missing optional field and an unknown extra field
    js2 = '{"When":[{"At":1.0},{"Scale":[{"Average":{"obs":{"Spot":"X"},"from_year":0.5,"extra":[1,{"a":2}]}},"Zero"]}]}'
    p = K.dsl._plan_raw(js2, 4, False)
    assert p[3] is True


def test_program_lowering_shapes():
    gbm = K.GBM(100, 0.2, 0.05, asset="X")
    grid = [i / 252 for i in range(253)]
    txt = K.describe_program(K.european_call("X", 100, 1.0), gbm, grid=grid)
    assert "steps 0..251: -" in txt and "emit" in txt
    txt = K.describe_program(K.up_and_out_call("X", 100, 200, 1.0), gbm, steps_per_year=252)
    assert "steps 0..251:" in txt and txt.count("until") == 2
    # `or` acquired at a deterministic step is always reached: error at compile time
    with pytest.raises(K.KontractError) as e:
        K.describe_program(K.one("USD").or_(K.one("USD")), gbm)
    assert e.value.kind == "Unsupported"
    with pytest.raises(K.KontractError) as e:
        K.describe_program(K.european_call("X", 100, 1.0), K.GBM(100, -0.2, 0.05, asset="X"))
    assert e.value.kind == "MalformedContract"
    with pytest.raises(K.KontractError) as e:
        K.describe_program(K.european_call("X", 100, 1.0), gbm, grid=[0.0, 0.5, 0.4])
    assert e.value.kind == "InconsistentPath"
    with pytest.raises(K.KontractError) as e:
        K.describe_program(K.european_call("X", 100, 1.0), K.heston(100, 0.04, 2, 0.04, 0.3, -1.5, 0.05, "X"))
    assert e.value.kind == "MalformedContract"


def test_average_at_plan_and_lowering():
    """AverageAt fixings join the grid; the accumulator is fed at the fixing
    points only and read with the static count of past fixings."""
    fix = [0.25, 0.5, 0.75, 1.0]
    c = K.scale((K.average_at(K.spot("X"), fix) - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0)
    p = K.plan(c)
    for d in fix:
        assert d in p.fixed_dates
    assert not p.needs_fine_grid
    text = K.describe_program(c, K.GBM(100.0, 0.2, 0.05, asset="X"), grid=[i / 8 for i in range(9)])
    fed = [ln for ln in text.splitlines() if "avg(" in ln]
    # fed at grid points 2, 4, 6, 8 (t = 0.25, 0.5, 0.75, 1.0) only
    assert any(ln.startswith("steps 2..2") for ln in fed) and not any(ln.startswith("steps 1..1") for ln in fed)
    assert "muli(5,5,0,0.25)" in text
    # round trip through the engine's reader
    assert K.Contract.from_json(c.to_json()) == c


def test_fixing_groups_classified():
    """A group of four points whose points either have no ops or fix one
    running sum (spot, then its Kahan update) is the fast kind "avg at
    points"; two sums in one group, or a fixing beside a knock-out monitor,
    stay interpreted; the partial last group always is."""
    gbm = K.GBM(100.0, 0.2, 0.05, asset="X")
    grid = [i / 16 for i in range(17)]

    def groups(c):
        line = K.describe_program(c, gbm, grid=grid).splitlines()[1]
        return {k.strip(): int(v) for k, v in (part.rsplit(" ", 1) for part in line[len("groups: "):].split(","))}

    avg_at = lambda dates: K.average_at(K.spot("X"), dates)
    # fixings at points 1 and 2 (group 0) and 6 (group 1); the payoff at 16
    g = groups(K.scale((avg_at([1 / 16, 2 / 16, 6 / 16]) - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0))
    assert g == {"generic": 1, "none": 2, "knock-out": 0, "avg": 0, "avg at points": 2, "rmax/rmin": 0}
    g = groups(K.scale((avg_at([1 / 16]) - avg_at([2 / 16])).clip(0.0), K.one("USD")) @ K.at(1.0))
    assert g["avg at points"] == 0 and g["generic"] == 2
    g = groups(K.until(K.spot("X") >= 150.0, K.scale(avg_at([5 / 16]), K.one("USD")) @ K.at(1.0)))
    assert g["avg at points"] == 0 and g["knock-out"] == 3 and g["generic"] == 2


def test_portfolio_book_generator():
    """The batched-portfolio book: deterministic, 250 of each kind, daily grids
    of 63..756 steps, 2^16..2^20 paths, single-asset models split GBM/Heston,
    and every program small enough for the persistent kernel."""
    from kontracts_amd import book as B
    jobs = B.book(1000)
    again = B.book(1000)
    assert [j[0].to_json() for j in jobs] == [j[0].to_json() for j in again]
    assert [j[0].to_json() for j in B.book(40)] == [j[0].to_json() for j in jobs[:40]]
    steps = [len(j[5]) - 1 for j in jobs]
    assert 63 <= min(steps) and max(steps) <= 756
    assert {j[2] for j in jobs} <= {1 << e for e in range(16, 21)}
    kinds = [B.KINDS[i % 4] for i in range(1000)]
    assert all(kinds.count(k) == 250 for k in B.KINDS)
    n_heston = sum(1 for j in jobs if isinstance(j[1], K.Model) and j[1].label == "Heston")
    assert 300 < n_heston < 450
    assert 1.4e11 < B.path_steps(jobs) < 1.9e11
    for j in jobs[:8]:
        first = K.describe_program(j[0], j[1], grid=j[5]).splitlines()[0]
        assert int(first.split()[1].rstrip(",")) <= 16


def test_job_array_reused_only_while_the_book_is_unchanged():
    """price_portfolio marshals a book once and reuses the kt_job array while
    the same job tuples come back with unchanged models; a changed model, a
    different job or a grid the engine cannot read in place (a list, copied
    on marshalling) gives a fresh array with the current values."""
    from kontracts_amd import book as B
    from kontracts_amd import dsl as D
    jobs = B.book(12)
    a1, _ = D._jobs_array(jobs)
    a2, _ = D._jobs_array(jobs)
    assert a1 is a2
    g = next(i for i, j in enumerate(jobs) if isinstance(j[1], D.GBM))
    h = next(i for i, j in enumerate(jobs) if isinstance(j[1], D.Model) and j[1].kind == N.KT_MODEL_HESTON)
    jobs[g][1].sigma = 0.31
    a3, _ = D._jobs_array(jobs)
    assert a3 is not a2 and a3[g].model.sigma[0] == 0.31
    jobs[h][1].params["kappa"] = 2.5
    a4, _ = D._jobs_array(jobs)
    assert a4 is not a3 and a4[h].model.kappa == 2.5
    other = list(jobs)
    other[0] = other[1]
    a5, _ = D._jobs_array(other)
    assert a5 is not a4 and a5[0].cfg.seed == jobs[1][3]
    listed = [j[:5] + (list(j[5]),) for j in jobs]
    b1, _ = D._jobs_array(listed)
    b2, _ = D._jobs_array(listed)
    assert b1 is not b2
    listed[0][5][1] = 0.5
    b3, _ = D._jobs_array(listed)
    assert b3[0].cfg.grid[1] == 0.5
