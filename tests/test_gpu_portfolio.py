"""Persistent portfolio launch (kt_price_portfolio) against per-contract
pricing.  The reference prices a book as a loop of price_gbm calls
(src/pricer.rs:97-115); every job of one persistent launch must give exactly
what pricing that job alone gives (same paths, same per-block statistics,
same reduction order), whatever the tile configuration, and errors surface as
the loop would raise them.
"""
import pytest

import kontracts_amd as K
from kontracts_amd import _native as N
from kontracts_amd import book as B

pytestmark = pytest.mark.gpu


def small_book(n=36, paths=(4096, 9000, 16384, 300)):
    jobs = []
    for i, job in enumerate(B.book(n)):
        con, model, _, seed, spy, grid = job
        jobs.append((con, model, paths[i % len(paths)], seed, spy, grid))
    return jobs


def edge_jobs():
    g = K.GBM(100.0, 0.25, 0.03, asset="X")
    h = K.heston(100.0, 0.04, 1.5, 0.05, 0.4, -0.6, 0.03, asset="X")
    f3 = [K.GbmFactor(n, 100.0, 0.03, s) for n, s in zip("ABC", (0.2, 0.3, 0.25))]
    b3 = K.correlated_gbm(f3, [[1.0, 0.2, 0.1], [0.2, 1.0, 0.3], [0.1, 0.3, 1.0]], 0.03)
    f7 = [K.GbmFactor(f"S{i}", 100.0, 0.03, 0.2) for i in range(7)]
    b7 = K.correlated_gbm(f7, [[1.0 if i == j else 0.25 for j in range(7)] for i in range(7)], 0.03)
    b7_pay = K.scale((sum((K.spot(f"S{i}") for i in range(1, 7)), K.spot("S0")) / 7.0 - 100.0).clip(0.0),
                     K.one("USD")) @ K.at(1.0)
    many = K.european_call("X", 90.0, 1.0)
    for k in range(95, 160, 5):  # a program that needs more than 16 registers
        many = many + K.when(K.spot("X") >= float(k), K.one("USD")) + K.european_call("X", float(k), 1.0)
    return [
        (K.european_call("X", 100.0, 1.0), g, 0, 1, 50),                        # no paths
        (K.european_call("X", 100.0, 1.0), g, 1, 2, 50),                        # one path
        (K.scale(K.spot("A") - K.spot("C"), K.one("USD")) @ K.at(0.5), b3, 5000, 3, 12),
        (b7_pay, b7, 3000, 4, 12),                                              # 7 assets: per-contract path
        (many, g, 2000, 5, 50),                                                 # > 16 registers: per-contract
        (K.up_and_out_call("X", 100.0, 130.0, 1.0), h, 7000, 6, 100),
        (K.down_and_out_call("X", 100.0, 80.0, 1.0), g, 6000, 7, 252),
        (K.anytime(K.spot("X") >= 115.0, K.one("USD")) @ K.at(0.1), g, 5000, 8, 50),
    ]


def price_each(jobs):
    out = []
    for job in jobs:
        con, model, n, seed, spy = job[:5]
        grid = job[5] if len(job) > 5 else None
        out.append(con.price(model, n_paths=n, seed=seed, steps_per_year=spy, grid=grid))
    return out


def assert_identical(got, want):
    for i, (a, b) in enumerate(zip(got, want)):
        assert a.n_paths == b.n_paths, i
        assert a.price == b.price, (i, a.price, b.price)
        assert a.std_error == b.std_error, (i, a.std_error, b.std_error)


def test_portfolio_equals_per_contract():
    jobs = small_book() + edge_jobs()
    assert_identical(K.price_portfolio(jobs), price_each(jobs))


def test_portfolio_tile_configs_and_graph_replay():
    eng = N.engine(0)
    jobs = small_book(24)
    want = price_each(jobs)
    try:
        for graph in (1, 0):
            eng.set_option("portfolio_graph", graph)
            for bpc, chunk in ((0, 1), (2, 4), (8, 1), (1, 64)):
                eng.set_option("portfolio_blocks_per_cu", bpc)
                eng.set_option("portfolio_chunk", chunk)
                assert_identical(K.price_portfolio(jobs), want)
                assert_identical(K.price_portfolio(jobs), want)  # replay of the captured graph
            eng.set_option("portfolio_blocks_per_cu", 0)
            eng.set_option("portfolio_chunk", 1)
            other = small_book(10, paths=(2048,))  # another shape: a new graph
            assert_identical(K.price_portfolio(other), price_each(other))
    finally:
        eng.set_option("portfolio_blocks_per_cu", 0)
        eng.set_option("portfolio_chunk", 1)
        eng.set_option("portfolio_graph", 0)
    for bad in (("portfolio_chunk", 0), ("portfolio_blocks_per_cu", -1), ("portfolio_graph", 2),
                ("portfolio_waves", 5), ("portfolio_adjoint_waves", 4), ("lsm_graph", 2),
                ("portfolio_unroll", 4), ("portfolio_lanes", 64)):
        with pytest.raises(N.KontractError):
            eng.set_option(*bad)


def test_portfolio_register_budget_tiles():
    """The register-budget variants of the persistent kernels run the same
    code: every job stays identical to per-contract pricing (and Greeks)."""
    eng = N.engine(0)
    jobs = small_book(24) + edge_jobs()
    want = price_each(jobs)
    gjobs = small_book(12)
    gwant = per_contract_greeks(gjobs)
    try:
        for pw in (6, 8):
            eng.set_option("portfolio_waves", pw)
            assert_identical(K.price_portfolio(jobs), want)
        eng.set_option("portfolio_waves", 0)
        # step unroll (8 points per iteration) and 128-lane workgroups with
        # two paths per lane: same statistics, combined in the same order
        for unroll, lanes in ((2, 256), (1, 128), (2, 128)):
            eng.set_option("portfolio_unroll", unroll)
            eng.set_option("portfolio_lanes", lanes)
            assert_identical(K.price_portfolio(jobs), want)
            eng.set_option("portfolio_waves", 8)
            assert_identical(K.price_portfolio(jobs), want)
            eng.set_option("portfolio_waves", 0)
        eng.set_option("portfolio_adjoint_waves", 5)
        assert_greeks_identical(K.price_portfolio(gjobs, greeks=True), gwant)
    finally:
        eng.set_option("portfolio_waves", 0)
        eng.set_option("portfolio_adjoint_waves", 0)
        eng.set_option("portfolio_unroll", 1)
        eng.set_option("portfolio_lanes", 256)


def test_autotuner_sets_a_measured_tile():
    from kontracts_amd import tune as T
    eng = N.engine(0)
    jobs = small_book(16)
    want = price_each(jobs)
    try:
        r = T.autotune_portfolio(jobs, candidates=((0, 1), (2, 4)), repeats=1)
        assert T.tile_name(r["best"]) in r["device_ms"]
        # two (grid, chunk) pairs, price budgets 6 and 8, three (unroll, lanes) variants
        assert len(r["device_ms"]) == 2 + 2 + 3
        assert all(v > 0 for v in r["device_ms"].values())
        assert_identical(K.price_portfolio(jobs), want)  # under the tuned tile
    finally:
        T.set_tile(eng, T.DEFAULT_TILE)


def test_portfolio_errors_like_a_loop():
    g = K.GBM(100.0, 0.2, 0.05, asset="X")
    ok = (K.european_call("X", 100.0, 1.0), g, 1000, 1, 50)
    unknown = (K.european_call("Y", 100.0, 1.0), g, 1000, 1, 50)
    bad_model = (K.european_call("X", 100.0, 1.0), K.heston(100, 0.04, 2, 0.04, 0.3, -1.5, 0.05, asset="X"),
                 1000, 1, 50)
    with pytest.raises(N.KontractError) as e:
        K.price_portfolio([ok, unknown, ok])
    assert e.value.kind == "UnknownAsset"
    with pytest.raises(N.KontractError) as e:
        K.price_portfolio([ok, bad_model, unknown])
    assert e.value.kind == "MalformedContract"
    assert K.price_portfolio([]) == []


# ---------------------------------------------------------------------------
# Pathwise Greeks of a portfolio (kt_price_portfolio_greeks)
# ---------------------------------------------------------------------------
def per_contract_greeks(jobs):
    """The per-contract equivalent: pathwise Greeks where the payoff admits
    them, a price otherwise."""
    want = []
    for job in jobs:
        con, model, n, seed, spy = job[:5]
        grid = job[5] if len(job) > 5 else None
        try:
            g = con.pathwise_greeks(model, n_paths=n, seed=seed, steps_per_year=spy, grid=grid)
        except K.KontractError as e:
            assert e.kind == "Unsupported"
            want.append((con.price(model, n_paths=n, seed=seed, steps_per_year=spy, grid=grid), None))
            continue
        want.append((g.price, g))
    return want


def assert_greeks_identical(got, want):
    for i, ((pr, g), (wp, wg)) in enumerate(zip(got, want)):
        if wg is None:
            assert g is None, i
            assert pr.price == wp.price and pr.std_error == wp.std_error, (i, pr.price, wp.price)
        else:
            assert g is not None, i
            assert pr.price == wg.price and pr.std_error == wg.price_std_error, (i, pr.price, wg.price)
            assert g.names == wg.names
            assert g.values == wg.values, (i, g.values, wg.values)
            assert g.std_errors == wg.std_errors, i


def test_portfolio_greeks_equal_per_contract():
    g = K.GBM(100.0, 0.25, 0.03, asset="X")
    f3 = [K.GbmFactor(n, 100.0, 0.03, s) for n, s in zip("ABC", (0.2, 0.3, 0.25))]
    b3 = K.correlated_gbm(f3, [[1.0, 0.2, 0.1], [0.2, 1.0, 0.3], [0.1, 0.3, 1.0]], 0.03)
    big = K.european_call("X", 90.0, 1.0)
    for k in range(95, 130, 5):  # a Lipschitz program beyond 16 registers: per-contract Greeks
        big = big + K.european_call("X", float(k), 1.0)
    jobs = small_book(32) + [
        (K.scale(K.spot("A") - K.spot("C"), K.one("USD")) @ K.at(0.5), b3, 5000, 3, 12),   # 3 assets: per contract
        (big, g, 3000, 9, 50),
        (K.european_call("X", 100.0, 1.0), g, 0, 1, 50),                                      # no paths
        (K.scale((K.running_max(K.spot("X")) - 95.0).clip(0.0), K.one("USD")) @ K.at(1.0), g, 4000, 5, 50),
    ]
    got = K.price_portfolio(jobs, greeks=True)
    want = per_contract_greeks(jobs)
    assert_greeks_identical(got, want)
    kinds = {B.KINDS[i % 4]: got[i][1] is not None for i in range(32)}
    assert kinds == {"european": True, "asian": True, "barrier": False, "basket": True}
    assert got[-2][0].n_paths == 0 and got[-2][1].values == [0.0, 0.0, 0.0]
    # a second call (same shapes) gives the same numbers
    assert_greeks_identical(K.price_portfolio(jobs, greeks=True), want)


def test_portfolio_greeks_errors_and_tiles():
    g = K.GBM(100, 0.2, 0.05, asset="X")
    ok = (K.european_call("X", 100.0, 1.0), g, 1000, 1, 50)
    with pytest.raises(N.KontractError) as e:
        K.price_portfolio([ok, (K.european_call("Y", 100.0, 1.0), g, 1000, 1, 50)], greeks=True)
    assert e.value.kind == "UnknownAsset"
    horizon = K.one("USD") @ K.at(1.0)
    bad_or = (horizon + K.when(K.spot("X") >= 101.0, K.one("USD").or_(K.one("USD"))), g, 1000, 1, 12)
    with pytest.raises(N.KontractError) as e:
        K.price_portfolio([ok, ok, bad_or], greeks=True)
    assert e.value.kind == "Unsupported" and "job 2" in str(e.value)
    eng = N.engine(0)
    jobs = small_book(12)
    want = per_contract_greeks(jobs)
    try:
        for bpc, chunk in ((1, 1), (2, 8)):
            eng.set_option("portfolio_blocks_per_cu", bpc)
            eng.set_option("portfolio_chunk", chunk)
            assert_greeks_identical(K.price_portfolio(jobs, greeks=True), want)
    finally:
        eng.set_option("portfolio_blocks_per_cu", 0)
        eng.set_option("portfolio_chunk", 1)


# ---------------------------------------------------------------------------
# The 1,000-contract book itself (1.63e11 path-steps)
# ---------------------------------------------------------------------------
FULL_TILES = ((4, 4), (0, 1), (1, 16))  # the autotuner's pick for this book, the default, a long chunk


def _set_tile(eng, bpc, chunk):
    eng.set_option("portfolio_blocks_per_cu", bpc)
    eng.set_option("portfolio_chunk", chunk)


def test_full_book_equals_per_contract_pricing():
    """Every job of the whole book, priced in one persistent launch under
    each tile, is bit-identical to pricing it alone (the reference's loop of
    price_gbm calls, src/pricer.rs:97-115).  The book has ~1.6e6 (job, block)
    items, so a chunk of 4 or 16 items per queue fetch takes effect on the
    1,024 resident workgroups (the chunk is capped at items per workgroup)."""
    eng = N.engine(0)
    jobs = B.book(1000)
    assert len(jobs) == 1000
    n_items = sum((job[2] + 255) // 256 for job in jobs)
    assert n_items // (256 * 4) >= 16
    want = price_each(jobs)
    try:
        for bpc, chunk in FULL_TILES:
            _set_tile(eng, bpc, chunk)
            assert_identical(K.price_portfolio(jobs), want)
    finally:
        _set_tile(eng, 0, 1)


def test_full_book_greeks_equal_per_contract():
    eng = N.engine(0)
    jobs = B.book(1000)
    want = per_contract_greeks(jobs)
    try:
        for bpc, chunk in FULL_TILES[:2]:
            _set_tile(eng, bpc, chunk)
            assert_greeks_identical(K.price_portfolio(jobs, greeks=True), want)
    finally:
        _set_tile(eng, 0, 1)


def test_book_slice_vs_oracle():
    """The first 16 jobs of the book (4 of each kind, GBM and Heston, monthly
    fixings, 5-asset baskets) at 8,192 paths each, in one persistent launch
    under the tuned tile, against the oracle on the same Philox stream:
    continuous payoffs agree to FP32 path rounding, knock-outs within half a
    standard error (paths on the level may flip)."""
    import oracle_lib as O
    from test_gpu_parity import oracle_model

    eng = N.engine(0)
    jobs = [(c, m, 8192, seed, spy, grid) for c, m, _, seed, spy, grid in B.book(16)]
    try:
        _set_tile(eng, 4, 4)
        got = K.price_portfolio(jobs)
    finally:
        _set_tile(eng, 0, 1)
    for i, (job, g) in enumerate(zip(jobs, got)):
        c, m, n, seed, spy, grid = job
        want, _, _ = O.price(c.to_json(), oracle_model(m), O.cfg(n, seed, spy, m.rate, grid=grid, threads=16))
        smooth = B.KINDS[i % 4] != "barrier"
        tol = 5e-5 * (abs(want["price"]) + 100.0) if smooth else 0.5 * want["std_error"]
        print(f"job {i} {B.KINDS[i % 4]}: gpu {g.price:.8f} oracle {want['price']:.8f} se {want['std_error']:.2e}")
        assert abs(g.price - want["price"]) <= tol, i
        assert g.n_paths == n
