"""kt_price_noalloc (the allocation-free kt_price) through the C ABI: on
consecutive calls, the second one smaller, its results equal kt_price's on
the same inputs to the last bit; sizes beyond its static scratch array and
variance-reduction flags are refused before any launch, and the array still
serves the next call."""
import ctypes as C

import pytest

import kontracts_amd as K
from kontracts_amd import _native as N
from kontracts_amd.dsl import _cfg, _model_struct

pytestmark = pytest.mark.gpu

MAX_PATHS = 1 << 26


def _lib():
    return N.lib()


def call(fn, eng, contract, model, n, seed, spy=50, grid=None, anti=False):
    m = _model_struct(model)
    cfg = _cfg(n, seed, spy, model.rate, grid, anti)
    out = N.KtPriceResult()
    rc = fn(eng.handle, contract.to_json().encode(), C.byref(m.s), C.byref(cfg.s), C.byref(out))
    return rc, out


def fields(r):
    return (r.price, r.sample_std, r.std_error, r.ci95_low, r.ci95_high, r.n_paths)


CASES = [
    # (contract, model, paths, seed, steps a year): each pair of rows runs
    # back to back, the second smaller than the first
    (K.up_and_out_call("X", 100.0, 200.0, 1.0), K.GBM(180.0, 0.25, 0.05, asset="X"), 4_194_304, 42, 252),
    (K.european_call("X", 100.0, 1.0), K.GBM(100.0, 0.2, 0.05, asset="X"), 5_000, 7, 50),
    (K.scale((K.average(K.spot("X")) - 100.0).clip(0.0), K.one("USD")) @ K.at(1.0),
     K.heston(100.0, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X"), 1 << 20, 3, 365),
    (K.european_call("X", 95.0, 0.5), K.heston(100.0, 0.04, 2.0, 0.04, 0.3, -0.5, 0.05, asset="X"), 257, 4, 12),
]


def test_noalloc_equals_kt_price_on_consecutive_calls():
    eng = N.engine(0)
    L = _lib()
    for contract, model, n, seed, spy in CASES:
        rc, got = call(L.kt_price_noalloc, eng, contract, model, n, seed, spy)
        assert rc == 0, N.lib().kt_last_error(eng.handle)
        rc2, want = call(L.kt_price, eng, contract, model, n, seed, spy)
        assert rc2 == 0
        assert fields(got) == fields(want), (n, fields(got), fields(want))
    # a 5-asset basket, then a smaller one, again back to back
    fs = [K.GbmFactor(f"A{i}", 100.0, 0.03, 0.15 + 0.05 * i) for i in range(5)]
    b5 = K.correlated_gbm(fs, [[1.0 if i == j else 0.3 for j in range(5)] for i in range(5)], 0.03)
    pay = K.scale((sum((K.spot(f"A{i}") for i in range(1, 5)), K.spot("A0")) / 5.0 - 100.0).clip(0.0),
                  K.one("USD")) @ K.at(1.0)
    for n in (1 << 20, 1000):
        rc, got = call(L.kt_price_noalloc, eng, pay, b5, n, 11, 252)
        assert rc == 0
        _, want = call(L.kt_price, eng, pay, b5, n, 11, 252)
        assert fields(got) == fields(want)


def test_noalloc_limits_and_errors():
    eng = N.engine(0)
    L = _lib()
    c, m = K.european_call("X", 100.0, 1.0), K.GBM(100.0, 0.2, 0.05, asset="X")
    rc, _ = call(L.kt_price_noalloc, eng, c, m, MAX_PATHS + 1, 1)
    assert rc == 100  # InvalidArgument
    assert b"paths fit its scratch array" in N.lib().kt_last_error(eng.handle)
    rc, _ = call(L.kt_price_noalloc, eng, c, m, 4096, 1, anti=True)
    assert rc == 7  # Unsupported
    rc, _ = call(L.kt_price_noalloc, eng, K.european_call("Y", 100.0, 1.0), m, 4096, 1)
    assert rc == 1  # UnknownAsset, as kt_price
    # a path error raised inside the kernel is reported, and cleared for the next call
    horizon = K.one("USD") @ K.at(1.0)
    bad = horizon + K.when(K.spot("X") >= 101.0, K.one("USD").or_(K.one("USD")))
    rc, _ = call(L.kt_price_noalloc, eng, bad, m, 4096, 1, spy=12)
    assert rc == 7
    # the array serves the next call normally
    rc, got = call(L.kt_price_noalloc, eng, c, m, 4096, 1)
    assert rc == 0
    _, want = call(L.kt_price, eng, c, m, 4096, 1)
    assert fields(got) == fields(want)
    rc, got = call(L.kt_price_noalloc, eng, c, m, 0, 1)
    assert rc == 0 and got.n_paths == 0
