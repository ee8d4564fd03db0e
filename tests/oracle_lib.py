"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) — the checker.

Only tests, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py
use this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")


class OrcModel(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("n_assets", C.c_int), ("names", C.POINTER(C.c_char_p)),
        ("s0", C.POINTER(C.c_double)), ("mu", C.POINTER(C.c_double)), ("sigma", C.POINTER(C.c_double)),
        ("corr", C.POINTER(C.c_double)), ("v0", C.c_double), ("kappa", C.c_double), ("theta", C.c_double),
        ("sigma_v", C.c_double), ("rho", C.c_double), ("heston_scheme", C.c_int),
    ]


class OrcCfg(C.Structure):
    _fields_ = [
        ("n_paths", C.c_uint64), ("seed", C.c_uint64), ("steps_per_year", C.c_uint64), ("rate", C.c_double),
        ("grid", C.POINTER(C.c_double)), ("n_grid", C.c_uint64), ("threads", C.c_int),
    ]


class OrcResult(C.Structure):
    _fields_ = [("price", C.c_double), ("sample_std", C.c_double), ("std_error", C.c_double),
                ("ci95_low", C.c_double), ("ci95_high", C.c_double), ("n_paths", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OracleError(Exception):
    def __init__(self, code, msg):
        self.code = code
        super().__init__(f"oracle error {code}: {msg}")


_L = None


def load():
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    L = C.CDLL(ORACLE_LIB)
    d = C.c_double
    for f in ["orc_bs_call", "orc_bs_put", "orc_uo_call", "orc_heston_call", "orc_crr_american", "orc_inv_norm",
              "orc_uniform_from_word", "orc_bs_call_as", "orc_norm_cdf_as"]:
        getattr(L, f).restype = d
    L.orc_bs_call.argtypes = [d] * 5
    L.orc_bs_call_as.argtypes = [d] * 5
    L.orc_norm_cdf_as.argtypes = [d]
    L.orc_bs_put.argtypes = [d] * 5
    L.orc_uo_call.argtypes = [d] * 6
    L.orc_heston_call.argtypes = [d] * 9
    L.orc_crr_american.argtypes = [d, d, d, d, d, C.c_int, C.c_int]
    L.orc_inv_norm.argtypes = [d]
    L.orc_uniform_from_word.argtypes = [C.c_uint32]
    L.orc_bs_call_greeks.argtypes = [d] * 5 + [C.POINTER(d)]
    L.orc_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(d)]
    _L = L
    return L


def _dbl(xs):
    xs = list(xs)
    return (C.c_double * max(len(xs), 1))(*xs)


class Model:
    """Holds an OrcModel and its buffers."""

    def __init__(self, kind, names, s0, mu, sigma, corr=None, v0=0.0, kappa=0.0, theta=0.0, sigma_v=0.0, rho=0.0,
                 scheme=0):
        self._names = (C.c_char_p * len(names))(*[n.encode() for n in names])
        self._s0, self._mu, self._sg = _dbl(s0), _dbl(mu), _dbl(sigma)
        self._corr = _dbl([x for row in corr for x in row]) if corr is not None else None
        self.s = OrcModel(kind, len(names), self._names, self._s0, self._mu, self._sg, self._corr, v0, kappa, theta,
                          sigma_v, rho, scheme)


def gbm(asset, s0, mu, sigma):
    return Model(0, [asset], [s0], [mu], [sigma])


def heston(asset, s0, r, v0, kappa, theta, sigma_v, rho, scheme=0):
    return Model(1, [asset], [s0], [r], [0.0], None, v0, kappa, theta, sigma_v, rho, scheme)


def basket(names, s0, mu, sigma, corr):
    return Model(2, names, s0, mu, sigma, corr)


def cfg(n_paths, seed, steps_per_year, rate, grid=None, threads=0):
    g = _dbl(grid) if grid is not None else None
    c = OrcCfg(n_paths, seed, steps_per_year, rate, g, len(grid) if grid is not None else 0, threads)
    c._keep = g
    return c


def _err():
    return C.create_string_buffer(512)


def price(json_s, model, c, normals=None, want_pv=False):
    """Returns (result dict, pv array or None, grid list)."""
    L = load()
    res = OrcResult()
    e = _err()
    pv = np.zeros(c.n_paths, dtype=np.float64) if want_pv else None
    nz = None
    if normals is not None:
        nz = np.ascontiguousarray(normals, dtype=np.float32)
    grid = (C.c_double * 100000)()
    ng = C.c_uint64()
    rc = L.orc_price(json_s.encode(), C.byref(model.s), C.byref(c),
                     nz.ctypes.data_as(C.POINTER(C.c_float)) if nz is not None else None,
                     pv.ctypes.data_as(C.POINTER(C.c_double)) if pv is not None else None,
                     grid, C.byref(ng), 100000, C.byref(res), e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    return res.as_dict(), pv, list(grid[: ng.value])


def price_vr(json_s, model, c, antithetic=False, control_variate=False):
    """price_gbm with a VarianceReductionConfig (src/pricer.rs:117-228)."""
    L = load()
    res = OrcResult()
    e = _err()
    rc = L.orc_price_vr(json_s.encode(), C.byref(model.s), C.byref(c), int(antithetic), int(control_variate),
                        C.byref(res), e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    return res.as_dict()


def price_batch(jsons, model, c):
    L = load()
    arr = (C.c_char_p * len(jsons))(*[j.encode() for j in jsons])
    out = (OrcResult * len(jsons))()
    e = _err()
    rc = L.orc_price_batch(arr, len(jsons), C.byref(model.s), C.byref(c), out, e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    return [o.as_dict() for o in out]


def greeks_gbm(json_s, asset, s0, sigma, c, h_spot=1e-2, h_vol=1e-2, h_rate=1e-4):
    L = load()
    out = (C.c_double * 5)()
    e = _err()
    rc = L.orc_greeks_gbm(json_s.encode(), asset.encode(), C.c_double(s0), C.c_double(sigma), C.byref(c),
                          C.c_double(h_spot), C.c_double(h_vol), C.c_double(h_rate), out, e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    return dict(zip(["price", "delta", "gamma", "vega", "rho"], list(out)))


def greek_surface(json_s, asset, spots, vols, c, h_spot=1e-2, h_vol=1e-2, h_rate=1e-4):
    """Cells row-major [spot][vol], each a greeks_gbm dict."""
    L = load()
    ns, nv = len(spots), len(vols)
    out = (C.c_double * max(5 * ns * nv, 1))()
    e = _err()
    rc = L.orc_greek_surface(json_s.encode(), asset.encode(), _dbl(spots) if ns else None, ns,
                             _dbl(vols) if nv else None, nv, C.byref(c), C.c_double(h_spot), C.c_double(h_vol),
                             C.c_double(h_rate), out, e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    keys = ["price", "delta", "gamma", "vega", "rho"]
    return [dict(zip(keys, out[5 * i:5 * i + 5])) for i in range(ns * nv)]


def greeks_heston_bump(json_s, model, c, h):
    L = load()
    out = (C.c_double * 7)()
    e = _err()
    rc = L.orc_greeks_heston_bump(json_s.encode(), C.byref(model.s), C.byref(c), _dbl(h), out, e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    return list(out)


def price_american(json_s, dates, model, c, n_basis=3, want_pv=False):
    L = load()
    res = OrcResult()
    e = _err()
    pv = np.zeros(c.n_paths) if want_pv else None
    rc = L.orc_price_american(json_s.encode(), _dbl(dates), len(dates), C.byref(model.s), C.byref(c), n_basis,
                              pv.ctypes.data_as(C.POINTER(C.c_double)) if pv is not None else None, C.byref(res),
                              e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    return res.as_dict(), pv


def compile_plan(json_s, spy=0, force=False):
    L = load()
    grid = (C.c_double * 100000)()
    dates = (C.c_double * 100000)()
    ng, nd = C.c_uint64(), C.c_uint64()
    fine, hz = C.c_int(), C.c_double()
    assets = C.create_string_buffer(4096)
    e = _err()
    rc = L.orc_compile(json_s.encode(), C.c_uint64(spy), 1 if force else 0, grid, C.byref(ng), C.c_uint64(100000),
                       C.byref(fine), C.byref(hz), dates, C.byref(nd), assets, 4096, e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    a = assets.value.decode()
    return {"grid": list(grid[: ng.value]), "fixed_dates": list(dates[: nd.value]), "horizon": hz.value,
            "needs_fine_grid": bool(fine.value), "assets": a.split("\n") if a else []}


def eval_path(json_s, times, spots: dict, rate=0.0):
    """Cashflows of a contract on a hand-built path: (pv, [(amount, step)])."""
    L = load()
    names = list(spots)
    n = len(times)
    arr = np.array([spots[k] for k in names], dtype=np.float64).reshape(-1)
    nm = (C.c_char_p * max(len(names), 1))(*[x.encode() for x in names])
    pv = C.c_double()
    amt = (C.c_double * 1024)()
    idx = (C.c_uint64 * 1024)()
    nf = C.c_uint64()
    e = _err()
    rc = L.orc_eval_path(json_s.encode(), _dbl(times), n, nm, len(names), arr.ctypes.data_as(C.POINTER(C.c_double)),
                         C.c_double(rate), C.byref(pv), amt, idx, C.byref(nf), 1024, e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    return pv.value, [(amt[i], idx[i]) for i in range(nf.value)]


def eval_observable(json_s, times, spots: dict, t):
    L = load()
    names = list(spots)
    arr = np.array([spots[k] for k in names], dtype=np.float64).reshape(-1)
    nm = (C.c_char_p * max(len(names), 1))(*[x.encode() for x in names])
    out = C.c_double()
    e = _err()
    rc = L.orc_eval_observable(json_s.encode(), _dbl(times), len(times), nm, len(names),
                               arr.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(t), C.byref(out), e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    return out.value


def philox(ctr, key):
    L = load()
    out = (C.c_uint32 * 4)()
    L.orc_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return list(out)


def normals(seed, path, first, count):
    L = load()
    out = (C.c_double * count)()
    L.orc_normals(seed, path, first, count, out)
    return np.array(out[:])


def bs_call(s, k, t, r, sg):
    return load().orc_bs_call(s, k, t, r, sg)


def bs_call_as(s, k, t, r, sg):
    """The reference's black_scholes_call (Abramowitz-Stegun normal CDF)."""
    return load().orc_bs_call_as(s, k, t, r, sg)


def norm_cdf_as(x):
    return load().orc_norm_cdf_as(x)


def bs_put(s, k, t, r, sg):
    return load().orc_bs_put(s, k, t, r, sg)


def bs_call_greeks(s, k, t, r, sg):
    out = (C.c_double * 4)()
    load().orc_bs_call_greeks(s, k, t, r, sg, out)
    return dict(zip(["delta", "gamma", "vega", "rho"], list(out)))


def uo_call(s, k, h, t, r, sg):
    return load().orc_uo_call(s, k, h, t, r, sg)


def heston_call(s0, k, t, r, v0, kappa, theta, sv, rho):
    return load().orc_heston_call(s0, k, t, r, v0, kappa, theta, sv, rho)


def crr_american(s0, k, t, r, sg, n, is_call):
    return load().orc_crr_american(s0, k, t, r, sg, n, 1 if is_call else 0)


def summarize(pvs):
    L = load()
    a = np.ascontiguousarray(pvs, dtype=np.float64)
    res = OrcResult()
    L.orc_summarize(a.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(len(a)), C.byref(res))
    return res.as_dict()


def simulate(model, grid, n_paths, seed, threads=0):
    """Simulated spots [n_paths][n_assets][n_grid] (Simulator::simulate_paths)."""
    L = load()
    K = model.s.n_assets
    out = np.zeros((n_paths, K, len(grid)), dtype=np.float64)
    e = _err()
    rc = L.orc_simulate(C.byref(model.s), _dbl(grid), C.c_uint64(len(grid)), C.c_uint64(n_paths), C.c_uint64(seed),
                        threads, out.ctypes.data_as(C.POINTER(C.c_double)), e, 512)
    if rc:
        raise OracleError(rc, e.value.decode())
    return out
