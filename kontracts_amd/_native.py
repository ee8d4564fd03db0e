"""ctypes binding of the engine's C ABI (include/kontracts.h).

The shared library is built in-tree by ``python -m kontracts_amd.build``.  There
is no fallback: if the library or a GPU is missing, the pricing calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkontracts_amd.so")

KT_MODEL_GBM = 0
KT_MODEL_HESTON = 1
KT_MODEL_CORRELATED_GBM = 2

STATUS_NAMES = {
    0: "Ok",
    1: "UnknownAsset",
    2: "TimeOutOfRange",
    3: "StepOutOfRange",
    4: "InconsistentPath",
    5: "Serde",
    6: "MalformedContract",
    7: "Unsupported",
    100: "InvalidArgument",
    101: "Device",
    102: "OutOfMemory",
}


class KontractError(ValueError):
    """A KontractError from the engine (the reference maps these to ValueError,
    src/bindings.rs:50-52).  ``kind`` is the variant name."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = STATUS_NAMES.get(code, str(code))
        super().__init__(f"{self.kind}: {message}")


class DeviceError(RuntimeError):
    """The HIP runtime or a kernel failed, or no GPU is available."""


class KtModel(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("n_assets", C.c_int32),
        ("asset_names", C.POINTER(C.c_char_p)),
        ("s0", C.POINTER(C.c_double)),
        ("mu", C.POINTER(C.c_double)),
        ("sigma", C.POINTER(C.c_double)),
        ("corr", C.POINTER(C.c_double)),
        ("v0", C.c_double),
        ("kappa", C.c_double),
        ("theta", C.c_double),
        ("sigma_v", C.c_double),
        ("rho", C.c_double),
        ("heston_scheme", C.c_int32),
        ("reserved", C.c_int32),
    ]


class KtMcConfig(C.Structure):
    _fields_ = [
        ("n_paths", C.c_uint64),
        ("seed", C.c_uint64),
        ("steps_per_year", C.c_uint64),
        ("rate", C.c_double),
        ("grid", C.POINTER(C.c_double)),
        ("n_grid", C.c_uint64),
        ("antithetic", C.c_uint32),
        ("control_variate", C.c_uint32),
    ]


class KtPriceResult(C.Structure):
    _fields_ = [
        ("price", C.c_double),
        ("sample_std", C.c_double),
        ("std_error", C.c_double),
        ("ci95_low", C.c_double),
        ("ci95_high", C.c_double),
        ("n_paths", C.c_uint64),
    ]


class KtBumps(C.Structure):
    _fields_ = [("spot", C.c_double), ("vol", C.c_double), ("rate", C.c_double)]


class KtGreeks(C.Structure):
    _fields_ = [
        ("price", C.c_double),
        ("delta", C.c_double),
        ("gamma", C.c_double),
        ("vega", C.c_double),
        ("rho", C.c_double),
        ("std_error", C.c_double * 5),
    ]


class KtJob(C.Structure):
    _fields_ = [("contract_json", C.c_char_p), ("model", KtModel), ("cfg", KtMcConfig)]


# Every symbol the header declares, with its signature.
_P = C.POINTER
_SIGNATURES = {
    "kt_engine_create": (C.c_int, [C.c_int, _P(C.c_void_p)]),
    "kt_engine_destroy": (None, [C.c_void_p]),
    "kt_last_error": (C.c_char_p, [C.c_void_p]),
    "kt_last_error_global": (C.c_char_p, []),
    "kt_price": (C.c_int, [C.c_void_p, C.c_char_p, _P(KtModel), _P(KtMcConfig), _P(KtPriceResult)]),
    "kt_price_noalloc": (C.c_int, [C.c_void_p, C.c_char_p, _P(KtModel), _P(KtMcConfig), _P(KtPriceResult)]),
    "kt_price_batch": (
        C.c_int,
        [C.c_void_p, _P(C.c_char_p), C.c_uint64, _P(KtModel), _P(KtMcConfig), _P(KtPriceResult)],
    ),
    "kt_greeks_gbm": (
        C.c_int,
        [C.c_void_p, C.c_char_p, C.c_char_p, C.c_double, C.c_double, _P(KtMcConfig), _P(KtBumps), _P(KtGreeks)],
    ),
    "kt_greeks_bump": (
        C.c_int,
        [C.c_void_p, C.c_char_p, _P(KtModel), _P(KtMcConfig), _P(C.c_double), C.c_uint64, _P(C.c_double),
         _P(C.c_double)],
    ),
    "kt_greeks_adjoint": (
        C.c_int,
        [C.c_void_p, C.c_char_p, _P(KtModel), _P(KtMcConfig), _P(C.c_double), _P(C.c_double), C.c_uint64],
    ),
    "kt_price_american": (
        C.c_int,
        [C.c_void_p, C.c_char_p, _P(C.c_double), C.c_uint64, _P(KtModel), _P(KtMcConfig), C.c_int32,
         _P(KtPriceResult)],
    ),
    "kt_debug_american": (
        C.c_int,
        [C.c_void_p, C.c_char_p, _P(C.c_double), C.c_uint64, _P(KtModel), _P(KtMcConfig), C.c_int32,
         _P(C.c_double), _P(C.c_double), _P(KtPriceResult)],
    ),
    "kt_price_portfolio": (C.c_int, [C.c_void_p, _P(KtJob), C.c_uint64, _P(KtPriceResult)]),
    "kt_price_portfolio_greeks": (
        C.c_int, [C.c_void_p, _P(KtJob), C.c_uint64, _P(KtPriceResult), _P(C.c_double), _P(C.c_double), C.c_uint64]),
    "kt_greek_surface": (
        C.c_int,
        [C.c_void_p, C.c_char_p, C.c_char_p, _P(C.c_double), C.c_uint64, _P(C.c_double), C.c_uint64,
         _P(KtMcConfig), _P(KtBumps), _P(KtGreeks)],
    ),
    "kt_engine_set_profiling": (C.c_int, [C.c_void_p, C.c_int32]),
    "kt_engine_set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "kt_engine_kernel_time": (
        C.c_int, [C.c_void_p, _P(C.c_double), _P(C.c_uint64), _P(C.c_double), C.c_int32]),
    "kt_plan": (
        C.c_int,
        [C.c_char_p, C.c_uint64, C.c_int32, _P(C.c_double), _P(C.c_uint64), _P(C.c_double), _P(C.c_uint64),
         C.c_uint64, _P(C.c_int32), _P(C.c_double), C.c_char_p, C.c_uint64],
    ),
    "kt_describe_program": (C.c_int, [C.c_char_p, _P(KtModel), _P(KtMcConfig), C.c_char_p, C.c_uint64]),
    "kt_black_scholes_call": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
    "kt_debug_philox": (
        C.c_int, [C.c_void_p, _P(C.c_uint32), _P(C.c_uint32), C.c_uint64, _P(C.c_uint32)]),
    "kt_debug_normals": (C.c_int, [C.c_void_p, _P(C.c_uint32), C.c_uint64, _P(C.c_float)]),
    "kt_debug_copy": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, _P(C.c_double)]),
    "kt_debug_valu": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _P(C.c_double), _P(C.c_double)]),
    "kt_debug_rng_rate": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _P(C.c_double), _P(C.c_double)]),
    "kt_debug_paths": (
        C.c_int,
        [C.c_void_p, C.c_char_p, _P(KtModel), _P(KtMcConfig), _P(C.c_double), _P(C.c_float), _P(C.c_double),
         C.c_uint64, _P(C.c_uint64), _P(KtPriceResult)],
    ),
}

_lib = None
_lib_lock = threading.Lock()


def lib():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"kontracts_amd: {LIB_PATH} is missing; build it with `python -m kontracts_amd.build`")
            handle = C.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def check(status: int, engine=None) -> None:
    if status == 0:
        return
    L = lib()
    msg = (L.kt_last_error(engine) if engine else L.kt_last_error_global()) or b""
    text = msg.decode("utf-8", "replace")
    if status == 101:
        raise DeviceError(text)
    if status == 102:
        raise MemoryError(text)
    raise KontractError(status, text)


class Engine:
    """One engine (HIP stream + device buffers) on one GPU."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        check(L.kt_engine_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def set_profiling(self, on: bool) -> None:
        check(lib().kt_engine_set_profiling(self.handle, 1 if on else 0), self.handle)

    def set_option(self, name: str, value: int) -> None:
        """Engine tuning knob (see kt_engine_set_option in include/kontracts.h)."""
        check(lib().kt_engine_set_option(self.handle, name.encode(), int(value)), self.handle)

    def kernel_time(self, reset: bool = False):
        """(total ms, launches, path-steps) of the path kernels since the last reset."""
        ms, n, ps = C.c_double(), C.c_uint64(), C.c_double()
        check(lib().kt_engine_kernel_time(self.handle, C.byref(ms), C.byref(n), C.byref(ps), 1 if reset else 0),
              self.handle)
        return ms.value, n.value, ps.value

    def close(self):
        if getattr(self, "handle", None):
            lib().kt_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_engines: dict[int, Engine] = {}
_engines_lock = threading.Lock()


def engine(device: int | None = None) -> Engine:
    if device is None:
        device = int(os.environ.get("KT_DEVICE", "0"))
    with _engines_lock:
        e = _engines.get(device)
        if e is None:
            e = Engine(device)
            _engines[device] = e
        return e


def exported_symbols() -> list[str]:
    return sorted(_SIGNATURES)
