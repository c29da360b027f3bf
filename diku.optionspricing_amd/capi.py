"""ctypes binding of the C ABI in include/hwtree.h (libhwtree.so).

The library is the product: it must be built (``__graft_entry__.build()``)
before use, and every call here raises if it is missing -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .formats import Options, Yield

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libhwtree.so"

HWT_OK, HWT_EINVAL, HWT_EUNSUPPORTED, HWT_ENOMEM, HWT_EHIP = 0, -1, -2, -3, -4
FORM_AUTO, FORM_LANE, FORM_WAVE = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
FLAG_NO_GRAPH, FLAG_NO_ROWS, FLAG_CONCURRENT, FLAG_FUSED, FLAG_NO_FLAT, FLAG_DIRECT = 1, 2, 4, 8, 16, 32

# The exported C symbols, in include/hwtree.h order.
EXPORTS = (
    "hwt_config_init", "hwt_price_f64", "hwt_plan_create", "hwt_plan_execute",
    "hwt_plan_sync", "hwt_plan_destroy", "hwt_price_f32", "hwt_price_constants_f64", "hwt_price_constants_f32",
    "hwt_batch_create", "hwt_batch_create_f32", "hwt_batch_price", "hwt_batch_device_memory", "hwt_batch_destroy",
    "hwt_price_device_noalloc", "hwt_alphas_f64", "hwt_autotune", "hwt_max_width", "hwt_strerror",
    "hwt_last_error", "hwt_abi_version",
)


class HwtError(RuntimeError):
    def __init__(self, code: int, detail: str):
        super().__init__(f"hwtree error {code}: {detail}")
        self.code = code
        self.detail = detail


class OptionsC(C.Structure):
    _fields_ = [("n", C.c_int32), ("strike", C.c_void_p), ("maturity", C.c_void_p),
                ("length", C.c_void_p), ("term_unit", C.c_void_p), ("term_steps", C.c_void_p),
                ("a", C.c_void_p), ("sigma", C.c_void_p), ("type", C.c_void_p)]


class YieldC(C.Structure):
    _fields_ = [("n", C.c_int32), ("rate", C.c_void_p), ("day", C.c_void_p)]


# hwt_options_f32 / hwt_yield_f32: the same layouts with float arrays
OptionsF32C, YieldF32C = OptionsC, YieldC


class ConstantsC(C.Structure):
    """hwt_option_constants_f64 (trinom::OptionConstants)."""
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("dr", C.c_double), ("X", C.c_double),
                ("M", C.c_double), ("jmax", C.c_int32), ("n", C.c_int32), ("width", C.c_int32),
                ("term_unit", C.c_uint16), ("type", C.c_int8)]


class ConstantsF32C(C.Structure):
    """hwt_option_constants_f32 (trinom::OptionConstants with real = float)."""
    _fields_ = [("t", C.c_float), ("dt", C.c_float), ("dr", C.c_float), ("X", C.c_float),
                ("M", C.c_float), ("jmax", C.c_int32), ("n", C.c_int32), ("width", C.c_int32),
                ("term_unit", C.c_uint16), ("type", C.c_int8)]


class ConfigC(C.Structure):
    _fields_ = [("form", C.c_int32), ("precision", C.c_int32), ("block_size", C.c_int32),
                ("sort", C.c_int32), ("device", C.c_int32), ("flags", C.c_int32),
                ("wide_per_cu", C.c_int32), ("narrow_per_cu", C.c_int32), ("lane_max_width", C.c_int32)]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["sort"] = chr(d["sort"]) if d["sort"] else "L"
        return d


class TuneReportC(C.Structure):
    _fields_ = [("ms", C.c_double), ("default_ms", C.c_double), ("tune_ms", C.c_double),
                ("candidates", C.c_int32), ("cached", C.c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class StatsC(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("wide_ms", C.c_double),
                ("narrow_ms", C.c_double), ("device_memory", C.c_int64), ("node_updates", C.c_int64),
                ("wide_node_updates", C.c_int64), ("seq_node_updates", C.c_int64), ("tasks", C.c_int32),
                ("wide_tasks", C.c_int32), ("unsupported", C.c_int32), ("fp64_repriced", C.c_int32),
                ("wide_step_ops", C.c_int64), ("wide_unit_slots", C.c_int64), ("wide_plain_slots", C.c_int64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        L.hwt_config_init.argtypes = [C.POINTER(ConfigC)]
        L.hwt_config_init.restype = None
        L.hwt_price_f64.argtypes = [C.POINTER(OptionsC), C.POINTER(YieldC), P, C.POINTER(ConfigC), C.POINTER(StatsC)]
        L.hwt_plan_create.argtypes = [C.POINTER(P), C.c_int32, C.c_int32, C.POINTER(ConfigC)]
        L.hwt_plan_execute.argtypes = [P, C.POINTER(OptionsC), C.POINTER(YieldC), P, P]
        L.hwt_plan_sync.argtypes = [P, C.POINTER(StatsC)]
        L.hwt_plan_destroy.argtypes = [P]
        L.hwt_plan_destroy.restype = None
        L.hwt_price_f32.argtypes = [C.POINTER(OptionsF32C), C.POINTER(YieldF32C), P, C.POINTER(ConfigC),
                                    C.POINTER(StatsC)]
        L.hwt_price_constants_f64.argtypes = [P, C.c_int32, C.POINTER(YieldC), P, C.POINTER(ConfigC),
                                              C.POINTER(StatsC)]
        L.hwt_price_constants_f32.argtypes = [P, C.c_int32, C.POINTER(YieldF32C), P, C.POINTER(ConfigC),
                                              C.POINTER(StatsC)]
        L.hwt_batch_create.argtypes = [C.POINTER(P), C.POINTER(OptionsC), C.POINTER(YieldC)]
        L.hwt_batch_create_f32.argtypes = [C.POINTER(P), C.POINTER(OptionsF32C), C.POINTER(YieldF32C)]
        L.hwt_batch_price.argtypes = [P, P, C.POINTER(ConfigC), C.POINTER(StatsC)]
        L.hwt_batch_device_memory.argtypes = [P]
        L.hwt_batch_device_memory.restype = C.c_int64
        L.hwt_batch_destroy.argtypes = [P]
        L.hwt_batch_destroy.restype = None
        L.hwt_price_device_noalloc.argtypes = [C.POINTER(OptionsC), C.POINTER(YieldC), P, P]
        L.hwt_alphas_f64.argtypes = [C.POINTER(OptionsC), C.POINTER(YieldC), P, P]
        L.hwt_autotune.argtypes = [C.POINTER(OptionsC), C.POINTER(YieldC), P, C.POINTER(ConfigC),
                                   C.POINTER(TuneReportC)]
        L.hwt_max_width.argtypes = [C.c_int32]
        L.hwt_max_width.restype = C.c_int32
        L.hwt_strerror.argtypes = [C.c_int]
        L.hwt_strerror.restype = C.c_char_p
        L.hwt_last_error.restype = C.c_char_p
        L.hwt_abi_version.restype = C.c_int32
        _lib = L
    return _lib


def _check(rc: int):
    if rc != HWT_OK:
        raise HwtError(rc, lib().hwt_last_error().decode() or lib().hwt_strerror(rc).decode())


def config(form=FORM_AUTO, precision=PREC_FP64, block_size=0, sort="L", device=-1, flags=0) -> ConfigC:
    c = ConfigC()
    lib().hwt_config_init(C.byref(c))
    c.form, c.precision, c.block_size, c.device, c.flags = form, precision, block_size, device, flags
    c.sort = ord(sort) if isinstance(sort, str) else int(sort)
    return c


def _host_opts(o: Options) -> OptionsC:
    return OptionsC(o.N, o.strike.ctypes.data, o.maturity.ctypes.data, o.length.ctypes.data,
                    o.term_unit.ctypes.data, o.term_steps.ctypes.data, o.a.ctypes.data,
                    o.sigma.ctypes.data, o.type.ctypes.data)


def _host_yield(y: Yield) -> YieldC:
    return YieldC(y.N, y.rates.ctypes.data, y.days.ctypes.data)


def price(opts: Options, y: Yield, cfg: ConfigC | None = None, stats: bool = False):
    """hwt_price_f64 on host arrays; returns prices (and the stats dict)."""
    out = np.empty(opts.N, dtype=np.float64)
    st = StatsC()
    rc = lib().hwt_price_f64(C.byref(_host_opts(opts)), C.byref(_host_yield(y)), out.ctypes.data,
                             C.byref(cfg) if cfg is not None else None, C.byref(st))
    _check(rc)
    return (out, st.as_dict()) if stats else out


def price_status(opts: Options, y: Yield, cfg: ConfigC | None = None):
    """Like price() but returns (return code, prices) instead of raising."""
    out = np.empty(opts.N, dtype=np.float64)
    rc = lib().hwt_price_f64(C.byref(_host_opts(opts)), C.byref(_host_yield(y)), out.ctypes.data,
                             C.byref(cfg) if cfg is not None else None, None)
    return rc, out


def price_f32(opts: Options, y: Yield, cfg: ConfigC | None = None, stats: bool = False):
    """hwt_price_f32: the batch's real fields as float32 host arrays (a
    default-built reference caller's data), float32 prices back."""
    f = {k: np.ascontiguousarray(getattr(opts, k), dtype=np.float32)
         for k in ("strike", "maturity", "length", "a", "sigma")}
    rates = np.ascontiguousarray(y.rates, dtype=np.float32)
    o = OptionsF32C(opts.N, f["strike"].ctypes.data, f["maturity"].ctypes.data, f["length"].ctypes.data,
                    opts.term_unit.ctypes.data, opts.term_steps.ctypes.data, f["a"].ctypes.data,
                    f["sigma"].ctypes.data, opts.type.ctypes.data)
    yc = YieldF32C(y.N, rates.ctypes.data, y.days.ctypes.data)
    out = np.empty(opts.N, dtype=np.float32)
    st = StatsC()
    _check(lib().hwt_price_f32(C.byref(o), C.byref(yc), out.ctypes.data, C.byref(cfg) if cfg is not None else None,
                               C.byref(st)))
    return (out, st.as_dict()) if stats else out


def constants_of(opts: Options) -> np.ndarray:
    """Per-option trinom::OptionConstants (common/OptionConstants.hpp:29-53),
    computed on the host in FP64 as an array of ConstantsC."""
    arr = (ConstantsC * opts.N)()
    for i in range(opts.N):
        tuy = int(np.ceil(365.0 / float(opts.term_unit[i])))
        steps = int(opts.term_steps[i])
        n = int(steps * tuy * float(opts.maturity[i]))
        dt = tuy / float(steps)
        a, sg = float(opts.a[i]), float(opts.sigma[i])
        V = sg * sg * (1.0 - np.exp(-2.0 * a * dt)) / (2.0 * a)
        M = float(np.exp(-a * dt) - 1.0)
        jmax = int(-0.184 / M) + 1
        arr[i] = ConstantsC(float(opts.length[i]), dt, float(np.sqrt(3.0 * V)), float(opts.strike[i]), M, jmax, n,
                            2 * jmax + 1, int(opts.term_unit[i]), int(opts.type[i]))
    return arr


def price_constants(consts, y: Yield, cfg: ConfigC | None = None, stats: bool = False):
    """hwt_price_constants_f64 over an array of ConstantsC."""
    n = len(consts)
    out = np.empty(n, dtype=np.float64)
    st = StatsC()
    _check(lib().hwt_price_constants_f64(C.cast(consts, C.c_void_p), n, C.byref(_host_yield(y)), out.ctypes.data,
                                         C.byref(cfg) if cfg is not None else None, C.byref(st)))
    return (out, st.as_dict()) if stats else out


def price_constants_f32(consts, y: Yield, cfg: ConfigC | None = None):
    """hwt_price_constants_f32 over an array of ConstantsF32C (float prices)."""
    n = len(consts)
    out = np.empty(n, dtype=np.float32)
    rates = np.ascontiguousarray(y.rates, dtype=np.float32)
    yc = YieldF32C(y.N, rates.ctypes.data, y.days.ctypes.data)
    _check(lib().hwt_price_constants_f32(C.cast(consts, C.c_void_p), n, C.byref(yc), out.ctypes.data,
                                         C.byref(cfg) if cfg is not None else None, None))
    return out


class Batch:
    """hwt_batch: a batch uploaded once and priced any number of times."""

    def __init__(self, opts: Options, y: Yield):
        self.N = opts.N
        self._keep = (opts, y)
        self._b = C.c_void_p()
        _check(lib().hwt_batch_create(C.byref(self._b), C.byref(_host_opts(opts)), C.byref(_host_yield(y))))

    def price(self, cfg: ConfigC | None = None, stats: bool = False):
        out = np.empty(self.N, dtype=np.float64)
        st = StatsC()
        _check(lib().hwt_batch_price(self._b, out.ctypes.data, C.byref(cfg) if cfg is not None else None,
                                     C.byref(st)))
        return (out, st.as_dict()) if stats else out

    @property
    def device_memory(self) -> int:
        return int(lib().hwt_batch_device_memory(self._b))

    def close(self):
        if self._b:
            lib().hwt_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def alphas(opts: Options, y: Yield) -> list[np.ndarray]:
    """Per-step alphas of each option (hwt_alphas_f64)."""
    offs = np.zeros(opts.N + 1, dtype=np.int64)
    ho, hy = _host_opts(opts), _host_yield(y)
    _check(lib().hwt_alphas_f64(C.byref(ho), C.byref(hy), offs.ctypes.data, None))
    buf = np.empty(int(offs[-1]), dtype=np.float64)
    _check(lib().hwt_alphas_f64(C.byref(ho), C.byref(hy), offs.ctypes.data, buf.ctypes.data))
    return [buf[offs[i]:offs[i + 1]] for i in range(opts.N)]


class Plan:
    """hwt_plan over device-resident inputs (see ``to_device``)."""

    FIELDS = ("strike", "maturity", "length", "term_unit", "term_steps", "a", "sigma", "type")

    def __init__(self, capacity: int, yield_capacity: int, cfg: ConfigC | None = None):
        self._p = C.c_void_p()
        _check(lib().hwt_plan_create(C.byref(self._p), int(capacity), int(yield_capacity),
                                     C.byref(cfg) if cfg is not None else None))

    def execute(self, dev: dict, stream=None, n: int | None = None):
        """dev: device buffers keyed by Options field names plus 'rates', 'days',
        'out' (anything with an integer ``.value`` device address)."""
        n = int(dev["n"] if n is None else n)
        o = OptionsC(n, *[dev[f].value for f in self.FIELDS])
        y = YieldC(int(dev["ny"]), dev["rates"].value, dev["days"].value)
        _check(lib().hwt_plan_execute(self._p, C.byref(o), C.byref(y), dev["out"].value,
                                      C.c_void_p(stream) if stream else None))

    def execute_status(self, dev: dict, n: int | None = None) -> int:
        n = int(dev["n"] if n is None else n)
        o = OptionsC(n, *[dev[f].value for f in self.FIELDS])
        y = YieldC(int(dev["ny"]), dev["rates"].value, dev["days"].value)
        return lib().hwt_plan_execute(self._p, C.byref(o), C.byref(y), dev["out"].value, None)

    def sync(self) -> dict:
        st = StatsC()
        _check(lib().hwt_plan_sync(self._p, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self._p:
            lib().hwt_plan_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(dev: dict, cfg: ConfigC | None = None):
    """hwt_autotune on a device batch (see ``to_device``): (tuned config, report)."""
    c = cfg if cfg is not None else config()
    n = int(dev["n"])
    o = OptionsC(n, *[dev[f].value for f in Plan.FIELDS])
    y = YieldC(int(dev["ny"]), dev["rates"].value, dev["days"].value)
    rep = TuneReportC()
    _check(lib().hwt_autotune(C.byref(o), C.byref(y), dev["out"].value, C.byref(c), C.byref(rep)))
    return c, rep.as_dict()


def to_device(opts: Options, y: Yield) -> dict:
    """Copies a batch and its yield curve to the device (DeviceArray buffers)."""
    from .devmem import DeviceArray
    dev = {f: DeviceArray.from_host(getattr(opts, f)) for f in Plan.FIELDS}
    dev["rates"] = DeviceArray.from_host(y.rates)
    dev["days"] = DeviceArray.from_host(y.days)
    dev["out"] = DeviceArray.empty((opts.N,), np.float64)
    dev["n"], dev["ny"] = opts.N, y.N
    return dev
