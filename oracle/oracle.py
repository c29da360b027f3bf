"""ctypes front end of the CPU oracle (oracle/seq_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or CPU baseline, never by the
product.  The oracle restates seq/Seq.hpp + common/Domain.hpp +
common/OptionConstants.hpp in FP64 with the reference's operation order.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_SEQ = HERE / "_ref" / "Seq"
REF_SEQ32 = HERE / "_ref" / "Seq32"


class Constants(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("dr", C.c_double),
                ("X", C.c_double), ("M", C.c_double), ("jmax", C.c_int32),
                ("n", C.c_int32), ("width", C.c_int32), ("termUnit", C.c_uint16),
                ("type", C.c_int8)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "lib"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        d, i32, u16, i8 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint16), C.POINTER(C.c_int8)
        L.orc_option_constants.argtypes = [C.c_double, C.c_double, C.c_double, C.c_uint16, C.c_uint16,
                                           C.c_double, C.c_double, C.c_int8, C.POINTER(Constants)]
        L.orc_yield_at_year.argtypes = [C.c_double, C.c_int, d, i32, C.c_int]
        L.orc_yield_at_year.restype = C.c_double
        L.orc_prob.argtypes = [C.c_char, C.c_int, C.c_int, C.c_double]
        L.orc_prob.restype = C.c_double
        L.orc_compute_alpha.argtypes = [C.c_double, C.c_int, C.c_double, C.c_int, d, i32, C.c_int]
        L.orc_compute_alpha.restype = C.c_double
        L.orc_call_value.argtypes = [C.c_int, C.c_double, C.c_int8, C.c_double]
        L.orc_call_value.restype = C.c_double
        L.orc_price_single.argtypes = [C.POINTER(Constants), d, i32, C.c_int, d]
        L.orc_price_single.restype = C.c_double
        L.orc_option_constants_f32.argtypes = [C.c_float, C.c_float, C.c_float, C.c_uint16, C.c_uint16,
                                               C.c_float, C.c_float, C.c_int8, C.POINTER(Constants)]
        L.orc_price_single_f32step.argtypes = [C.POINTER(Constants), d, i32, C.c_int]
        L.orc_price_single_f32step.restype = C.c_double
        L.orc_price_batch.argtypes = [C.c_int, d, d, d, u16, u16, d, d, i8, d, i32, C.c_int, d, C.c_int]
        L.orc_node_updates.argtypes = [C.POINTER(Constants), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _p(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def constants(opts, i: int) -> Constants:
    c = Constants()
    lib().orc_option_constants(float(opts.strike[i]), float(opts.maturity[i]), float(opts.length[i]),
                               int(opts.term_unit[i]), int(opts.term_steps[i]), float(opts.a[i]),
                               float(opts.sigma[i]), int(opts.type[i]), C.byref(c))
    return c


def yield_at_year(t: float, term_unit: int, y) -> float:
    return lib().orc_yield_at_year(float(t), int(term_unit), _p(y.rates, C.c_double), _p(y.days, C.c_int32), y.N)


def prob(branch: str, which: int, j: int, M: float) -> float:
    return lib().orc_prob(branch.encode(), which, j, M)


def call_value(is_maturity: bool, X: float, typ: int, res: float) -> float:
    return lib().orc_call_value(int(bool(is_maturity)), X, typ, res)


def price_single(opts, i: int, y, with_alphas: bool = False):
    c = constants(opts, i)
    alphas = np.zeros(max(c.n, 0) + 1) if with_alphas else None
    v = lib().orc_price_single(C.byref(c), _p(y.rates, C.c_double), _p(y.days, C.c_int32), y.N,
                               _p(alphas, C.c_double) if with_alphas else None)
    return (v, alphas) if with_alphas else v


def price(opts, y, threads: int | None = None) -> np.ndarray:
    """seq::computeOptions over the whole batch (seq/Seq.hpp:181-190)."""
    out = np.empty(opts.N, dtype=np.float64)
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    lib().orc_price_batch(opts.N, _p(opts.strike, C.c_double), _p(opts.maturity, C.c_double),
                          _p(opts.length, C.c_double), _p(opts.term_unit, C.c_uint16),
                          _p(opts.term_steps, C.c_uint16), _p(opts.a, C.c_double),
                          _p(opts.sigma, C.c_double), _p(opts.type, C.c_int8),
                          _p(y.rates, C.c_double), _p(y.days, C.c_int32), y.N,
                          _p(out, C.c_double), int(threads))
    return out


def constants_f32(opts, i: int) -> Constants:
    """The tree constants a real = float build of the reference derives
    (common/OptionConstants.hpp:29-53 in float), widened."""
    c = Constants()
    f = np.float32
    lib().orc_option_constants_f32(f(opts.strike[i]), f(opts.maturity[i]), f(opts.length[i]),
                                   int(opts.term_unit[i]), int(opts.term_steps[i]), f(opts.a[i]),
                                   f(opts.sigma[i]), int(opts.type[i]), C.byref(c))
    return c


def price_f32shape(opts, y) -> np.ndarray:
    """The reference's FP32 build's trees (float constants, exercise step in
    float) evaluated by seq's algorithm in double, on the float-rounded
    curve: what an FP32 caller's prices are without FP32 rounding."""
    rates = np.ascontiguousarray(y.rates.astype(np.float32).astype(np.float64))
    days = np.ascontiguousarray(y.days, dtype=np.int32)
    out = np.empty(opts.N)
    for i in range(opts.N):
        c = constants_f32(opts, i)
        out[i] = lib().orc_price_single_f32step(C.byref(c), _p(rates, C.c_double), _p(days, C.c_int32), y.N)
    return out


def run_reference_seq32(options_path, yield_path) -> list[str] | None:
    """The reference's own seq built with its default real = float
    (oracle/_ref/Seq32); None when that binary was not built here."""
    if not REF_SEQ32.exists():
        return None
    out = subprocess.run([str(REF_SEQ32), "-o", str(options_path), "-y", str(yield_path)],
                         check=True, capture_output=True, text=True).stdout
    line = out.strip().splitlines()[0]
    return [t.strip() for t in line.strip()[1:-1].split(",")]


def node_updates(opts, i: int):
    c = constants(opts, i)
    f, b = C.c_int64(), C.c_int64()
    lib().orc_node_updates(C.byref(c), C.byref(f), C.byref(b))
    return f.value, b.value


def run_reference_seq(options_path, yield_path) -> list[str] | None:
    """Run the reference's own seq (REAL=64) built into oracle/_ref; returns the
    result tokens, or None when that binary was not built here."""
    if not REF_SEQ.exists():
        return None
    out = subprocess.run([str(REF_SEQ), "-o", str(options_path), "-y", str(yield_path)],
                         check=True, capture_output=True, text=True,
                         env={**os.environ, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS", "8")}).stdout
    line = out.strip().splitlines()[0]
    return [t.strip() for t in line.strip()[1:-1].split(",")]
