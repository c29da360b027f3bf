"""Text I/O in the reference's array format.

An options file is 8 lines, each ``[v0, v1, ...]``, in the field order of
``trinom::Options`` (common/Options.hpp:87-94): strike, maturity, length,
term unit, term step count, reversion rate, volatility, type.  A yield file is
2 lines: rates, then day offsets (common/Yield.hpp:34-35).  Results are one
line written with ``std::fixed`` and ``max_digits10`` decimals: 17 for double,
9 for float (common/Arrays.hpp:36-46).
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

CALL = 0
PUT = 1

# (name, dtype) in file order, common/Options.hpp:48-58
OPTION_FIELDS = (
    ("strike", np.float64),
    ("maturity", np.float64),
    ("length", np.float64),
    ("term_unit", np.uint16),
    ("term_steps", np.uint16),
    ("a", np.float64),
    ("sigma", np.float64),
    ("type", np.int8),
)


@dataclass
class Options:
    """Structure-of-arrays batch, field for field ``trinom::Options``."""

    strike: np.ndarray
    maturity: np.ndarray
    length: np.ndarray
    term_unit: np.ndarray
    term_steps: np.ndarray
    a: np.ndarray
    sigma: np.ndarray
    type: np.ndarray

    def __post_init__(self):
        n = None
        for name, dt in OPTION_FIELDS:
            arr = np.ascontiguousarray(np.asarray(getattr(self, name)), dtype=dt)
            setattr(self, name, arr)
            if n is None:
                n = arr.shape[0]
            elif arr.shape != (n,):
                raise ValueError(f"field {name} has shape {arr.shape}, expected ({n},)")
        bad = (self.type != CALL) & (self.type != PUT)
        if bad.any():
            # common/Options.hpp:41-44 throws std::out_of_range on a bad type
            raise ValueError("Invalid OptionType read from stream.")

    @property
    def N(self) -> int:
        return int(self.strike.shape[0])

    def take(self, idx) -> "Options":
        return Options(**{name: getattr(self, name)[idx] for name, _ in OPTION_FIELDS})

    @staticmethod
    def concat(parts) -> "Options":
        return Options(**{name: np.concatenate([getattr(p, name) for p in parts])
                          for name, _ in OPTION_FIELDS})


@dataclass
class Yield:
    """``trinom::Yield``: continuously compounded zero rates at day offsets."""

    rates: np.ndarray
    days: np.ndarray

    def __post_init__(self):
        self.rates = np.ascontiguousarray(self.rates, dtype=np.float64)
        self.days = np.ascontiguousarray(self.days, dtype=np.int32)
        if self.rates.shape != self.days.shape or self.rates.ndim != 1 or self.rates.size == 0:
            raise ValueError("yield curve needs equally long, non-empty rate and day arrays")

    @property
    def N(self) -> int:
        return int(self.rates.shape[0])


def _parse_line(line: str):
    line = line.strip()
    if not (line.startswith("[") and line.endswith("]")):
        raise ValueError(f"not an array line: {line[:40]!r}")
    body = line[1:-1].strip()
    return [t.strip() for t in body.split(",")] if body else []


def read_arrays(path) -> list[list[str]]:
    text = Path(path).read_text()
    return [_parse_line(l) for l in text.splitlines() if l.strip()]


def read_options(path) -> Options:
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"File '{path}' does not exist.")
    lines = read_arrays(path)
    if len(lines) < 8:
        raise ValueError(f"{path}: expected 8 array lines, found {len(lines)}")
    fields = {}
    for (name, dt), toks in zip(OPTION_FIELDS, lines[:8]):
        if np.issubdtype(dt, np.integer):
            fields[name] = np.array([int(t) for t in toks], dtype=dt)
        else:
            fields[name] = np.array([float(t) for t in toks], dtype=dt)
    return Options(**fields)


def read_yield(path) -> Yield:
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError("File does not exist.")
    lines = read_arrays(path)
    return Yield(np.array([float(t) for t in lines[0]]), np.array([int(t) for t in lines[1]]))


def read_results(path) -> np.ndarray:
    return np.array([float(t) for t in read_arrays(path)[0]], dtype=np.float64)


def format_array(values, decimals: int = 17) -> str:
    """One ``[a, b, ...]`` line, as ``Arrays::write_array`` prints reals."""
    return "[" + ", ".join(f"{float(v):.{decimals}f}" for v in values) + "]"


def write_options(path, opts: Options) -> None:
    out = []
    for name, dt in OPTION_FIELDS:
        arr = getattr(opts, name)
        if np.issubdtype(dt, np.integer):
            out.append("[" + ", ".join(str(int(v)) for v in arr) + "]")
        else:
            out.append(format_array(arr))
    Path(path).write_text("\n".join(out) + "\n")


def write_yield(path, y: Yield) -> None:
    Path(path).write_text(format_array(y.rates) + "\n" +
                          "[" + ", ".join(str(int(d)) for d in y.days) + "]\n")
