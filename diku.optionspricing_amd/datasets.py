"""Seeded synthetic workloads.

Field mapping (the reference's semantics, common/OptionConstants.hpp:33-35 and
seq/Seq.hpp:125): ``maturity`` is the underlying zero-coupon bond's maturity in
years and sets the tree height; ``length`` is the option's expiry in years and
sets the exercise step.  Workload descriptions that call the option's date
"maturity" and the bond's "length" are mapped onto these fields accordingly.

Workloads:
  uniform_1k     1,000 puts, 12 steps/yr, a=0.1, sigma=0.01, 10y bond, 5y option,
                 strike U[60,100]                        (width 47, height 120)
  skewed_60k     60,000 shuffled: 90% small (bond 1-5y, 12/yr, a U[0.1,0.2]),
                 10% large (bond 10-30y, 36/yr, a U[0.01,0.05]); sigma U[0.005,0.02],
                 strike U[60,100], call/put 1:1, option expiry U[0.2,0.8] x bond maturity
  box            surrogate training box: strike [60,100], option 1-30y, bond =
                 option + [1,10]y, 12 or 36 steps/yr, a [0.01,0.2], sigma [0.005,0.02]
  box_mix        95% inside the box, 5% outside (a in [0.2,0.4] or option 30-40y)
The yield curve for all of them has 120 monthly knots, zero rate rising
linearly from 1% to 4%.
"""
from __future__ import annotations

import numpy as np

from .formats import CALL, PUT, Options, Yield


def synthetic_yield() -> Yield:
    k = np.arange(1, 121)
    days = np.rint(365.0 * k / 12.0).astype(np.int32)
    rates = 0.01 + 0.03 * (k - 1) / 119.0
    return Yield(rates, days)


def _opts(strike, maturity, length, term_steps, a, sigma, typ) -> Options:
    n = len(strike)
    return Options(strike=strike, maturity=maturity, length=length,
                   term_unit=np.full(n, 365, np.uint16), term_steps=term_steps,
                   a=a, sigma=sigma, type=typ)


def uniform_1k(seed: int = 1, n: int = 1000) -> Options:
    rng = np.random.default_rng(seed)
    return _opts(rng.uniform(60, 100, n), np.full(n, 10.0), np.full(n, 5.0),
                 np.full(n, 12, np.uint16), np.full(n, 0.1), np.full(n, 0.01),
                 np.full(n, PUT, np.int8))


def skewed(seed: int = 2, n: int = 60000, large_frac: float = 0.1) -> Options:
    rng = np.random.default_rng(seed)
    nl = int(round(n * large_frac))
    ns = n - nl
    mat = np.concatenate([rng.uniform(1, 5, ns), rng.uniform(10, 30, nl)])
    steps = np.concatenate([np.full(ns, 12), np.full(nl, 36)]).astype(np.uint16)
    a = np.concatenate([rng.uniform(0.1, 0.2, ns), rng.uniform(0.01, 0.05, nl)])
    sigma = rng.uniform(0.005, 0.02, n)
    strike = rng.uniform(60, 100, n)
    typ = rng.integers(0, 2, n).astype(np.int8)
    length = rng.uniform(0.2, 0.8, n) * mat
    perm = rng.permutation(n)
    return _opts(strike[perm], mat[perm], length[perm], steps[perm], a[perm], sigma[perm], typ[perm])


def skewed_60k(seed: int = 2) -> Options:
    return skewed(seed, 60000)


BOX = {"strike": (60.0, 100.0), "length": (1.0, 30.0), "extra": (1.0, 10.0),
       "a": (0.01, 0.2), "sigma": (0.005, 0.02)}


def box(seed: int, n: int) -> Options:
    rng = np.random.default_rng(seed)
    length = rng.uniform(*BOX["length"], n)
    mat = length + rng.uniform(*BOX["extra"], n)
    steps = np.where(rng.integers(0, 2, n) == 0, 12, 36).astype(np.uint16)
    return _opts(rng.uniform(*BOX["strike"], n), mat, length, steps,
                 rng.uniform(*BOX["a"], n), rng.uniform(*BOX["sigma"], n),
                 rng.integers(0, 2, n).astype(np.int8))


def box_mix(seed: int, n: int, outside: float = 0.05) -> Options:
    rng = np.random.default_rng(seed)
    inside = box(seed + 7919, n)
    n_out = int(round(n * outside))
    idx = rng.choice(n, n_out, replace=False)
    half = n_out // 2
    a_out, t_out = idx[:half], idx[half:]
    inside.a[a_out] = rng.uniform(0.2, 0.4, half)
    new_len = rng.uniform(30.0, 40.0, len(t_out))
    inside.maturity[t_out] = new_len + (inside.maturity[t_out] - inside.length[t_out])
    inside.length[t_out] = new_len
    return inside


def book(steps=(1, 2, 3, 4, 5)) -> Options:
    """Hull's book example: 3-year put on a 9-year zero-coupon bond, strike 63,
    a = 0.1, sigma = 0.01 (test/TestCuda.cu:20-31)."""
    n = len(steps)
    return _opts(np.full(n, 63.0), np.full(n, 9.0), np.full(n, 3.0),
                 np.array(steps, np.uint16), np.full(n, 0.1), np.full(n, 0.01),
                 np.full(n, PUT, np.int8))
