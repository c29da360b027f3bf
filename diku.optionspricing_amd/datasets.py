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
  tiny           trees 3-21 nodes wide (a from 3.0 down to 0.09 at 12 steps/yr),
                 bond 1-30y, option U[0.2,0.8] x bond, strike U[60,100], call/put
The yield curve for all of them has 120 monthly knots, zero rate rising
linearly from 1% to 4%.

reference_set(kind, n) draws the reference's seven benchmark distributions
(data-generator/RandomDataGenerator.cpp:114-257): 0_UNIFORM, 1_RAND,
2_RANDCONSTHEIGHT, 3_RANDCONSTWIDTH, 4_SKEWED, 5_SKEWEDCONSTHEIGHT and
6_SKEWEDCONSTWIDTH, with the reference's widths, heights, skew fraction and
fixed fields, from this module's own seeded generator (the reference's C
rand() stream is not reproduced, so the rows differ; the distributions do
not).  They price on the reference's 15-knot curve (tests/golden/yield.in).
"""
from __future__ import annotations

import functools

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


def tiny(seed: int, n: int) -> Options:
    """Options on trees 3 to 21 nodes wide (widths uniform over the odd
    values), the flattened form's class."""
    rng = np.random.default_rng(seed)
    widths = rng.choice(np.arange(3, 22, 2), n)
    a = np.array([3.0 if w == 3 else reference_rate_for_width(int(w)) for w in widths])
    mat = rng.uniform(1.0, 30.0, n)
    return _opts(rng.uniform(60, 100, n), mat, rng.uniform(0.2, 0.8, n) * mat, np.full(n, 12, np.uint16), a,
                 rng.uniform(0.005, 0.02, n), rng.integers(0, 2, n).astype(np.int8))


def book(steps=(1, 2, 3, 4, 5)) -> Options:
    """Hull's book example: 3-year put on a 9-year zero-coupon bond, strike 63,
    a = 0.1, sigma = 0.01 (test/TestCuda.cu:20-31)."""
    n = len(steps)
    return _opts(np.full(n, 63.0), np.full(n, 9.0), np.full(n, 3.0),
                 np.array(steps, np.uint16), np.full(n, 0.1), np.full(n, 0.01),
                 np.full(n, PUT, np.int8))


# ------------------------------------------------------------------ reference sets
REFERENCE_SETS = ("0_UNIFORM", "1_RAND", "2_RANDCONSTHEIGHT", "3_RANDCONSTWIDTH", "4_SKEWED",
                  "5_SKEWEDCONSTHEIGHT", "6_SKEWEDCONSTWIDTH")


@functools.lru_cache(maxsize=None)
def reference_rate_for_width(width: int, steps: int = 12):
    """The reversion rate the reference's generator uses for a tree `width`
    nodes wide at `steps` per year: the smallest rate with five decimals that
    gives that width (jmax = (int)(0.184 / (1 - e^{-a dt})) + 1,
    common/OptionConstants.hpp:44-51).  That is the reference's own table
    (data-generator/RandomDataGenerator.cpp:54-55), entry for entry; None for
    the widths no five-decimal rate gives (957, 989, 1007 and 1021 at 12 steps),
    which the table lacks."""
    J = (int(width) - 1) // 2
    if width % 2 == 0 or J < 2:
        return None
    dt = 1.0 / steps
    lo = -np.log1p(-0.184 / J) / dt  # rates above lo have jmax <= J
    a = np.ceil(lo * 1e5 - 1e-6) / 1e5
    for cand in (a, a + 1e-5):
        M = np.exp(-cand * dt) - 1.0
        if 2 * (int(-0.184 / M) + 1) + 1 == width:
            return float(round(cand, 5))
    return None


def _rate_by_width_range(rng, lo: int, hi: int) -> float:
    """getRRByWidthRange (RandomDataGenerator.cpp:57-71): a width drawn
    uniformly from [lo, hi] until one has a rate, then that rate."""
    while True:
        w = int(rng.integers(lo, hi + 1))
        a = reference_rate_for_width(w)
        if a is not None:
            return a


def reference_set(kind, n: int = 65536, seed: int = 7, skew_percent: int = 1) -> Options:
    """One of the reference's seven benchmark distributions
    (RandomDataGenerator.cpp:114-257), shuffled.  Every option is a 3-year put
    (Lengths 3) on a bond of integer maturity, strike 63, term unit 365,
    12 steps a year, sigma 0.01; the distributions differ in maturity (height
    12 T + 1 levels) and width:
      0_UNIFORM            T = 9, a = 0.1 (width 47)
      1_RAND               T ~ U{1..100}, width ~ U[7, 511]
      2_RANDCONSTHEIGHT    T = 9, width ~ U[7, 511]
      3_RANDCONSTWIDTH     T ~ U{1..100}, a = 0.1
      4_SKEWED             skew_percent of the options T ~ U{70..100}, width ~ U[411, 511];
                           the rest T ~ U{1..30}, width ~ U[7, 107]
      5_SKEWEDCONSTHEIGHT  skew_percent T = 100, width ~ U[7, 107]; the rest as in 4
      6_SKEWEDCONSTWIDTH   skew_percent T ~ U{1..30} at one width drawn from U[509, 511];
                           the rest as in 4
    The skewed part holds ceil(skew_percent/100 * n) options, as the
    reference's loops count them."""
    k = REFERENCE_SETS.index(kind) if isinstance(kind, str) else int(kind)
    rng = np.random.default_rng(seed)
    nskew = int(np.ceil(skew_percent / 100.0 * n)) if k >= 4 else 0
    mat = np.empty(n)
    a = np.empty(n)
    for i in range(n):
        if k == 0:
            mat[i], a[i] = 9, 0.1
        elif k == 1:
            mat[i], a[i] = rng.integers(1, 101), _rate_by_width_range(rng, 7, 511)
        elif k == 2:
            mat[i], a[i] = 9, _rate_by_width_range(rng, 7, 511)
        elif k == 3:
            mat[i], a[i] = rng.integers(1, 101), 0.1
        elif i >= nskew:
            mat[i], a[i] = rng.integers(1, 31), _rate_by_width_range(rng, 7, 107)
        elif k == 4:
            mat[i], a[i] = rng.integers(70, 101), _rate_by_width_range(rng, 411, 511)
        elif k == 5:
            mat[i], a[i] = 100, _rate_by_width_range(rng, 7, 107)
        else:
            if i == 0:
                skew_rate = _rate_by_width_range(rng, 509, 511)
            mat[i], a[i] = rng.integers(1, 31), skew_rate
    perm = rng.permutation(n)
    return _opts(np.full(n, 63.0), mat[perm], np.full(n, 3.0), np.full(n, 12, np.uint16), a[perm],
                 np.full(n, 0.01), np.full(n, PUT, np.int8))
