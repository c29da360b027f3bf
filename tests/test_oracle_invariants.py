"""Pins each oracle function with facts the mathematics fixes, independent of
the oracle's own formulas:

* tree width vs the reference's reversion-rate -> width table
  (data-generator/RandomDataGenerator.cpp:54-55), and the height rule;
* yield interpolation hand cases (before / on / between / after knots);
* branching probabilities: rows sum to one and match the first two moments of
  the Hull-White increment (mean j*M, second moment 1/3 + j^2 M^2, in units
  of dr) for the standard, bottom and top branchings;
* the zero-coupon-bond identity and put-call parity on the tree;
* per-step alphas reprice the yield curve;
* an independent high-precision brute force on tiny trees (jmax 1..3) that
  derives the probabilities by solving the moment equations and calibrates
  every alpha by root finding.
"""
import json

import mpmath as mp
import numpy as np
import pytest

from conftest import GOLDEN


def test_width_table(pkg, oracle):
    t = json.loads((GOLDEN / "width_table.json").read_text())
    n = len(t["a"])
    opts = pkg.Options(strike=np.full(n, 63.0), maturity=np.full(n, 9.0), length=np.full(n, 3.0),
                       term_unit=np.full(n, t["term_unit"]), term_steps=np.full(n, t["term_steps"]),
                       a=t["a"], sigma=np.full(n, 0.01), type=np.full(n, 1))
    widths = [oracle.constants(opts, i).width for i in range(n)]
    assert widths == t["width"]


@pytest.mark.parametrize("steps,T,expect_n", [(12, 9.0, 108), (36, 30.0, 1080), (12, 2.5, 30), (1, 9.0, 9), (365, 9.0, 3285)])
def test_height_and_step(pkg, oracle, steps, T, expect_n):
    o = pkg.Options(strike=[63.0], maturity=[T], length=[3.0], term_unit=[365], term_steps=[steps],
                    a=[0.1], sigma=[0.01], type=[1])
    c = oracle.constants(o, 0)
    assert c.n == expect_n
    assert c.dt == pytest.approx(1.0 / steps, rel=1e-15)
    # jmax is the smallest integer above 0.184 / (1 - e^{-a dt})
    q = 0.184 / (1.0 - np.exp(-0.1 * c.dt))
    assert c.jmax == int(q) + 1 and c.width == 2 * c.jmax + 1


def test_yield_hand_cases(pkg, oracle):
    y = pkg.Yield(np.array([0.01, 0.02, 0.04]), np.array([100, 200, 400]))
    f = lambda t: oracle.yield_at_year(t, 365, y)
    assert f(10 / 365) == 0.01                 # before the first knot: flat
    assert f(100 / 365) == 0.01                # on the first knot
    assert f(200 / 365) == 0.02                # on an inner knot
    assert f(150 / 365) == pytest.approx(0.015, abs=1e-17)   # between knots
    assert f(300 / 365) == pytest.approx(0.03, abs=1e-17)
    assert f(500 / 365) == 0.04                # after the last knot: flat
    assert f(399.6 / 365) == 0.04              # rounds to day 400
    assert f(399.4 / 365) == pytest.approx(0.02 + (399 - 200) / 200 * 0.02, abs=1e-16)


@pytest.mark.parametrize("M", [-0.0083, -0.02, -0.08, -0.18, -0.25])
def test_probability_rows_and_moments(oracle, M):
    jmax = int(-0.184 / M) + 1
    for j in range(-jmax + 1, jmax):
        pu, pm, pd = (oracle.prob("A", w, j, M) for w in (1, 2, 3))
        assert pu + pm + pd == pytest.approx(1.0, abs=1e-15)
        assert pu - pd == pytest.approx(j * M, abs=1e-15)
        assert pu + pd == pytest.approx(1 / 3 + j * j * M * M, abs=1e-15)
        assert min(pu, pm, pd) > 0
    j = -jmax  # bottom: moves +2, +1, 0
    pu, pm, pd = (oracle.prob("B", w, j, M) for w in (1, 2, 3))
    assert pu + pm + pd == pytest.approx(1.0, abs=1e-14)
    assert 2 * pu + pm == pytest.approx(j * M, abs=1e-14)
    assert 4 * pu + pm == pytest.approx(1 / 3 + j * j * M * M, abs=1e-14)
    j = jmax  # top: moves 0, -1, -2
    pu, pm, pd = (oracle.prob("C", w, j, M) for w in (1, 2, 3))
    assert pu + pm + pd == pytest.approx(1.0, abs=1e-14)
    assert -pm - 2 * pd == pytest.approx(j * M, abs=1e-14)
    assert pm + 4 * pd == pytest.approx(1 / 3 + j * j * M * M, abs=1e-14)


def test_call_value_clamps(oracle):
    assert oracle.call_value(True, 63.0, 1, 60.0) == 3.0          # put
    assert oracle.call_value(True, 63.0, 1, 70.0) == 0.0
    assert oracle.call_value(True, 63.0, 0, 70.0) == 7.0          # call
    assert oracle.call_value(False, 63.0, 0, 70.0) == 70.0        # not the exercise step
    assert oracle.call_value(False, 63.0, 0, float("inf")) == np.finfo(float).max
    assert oracle.call_value(False, 63.0, 0, float("-inf")) == np.finfo(float).tiny


def _discount(oracle, y, t, tu=365):
    return np.exp(-oracle.yield_at_year(t, tu, y) * t)


def test_zero_coupon_bond_and_put_call_parity(pkg, oracle):
    y = pkg.datasets.synthetic_yield()
    rng = np.random.default_rng(11)
    n = 24
    mat = rng.uniform(1, 8, n)
    base = dict(maturity=mat, term_unit=np.full(n, 365), term_steps=rng.choice([4, 12, 36], n),
                a=rng.uniform(0.02, 0.3, n), sigma=rng.uniform(0.005, 0.02, n), strike=rng.uniform(60, 100, n))
    put = pkg.Options(length=mat * rng.uniform(0.1, 0.9, n), type=np.full(n, 1), **base)
    call = pkg.Options(length=put.length, type=np.full(n, 0), **base)
    zcb = pkg.Options(length=mat + 1.0, type=np.full(n, 1), **base)
    pp, pc, pz = oracle.price(put, y), oracle.price(call, y), oracle.price(zcb, y)
    for i in range(n):
        c = oracle.constants(put, i)
        e = int(c.t / c.dt)
        Pn = _discount(oracle, y, c.n * c.dt)
        Pe = _discount(oracle, y, e * c.dt) if e > 0 else 1.0
        assert pz[i] == pytest.approx(100 * Pn, rel=1e-12)
        assert pc[i] - pp[i] == pytest.approx(100 * Pn - c.X * Pe, abs=1e-10)


def test_alphas_reprice_the_curve(pkg, oracle):
    """With alpha calibrated, the tree's one-step discount of the root is
    P(dt) and alphas[0] is the dt-period zero rate."""
    y = pkg.datasets.synthetic_yield()
    o = pkg.datasets.book([12])
    price, al = oracle.price_single(o, 0, y, with_alphas=True)
    c = oracle.constants(o, 0)
    assert al[0] == oracle.yield_at_year(c.dt, 365, y)
    assert len(al) == c.n + 1 and np.all(np.isfinite(al))
    # alpha tracks the instantaneous forward rate, which for a zero curve
    # rising 1% -> 4% over ten years stays between 1% and R + t R' < 8%
    assert 0.009 < al.min() and al.max() < 0.08


# ---------------------------------------------------------------- brute force
def _brute_force_price(strike, T, L, steps, a, sigma, typ, y, dps=40):
    """Hull-White trinomial tree from first principles in high precision:
    probabilities solve the moment equations, alpha is root-found so the tree
    reprices the discount curve at every step."""
    mp.mp.dps = dps
    tuy = int(np.ceil(365.0 / 365))
    n = int(steps * tuy * T)
    dt_f = tuy / steps
    dt = mp.mpf(tuy) / steps
    a_m, s_m = mp.mpf(a), mp.mpf(sigma)
    V = s_m ** 2 * (1 - mp.e ** (-2 * a_m * dt)) / (2 * a_m)
    dr = mp.sqrt(3 * V)
    M = mp.e ** (-a_m * dt) - 1
    jmax = int(-0.184 / (np.exp(-a * dt_f) - 1.0)) + 1

    def probs(j):
        shifts = (1, 0, -1) if -jmax < j < jmax else ((2, 1, 0) if j == -jmax else (0, -1, -2))
        A = mp.matrix([[1, 1, 1], [shifts[0], shifts[1], shifts[2]],
                       [shifts[0] ** 2, shifts[1] ** 2, shifts[2] ** 2]])
        b = mp.matrix([1, j * M, mp.mpf(1) / 3 + (j * M) ** 2])
        p = mp.lu_solve(A, b)
        return [(j + s, p[k]) for k, s in enumerate(shifts)]

    def P(t_f):
        days = int(round(t_f * 365))
        R = mp.mpf(float(np.interp(days, y.days, y.rates)))
        return mp.e ** (-R * mp.mpf(t_f))

    Q = {0: mp.mpf(1)}
    alphas = []
    for i in range(n):
        target = P((i + 1) * dt_f)
        f = lambda al: sum(q * mp.e ** (-(al + j * dr) * dt) for j, q in Q.items()) - target
        al = mp.findroot(f, mp.mpf("0.03"))
        alphas.append(al)
        nq = {}
        for j, q in Q.items():
            disc = q * mp.e ** (-(al + j * dr) * dt)
            for k, p in probs(j):
                nq[k] = nq.get(k, 0) + p * disc
        Q = nq
    e = int(L / dt_f)
    Vn = {j: mp.mpf(100) for j in range(-jmax, jmax + 1)}
    for i in range(n - 1, -1, -1):
        hi = min(i, jmax)
        nv = {}
        for j in range(-hi, hi + 1):
            v = sum(p * Vn[k] for k, p in probs(j)) * mp.e ** (-(alphas[i] + j * dr) * dt)
            if i == e:
                v = max(strike - v, 0) if typ == 1 else max(v - strike, 0)
            nv[j] = v
        Vn = nv
    return float(Vn[0]), jmax


@pytest.mark.parametrize("a,steps,T,L,typ", [
    (3.0, 1, 6.0, 2.0, 1),    # jmax 1
    (2.5, 1, 7.0, 3.0, 0),    # jmax 1
    (0.15, 1, 8.0, 3.0, 1),   # jmax 2
    (0.08, 1, 9.0, 4.0, 1),   # jmax 3
    (0.08, 1, 9.0, 0.5, 0),   # exercise at step 0
    (0.30, 2, 5.0, 4.6, 1),   # exercise at the last step
])
def test_brute_force_tiny_trees(pkg, oracle, a, steps, T, L, typ):
    y = pkg.datasets.synthetic_yield()
    strike = 85.0 if typ == 1 else 80.0
    o = pkg.Options(strike=[strike], maturity=[T], length=[L], term_unit=[365], term_steps=[steps],
                    a=[a], sigma=[0.03], type=[typ])
    ref, jmax = _brute_force_price(strike, T, L, steps, a, 0.03, typ, y)
    assert oracle.constants(o, 0).jmax == jmax
    assert oracle.price_single(o, 0, y) == pytest.approx(ref, rel=1e-12, abs=1e-12)
