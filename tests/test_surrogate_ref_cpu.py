"""Pins the CPU reference of the surrogate (no GPU): the backward pass by
finite differences, the closed-form baseline by put-call parity, its
low-volatility limit and a fine FP64 tree from the oracle, and the BF16
rounding helper by hand-computed cases."""
import math

import numpy as np
import pytest

import surrogate_ref as R


def _rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    flat = np.zeros(R.NPAR)
    for k, o in R.OFF.items():
        size = int(np.prod(R.SHAPE[k]))
        fan_in = R.SHAPE[k][-1] if len(R.SHAPE[k]) == 2 else 256
        flat[o:o + size] = rng.uniform(-1, 1, size) * scale / math.sqrt(fan_in)
    return flat


def test_exact_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    flat = _rand_params(1, scale=2.0)
    x = rng.normal(size=(5, 8))
    t = rng.normal(size=5)
    g = R.grad_exact(flat, x, t)
    for name, o in R.OFF.items():
        size = int(np.prod(R.SHAPE[name]))
        for j in rng.choice(size, size=min(size, 6), replace=False):
            i = o + int(j)
            h = 1e-6
            fp, fm = flat.copy(), flat.copy()
            fp[i] += h
            fm[i] -= h
            fd = (R.loss_exact(fp, x, t) - R.loss_exact(fm, x, t)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (name, j, fd, g[i])


def test_rounded_backward_close_to_exact_for_small_rounding():
    # with weights and inputs that BF16 represents exactly the rounded and the
    # exact networks differ only by the activation roundings
    rng = np.random.default_rng(5)
    flat = R.bf16(_rand_params(2).astype(np.float32)).astype(np.float64)
    x = rng.normal(size=(64, 8)).astype(np.float32)
    t = rng.normal(size=64)
    y, cache = R.forward_bf16(flat, x)
    ye, _ = R.forward_exact(flat, x)
    assert np.max(np.abs(y - ye)) < 2e-2 * max(1.0, np.max(np.abs(ye)))
    g, ge = R.backward(cache, y, t), R.grad_exact(flat, x, t)
    assert np.linalg.norm(g - ge) <= 3e-2 * np.linalg.norm(ge)


def test_bf16_rounding_cases():
    v = np.array([1.0, 1.00390625, 1.005859375, 1.001953125, 1.005859376, -3.0, 0.0], np.float32)
    # BF16 spacing at 1 is 2^-7: 1 + 2^-8 is a tie and goes to the even 1;
    # 1 + 1.5*2^-8 rounds up to 1 + 2^-7; 1 + 2^-9 rounds down
    want = np.array([1.0, 1.0, 1.0078125, 1.0, 1.0078125, -3.0, 0.0], np.float32)
    np.testing.assert_array_equal(R.bf16(v), want)


def _one(pkg, X, Tb, L, steps, a, sigma, typ):
    return pkg.Options(strike=[X], maturity=[Tb], length=[L], term_unit=[365], term_steps=[steps],
                       a=[a], sigma=[sigma], type=[typ])


def test_baseline_put_call_parity_and_box(pkg):
    y = pkg.datasets.synthetic_yield()
    o = pkg.datasets.box(4, 50)
    for i in range(o.N):
        o.type[i] = 1
        x, put, ok = R.inputs(o, i, y)
        o.type[i] = 0
        _, call, _ = R.inputs(o, i, y)
        assert ok and x.shape == (8,) and np.all(np.abs(x[:6]) <= 1.0 + 1e-12)
        dt = 1.0 / o.term_steps[i]
        T, S = int(o.length[i] / dt) * dt, int(o.term_steps[i] * o.maturity[i]) * dt
        PT = math.exp(-R._yield_at(T, y.rates, y.days) * T)
        PS = math.exp(-R._yield_at(S, y.rates, y.days) * S)
        assert abs((call - put) - (100 * PS - o.strike[i] * PT)) < 1e-9
    out = _one(pkg, 80.0, 12.0, 5.0, 12, 0.3, 0.01, 1)  # a outside the box
    assert R.inputs(out, 0, y)[2] is False


def test_baseline_low_volatility_limit(pkg):
    y = pkg.datasets.synthetic_yield()
    T, S = 3.0, 8.0
    call = R.closed_form(70.0, T, S, 0.1, 1e-7, 0, y)
    fwd = 100 * math.exp(-R._yield_at(S, y.rates, y.days) * S) - 70 * math.exp(-R._yield_at(T, y.rates, y.days) * T)
    assert abs(call - max(fwd, 0.0)) < 1e-8


def test_baseline_against_fine_tree(pkg, oracle):
    """The closed form is the continuous-time limit of the tree: a tree with
    200 steps a year lands within 5e-3 of it (per 100 notional); the price
    itself is O(1), so a wrong sign or volatility term fails."""
    y = pkg.datasets.synthetic_yield()
    for (X, Tb, L, a, s, typ) in [(85.0, 4.0, 2.0, 0.1, 0.01, 1), (90.0, 5.0, 2.0, 0.05, 0.015, 0),
                                  (80.0, 6.0, 3.0, 0.15, 0.02, 1)]:
        o = _one(pkg, X, Tb, L, 200, a, s, typ)
        tree = oracle.price(o, y)[0]
        # the formula on the same dates, restated here for steps outside the box
        dt = 1 / 200
        T, S = int(L / dt) * dt, int(200 * Tb) * dt
        cf = R.closed_form(X, T, S, a, s, typ, y)
        assert abs(tree - cf) < 5e-3, (tree, cf)


def test_threshold_calibration_rule(pkg):
    """The gate threshold rule on hand-made cases: the largest spread below
    the first option (in spread order) whose error exceeds the bound, with
    ties at the threshold all required to pass, out-of-box options ignored."""
    f = pkg.surrogate.threshold_from
    tree = np.zeros(6)
    spread = np.array([1e-5, 2e-5, 3e-5, 3e-5, 5e-5, 1.0])
    mean = np.array([1e-4, 2e-4, 3e-4, 9e-4, 1e-4, 0.0])
    inbox = np.array([True, True, True, True, True, False])
    # 4th option (spread 3e-5) fails a 5e-4 bound: its tie partner cannot be kept either
    assert f(tree, mean, spread, inbox, 5e-4) == (2e-5, 2)
    assert f(tree, mean, spread, inbox, 1e-3) == (5e-5, 5)
    # the first option already fails: nothing is accepted
    assert f(tree, mean, spread, inbox, 5e-5) == (0.0, 0)
    # an out-of-box option with a huge error is ignored
    mean2 = mean.copy()
    mean2[5] = 10.0
    assert f(tree, mean2, spread, inbox, 1e-3) == (5e-5, 5)
