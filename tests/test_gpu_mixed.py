"""Mixed precision: both tree passes in FP32, discount factors, edge tables
and the duality sums in FP64 (compensated).

The mixed pricer is judged against the FP64 oracle (on a stratified subset)
and against the FP64 GPU pricer (whole batch, which itself matches the oracle
to 1e-10).  Bar: max abs error <= 1e-3 per 100 notional; the measured max and
mean errors are printed.  Exercise indices and tree heights are computed in
FP64 exactly as in FP64 mode, so ZCB-only options (exercise outside the tree)
agree to FP64 rounding.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ABS_BAR = 1e-3  # per 100 notional


def _mixed(pkg):
    return pkg.capi.config(precision=pkg.capi.PREC_MIXED)


def test_mixed_skewed_60k_error_report(pkg, oracle):
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed_60k()
    mixed = pkg.capi.price(opts, y, _mixed(pkg))
    fp64 = pkg.capi.price(opts, y)
    assert np.all(np.isfinite(mixed))
    err = np.abs(mixed - fp64)
    sub = np.arange(0, opts.N, 40)
    err_oracle = np.abs(mixed[sub] - oracle.price(opts.take(sub), y))
    big = opts.term_steps == 36
    print(f"\nmixed vs FP64 GPU, 60k skewed: max {err.max():.3e} mean {err.mean():.3e}; "
          f"large trees max {err[big].max():.3e}, small trees max {err[~big].max():.3e}; "
          f"vs oracle (every 40th): max {err_oracle.max():.3e} mean {err_oracle.mean():.3e}")
    assert err.max() <= ABS_BAR
    assert err_oracle.max() <= ABS_BAR
    # FP32 propagation genuinely differs from FP64 (the mode is not a no-op)
    assert err.max() > 1e-12


def test_mixed_golden_sets(pkg, golden_yield):
    from conftest import GOLDEN, GOLDEN_SETS
    for name in GOLDEN_SETS:
        opts = pkg.formats.read_options(GOLDEN / f"{name}.in")
        gold = pkg.formats.read_results(GOLDEN / f"{name}.out")
        got = pkg.capi.price(opts, golden_yield, _mixed(pkg))
        err = np.abs(got - gold)
        assert err.max() <= ABS_BAR, f"{name}: max abs err {err.max():.3e}"


def test_mixed_book_and_zcb_rows(pkg, oracle, golden_yield):
    opts = pkg.datasets.book(list(range(1, 101)))
    got = pkg.capi.price(opts, golden_yield, _mixed(pkg))
    ref = oracle.price(opts, golden_yield)
    assert np.abs(got - ref).max() <= ABS_BAR
    # exercise beyond the tree: the price is the FP64 bond formula in both modes
    o = pkg.Options(strike=[80.0, 80.0], maturity=[9.0, 9.0], length=[9.0, 12.0], term_unit=[365, 365],
                    term_steps=[12, 12], a=[0.1, 0.1], sigma=[0.01, 0.01], type=[pkg.PUT, pkg.CALL])
    np.testing.assert_allclose(pkg.capi.price(o, golden_yield, _mixed(pkg)), oracle.price(o, golden_yield),
                               rtol=1e-12, atol=1e-12)


def test_mixed_put_call_parity(pkg):
    """C - P = 100 P(n dt) - X P(e dt) holds in mixed mode to FP32 accuracy."""
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed(seed=11, n=4000)
    put = opts.take(np.arange(opts.N))
    put.type[:] = pkg.PUT
    call = opts.take(np.arange(opts.N))
    call.type[:] = pkg.CALL
    p64, c64 = pkg.capi.price(put, y), pkg.capi.price(call, y)
    p32, c32 = pkg.capi.price(put, y, _mixed(pkg)), pkg.capi.price(call, y, _mixed(pkg))
    np.testing.assert_allclose(c32 - p32, c64 - p64, rtol=0, atol=2 * ABS_BAR)


def test_mixed_rejected_for_lane_form(pkg):
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.uniform_1k()
    rc, _ = pkg.capi.price_status(opts, y, pkg.capi.config(form=pkg.capi.FORM_LANE,
                                                          precision=pkg.capi.PREC_MIXED))
    assert rc == pkg.capi.HWT_EUNSUPPORTED


def test_mixed_deterministic(pkg):
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed(seed=3, n=2000)
    a = pkg.capi.price(opts, y, _mixed(pkg))
    b = pkg.capi.price(opts, y, _mixed(pkg))
    assert np.array_equal(a, b)
