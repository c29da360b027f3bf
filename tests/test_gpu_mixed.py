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
    mixed, st = pkg.capi.price(opts, y, _mixed(pkg), stats=True)
    fp64 = pkg.capi.price(opts, y)
    assert np.all(np.isfinite(mixed))
    # the range guard leaves the benchmark batch to the FP32 kernels
    assert st["fp64_repriced"] <= opts.N // 1000, st["fp64_repriced"]
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


def _long_bonds(pkg):
    """Long bonds on weakly mean-reverting trees, where the undiscounted FP32
    backward state overflows inside the light cone of the exercise nodes
    (1e40 to 1e80 at the low-rate edge in FP64), followed by ordinary options
    of the same shapes with short bonds (no overflow)."""
    rows = [  # (bond years, expiry years, steps/yr, a, sigma, type)
        (100.0, 1.0, 12, 0.01, 0.02, 1),    # w = 443
        (100.0, 3.0, 12, 0.005, 0.01, 1),   # w = 885
        (150.0, 3.0, 12, 0.005, 0.01, 0),   # 1,800 steps
        (100.0, 3.0, 12, 0.009, 0.01, 1),   # reference 1_RAND-like row (w = 493)
        (97.0, 3.0, 12, 0.0095, 0.01, 1),
        (100.0, 1.0, 12, 0.009, 0.01, 0),
        (30.0, 1.0, 36, 0.01, 0.02, 1),     # the skewed batch's most extreme corner
        (10.0, 3.0, 12, 0.01, 0.02, 1),     # short bonds: FP32 is enough
        (9.0, 3.0, 12, 0.005, 0.01, 0),
    ]
    T, L, st, a, sg, ty = (np.array(c) for c in zip(*rows))
    n = len(rows)
    return pkg.Options(strike=np.full(n, 63.0), maturity=T, length=L, term_unit=np.full(n, 365),
                       term_steps=st, a=a, sigma=sg, type=ty)


def test_mixed_reprices_fp32_overflow_in_fp64(pkg, oracle, golden_yield):
    """The range guard: options whose FP32 state overflows come back finite,
    within the mixed bar of the FP64 oracle, and are counted as re-priced;
    the same options in FP64 mode need no guard."""
    for y in (golden_yield, pkg.datasets.synthetic_yield()):
        opts = _long_bonds(pkg)
        got, st = pkg.capi.price(opts, y, _mixed(pkg), stats=True)
        ref = oracle.price(opts, y)
        err = np.abs(got - ref)
        print(f"\nlong bonds, mixed: max abs err {err.max():.3e}, re-priced in FP64: {st['fp64_repriced']}")
        assert np.all(np.isfinite(got))
        assert err.max() <= ABS_BAR
        assert 3 <= st["fp64_repriced"] <= opts.N - 2
        fp64, st64 = pkg.capi.price(opts, y, stats=True)
        assert st64["fp64_repriced"] == 0
        from conftest import parity_tol
        assert np.all(np.abs(fp64 - ref) <= parity_tol(ref))


def test_mixed_guard_in_every_kernel_arrangement(pkg, oracle, golden_yield):
    """The long bonds mixed into a batch of narrow trees, through the serial,
    side-by-side and fused arrangements and one tree per wavefront."""
    y = golden_yield
    opts = pkg.Options.concat([_long_bonds(pkg), pkg.formats.read_options(
        __import__("conftest").GOLDEN / "2_RANDCONSTHEIGHT_1000.in")])
    ref = oracle.price(opts, y)
    C = pkg.capi
    for flags in (0, C.FLAG_CONCURRENT, C.FLAG_FUSED, C.FLAG_NO_ROWS, C.FLAG_NO_GRAPH):
        got, st = C.price(opts, y, C.config(precision=C.PREC_MIXED, flags=flags), stats=True)
        assert np.all(np.isfinite(got)), flags
        assert np.abs(got - ref).max() <= ABS_BAR, flags
        assert st["fp64_repriced"] >= 3, flags


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
