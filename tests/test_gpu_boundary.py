"""The boundary's other entry points on the GPU: FP32 callers
(hwt_price_f32), pricing from tree constants (hwt_price_constants_f64, behind
seq::computeSingleOption), and a batch uploaded once and priced many times
(hwt_batch_*, behind cuda::CudaOptions).

Bars: FP32 prices against the reference's FP32 seq output (data/out32) at the
reference's own FP32 comparison (test/Compare.hpp:8-16: Catch's Approx,
margin 1000 * FLT_EPSILON or 100 * FLT_EPSILON relative) and against the
oracle on the widened inputs to FP32 rounding; constants and batch paths
against the oracle at the FP64 bar (1e-10 relative, 1e-10 floor).
"""
import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_SETS, parity_tol
from test_drop_in_cpp import catch_approx_ok

pytestmark = pytest.mark.gpu


def _close_f32(got, ref):
    """float32 prices against a double reference of the same trees: one
    rounding to float apart (two ulps), with the FP64 bar's floor for tiny
    prices."""
    ref32 = ref.astype(np.float32)
    return np.abs(got - ref32) <= np.maximum(2 * np.spacing(np.abs(ref32)), 1e-10)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", GOLDEN_SETS)
def test_f32_entry_matches_fp32_goldens(pkg, oracle, golden_yield, name, precision):
    opts = pkg.formats.read_options(GOLDEN / f"{name}.in")
    gold32 = pkg.formats.read_results(GOLDEN / f"{name}.out32")
    cfg = pkg.capi.config(precision=pkg.capi.PREC_MIXED if precision == "mixed" else pkg.capi.PREC_FP64)
    got = pkg.capi.price_f32(opts, golden_yield, cfg)
    assert got.dtype == np.float32
    ok = catch_approx_ok(got.astype(np.float64), gold32)
    assert ok.all(), (name, np.nonzero(~ok)[0][:5], got[~ok][:5], gold32[~ok][:5])
    if precision == "fp64":
        # the FP32 build's trees, priced in FP64 and rounded once
        ref = oracle.price_f32shape(opts, golden_yield)
        assert _close_f32(got, ref).all(), name


def test_f32_entry_takes_the_exercise_step_in_float(pkg, oracle, golden_yield):
    """Step counts where the FP32 reference exercises one step earlier than
    FP64 (t / dt just below an integer in float): the FP32 entry prices the
    FP32 reference's tree, not the FP64 one."""
    steps = [15, 30, 60, 63, 75, 77, 12, 3]
    opts = pkg.datasets.book(steps)
    got = pkg.capi.price_f32(opts, golden_yield)
    assert _close_f32(got, oracle.price_f32shape(opts, golden_yield)).all()
    fp64 = oracle.price(opts, golden_yield)
    assert np.all(np.abs(got[:6] - fp64[:6]) > 1e-3)  # another tree than FP64's
    assert np.all(np.abs(got[6:] - fp64[6:]) < 1e-4)  # the same tree


def test_f32_invalid_and_unsupported_rows(pkg, golden_yield):
    o = pkg.Options(strike=[63.0, 63.0, 63.0], maturity=[9.0, 9.0, 9.0], length=[3.0, 3.0, 3.0],
                    term_unit=[365, 365, 365], term_steps=[12, 12, 12], a=[0.1, -0.1, 0.0001],
                    sigma=[0.01, 0.01, 0.01], type=[1, 1, 1])
    with pytest.raises(pkg.capi.HwtError) as e:
        pkg.capi.price_f32(o, golden_yield)
    assert e.value.code == pkg.capi.HWT_EINVAL


def test_constants_path_matches_oracle(pkg, oracle, golden_yield):
    """Trees given by their constants (the reference's OptionConstants) price
    as the options they were derived from."""
    y = pkg.datasets.synthetic_yield()
    opts = pkg.Options.concat([pkg.datasets.book(list(range(1, 41))), pkg.datasets.skewed(seed=23, n=600)])
    consts = pkg.capi.constants_of(opts)
    got = pkg.capi.price_constants(consts, y)
    ref = oracle.price(opts, y)
    assert np.all(np.abs(got - ref) <= parity_tol(ref))
    # and the same numbers as the option path
    np.testing.assert_allclose(got, pkg.capi.price(opts, y), rtol=1e-12, atol=1e-12)


def test_constants_path_takes_the_constants_as_given(pkg, oracle, golden_yield):
    """A height or exercise step from the constants, not re-derived: a longer
    tree priced from constants equals the option with the longer bond; an
    inconsistent width is rejected."""
    base = pkg.datasets.book([12])
    c = pkg.capi.constants_of(base)
    longer = pkg.datasets.book([12])
    longer.maturity[:] = 12.0
    c[0].n = pkg.capi.constants_of(longer)[0].n
    got = pkg.capi.price_constants(c, golden_yield)
    ref = oracle.price(longer, golden_yield)
    assert np.all(np.abs(got - ref) <= parity_tol(ref))
    c[0].width = c[0].width + 2
    with pytest.raises(pkg.capi.HwtError) as e:
        pkg.capi.price_constants(c, golden_yield)
    assert e.value.code == pkg.capi.HWT_EINVAL


def test_f32_constants_take_the_exercise_step_in_float(pkg, oracle, golden_yield):
    """OptionConstants of a real = float caller, priced from the constants:
    at 15 steps a year dt = 0.0666666701f and 3 / dt is 44.99999619 in float
    (exercise step 44), where the FP64 reference's dt = 1/15 gives 45; the
    FP32 entry prices the FP32 reference's tree."""
    opts = pkg.datasets.book([15, 30, 12, 3, 7])
    c32 = (pkg.capi.ConstantsF32C * opts.N)()
    for i in range(opts.N):
        c = oracle.constants_f32(opts, i)  # the FP32 build's constants, widened exactly
        c32[i] = pkg.capi.ConstantsF32C(c.t, c.dt, c.dr, c.X, c.M, c.jmax, c.n, c.width, c.termUnit, c.type)
    assert int(np.float32(3.0) / np.float32(c32[0].dt)) == 44 and int(3.0 / (1.0 / 15)) == 45
    got = pkg.capi.price_constants_f32(c32, golden_yield)
    assert _close_f32(got, oracle.price_f32shape(opts, golden_yield)).all()
    assert abs(got[0] - oracle.price(opts, golden_yield)[0]) > 0.1


def test_batch_uploaded_once_prices_many_times(pkg, oracle):
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed(seed=29, n=3000)
    ref = oracle.price(opts, y)
    b = pkg.capi.Batch(opts, y)
    assert b.device_memory >= opts.N * 45
    C = pkg.capi
    outs = [b.price(C.config(form=f, sort=s)) for f, s in ((C.FORM_AUTO, "L"), (C.FORM_WAVE, "H"),
                                                            (C.FORM_LANE, "-"), (C.FORM_AUTO, "W"))]
    for got in outs:
        assert np.all(np.abs(got - ref) <= parity_tol(ref))
    # the same settings again: the batch's kept plan is reused (no new device
    # memory) and replays the same prices bit for bit
    held = b.device_memory
    again = [b.price(C.config(form=C.FORM_AUTO, sort="W")) for _ in range(3)]
    assert b.device_memory == held
    for got in again:
        assert np.array_equal(got, outs[-1])
    b.close()


def test_lane_columns_sized_by_the_batch(pkg, oracle):
    """Form 1's tree columns are sized by the widest tree of a one-shot batch:
    a batch of narrow trees holds far less device memory than one with a
    1,327-wide tree; a plan limited to a width rejects wider trees."""
    y = pkg.datasets.synthetic_yield()
    small = pkg.datasets.uniform_1k()
    big = pkg.Options.concat([small, pkg.datasets.skewed(seed=2, n=200, large_frac=1.0).take([0])])
    C = pkg.capi
    _, st_s = C.price(small, y, C.config(form=C.FORM_LANE), stats=True)
    got, st_b = C.price(big, y, C.config(form=C.FORM_LANE), stats=True)
    assert st_s["device_memory"] * 5 < st_b["device_memory"]
    ref = oracle.price(big, y)
    assert np.all(np.abs(got - ref) <= parity_tol(ref))
    cfg = C.config(form=C.FORM_LANE)
    cfg.lane_max_width = 101
    rc, out = C.price_status(big, y, cfg)
    assert rc == C.HWT_EUNSUPPORTED
    assert np.isnan(out[-1]) and np.all(np.abs(out[:-1] - ref[:-1]) <= parity_tol(ref[:-1]))
