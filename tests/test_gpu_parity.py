"""GPU parity: the HIP pricer through its C ABI against the CPU oracle.

Bar (FP64): |gpu - oracle| <= 1e-10 * max(|oracle|, 1), i.e. 1e-10 relative
with an absolute floor of 1e-10 per 100 notional for prices below 1.
Full-size batches are checked through size-independent properties
(put-call parity on the tree, the zero-coupon-bond identity, determinism).
"""
import math

import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_SETS, parity_tol

pytestmark = pytest.mark.gpu


def assert_parity(got, ref, what=""):
    err = np.abs(got - ref)
    bad = np.nonzero(~(err <= parity_tol(ref)))[0]
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first {bad[:5]}: got {got[bad[:5]]} want {ref[bad[:5]]}"


@pytest.mark.parametrize("form", ["auto", "lane"])
@pytest.mark.parametrize("name", GOLDEN_SETS)
def test_golden_sets(pkg, oracle, golden_yield, name, form):
    opts = pkg.formats.read_options(GOLDEN / f"{name}.in")
    got = pkg.capi.price(opts, golden_yield, pkg.capi.config(form=FORMS[form]))
    assert_parity(got, oracle.price(opts, golden_yield), name)
    # and against the reference's own golden output (17 decimals, other libm)
    gold = pkg.formats.read_results(GOLDEN / f"{name}.out")
    np.testing.assert_allclose(got, gold, rtol=1e-10, atol=1e-11)


FORMS = {"auto": 0, "lane": 1, "wave": 2}


@pytest.mark.parametrize("form", ["auto", "wave", "lane"])
@pytest.mark.parametrize("sort", ["L", "-", "w", "W", "h", "H"])
def test_book_options_all_runners(pkg, oracle, golden_yield, form, sort):
    """The reference's only GPU test (test/TestCuda.cu:16-99): Hull's book
    option priced with steps/year 1..100, every runner against seq."""
    opts = pkg.datasets.book(list(range(1, 101)))
    f = FORMS[form]
    got = pkg.capi.price(opts, golden_yield, pkg.capi.config(form=f, sort=sort))
    assert_parity(got, oracle.price(opts, golden_yield))


def test_uniform_1k(pkg, oracle):
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.uniform_1k()
    assert_parity(pkg.capi.price(opts, y), oracle.price(opts, y))


def test_skewed_60k_stratified(pkg, oracle):
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed_60k()
    got = pkg.capi.price(opts, y)
    sub = np.arange(0, opts.N, 40)
    assert_parity(got[sub], oracle.price(opts.take(sub), y), "skewed_60k every 40th")
    assert np.all(np.isfinite(got)) and np.all(got >= 0)
    lane = pkg.capi.price(opts, y, pkg.capi.config(form=pkg.capi.FORM_LANE))
    assert_parity(lane, got, "lane form vs wave form, full batch")


def _yield_at(t, y, term_unit=365):
    d = math.floor(t * term_unit + 0.5)  # C round(): half away from zero (t >= 0)
    days, rates = y.days, y.rates
    if d <= days[0]:
        return rates[0]
    if d > days[-1]:
        return rates[-1]
    i = int(np.searchsorted(days, d))
    return rates[i - 1] + (d - days[i - 1]) / (days[i] - days[i - 1]) * (rates[i] - rates[i - 1])


def test_put_call_parity_full_size(pkg):
    """C - P = 100 P(n dt) - X P(e dt) holds exactly on the calibrated tree,
    for every option of the full 60k batch."""
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed_60k()
    put = opts.take(np.arange(opts.N))
    put.type[:] = pkg.PUT
    call = opts.take(np.arange(opts.N))
    call.type[:] = pkg.CALL
    p, c = pkg.capi.price(put, y), pkg.capi.price(call, y)
    dt = 1.0 / opts.term_steps.astype(np.float64)
    n = (opts.term_steps.astype(np.int64) * opts.maturity).astype(np.int64)
    e = (opts.length / dt).astype(np.int64)
    rhs = np.array([100 * math.exp(-_yield_at(n[i] * dt[i], y) * n[i] * dt[i])
                    - opts.strike[i] * math.exp(-_yield_at(e[i] * dt[i], y) * e[i] * dt[i])
                    for i in range(opts.N)])
    np.testing.assert_allclose(c - p, rhs, rtol=0, atol=1e-9)


def _one(pkg, **kw):
    base = dict(strike=[80.0], maturity=[9.0], length=[3.0], term_unit=[365], term_steps=[12],
                a=[0.1], sigma=[0.01], type=[pkg.PUT])
    base.update({k: [v] for k, v in kw.items()})
    return pkg.Options(**base)


@pytest.mark.parametrize("form", ["auto", "lane"])
@pytest.mark.parametrize("kw", [
    dict(a=3.0, term_steps=1, maturity=6.0, length=2.0, sigma=0.03),     # jmax 1
    dict(a=3.0, term_steps=1, maturity=6.0, length=2.0, sigma=0.03, type=0),
    dict(a=0.15, term_steps=1, maturity=8.0, length=3.0, sigma=0.03),    # jmax 2
    dict(a=0.08, term_steps=1, maturity=9.0, length=4.0, sigma=0.03),    # jmax 3
    dict(length=0.04),                                                   # exercise at step 0
    dict(length=0.25),                                                   # exercise at step 3 < jmax 23
    dict(length=1.0, type=0),                                            # exercise at step 12 < jmax 23
    dict(a=0.02, term_steps=36, maturity=5.0, length=2.0),               # e 72 < jmax 256: cone narrows
    dict(length=8.95),                                                   # exercise at the last step
    dict(length=9.0),                                                    # exercise == height: bond
    dict(length=12.0),                                                   # beyond the tree: bond
    dict(maturity=0.05),                                                 # zero steps
    dict(strike=0.0, type=0),                                            # deep in the money call
    dict(strike=1000.0, type=0),                                         # worthless call
    dict(a=0.0111, term_steps=36, maturity=12.0, length=6.0),            # width near the maximum
    dict(term_unit=30, term_steps=1),                                    # monthly term units
])
def test_edge_cases(pkg, oracle, kw, form):
    y = pkg.datasets.synthetic_yield()
    o = _one(pkg, **kw)
    assert_parity(pkg.capi.price(o, y, pkg.capi.config(form=FORMS[form])), oracle.price(o, y), str(kw))


def test_widest_supported_tree(pkg, oracle):
    y = pkg.datasets.synthetic_yield()
    # a such that jmax = 671, width 1343 (and 672 -> 1345, beyond the limit)
    dt = 1 / 36
    a_ok = -math.log(1 - 0.184 / 670.5) / dt
    o = _one(pkg, a=a_ok, term_steps=36, maturity=2.0, length=1.0)
    assert oracle.constants(o, 0).width == 1343
    assert_parity(pkg.capi.price(o, y), oracle.price(o, y))


def test_window_phase_boundaries(pkg, oracle):
    """Wide trees whose register window changes exactly where the light cone
    reaches a window edge: half-widths J at and around multiples of 128 nodes
    (a window grows four nodes per lane, 128 nodes of half-width, at a time),
    exercise steps at and around 32-step multiples and at J, and heights that
    end the backward pass just before, at and after a shrink."""
    y = pkg.datasets.synthetic_yield()
    dt = 1 / 36
    cols = {k: [] for k in ("strike", "maturity", "length", "term_unit", "term_steps", "a", "sigma", "type")}
    for J in (127, 128, 129, 255, 256, 257, 383, 384, 385, 511, 512, 513, 663):
        a = -math.log(1 - 0.184 / (J - 0.5)) / dt
        for e in sorted({0, 1, 31, 32, 127, 128, 129, 255, 256, 257, J - 1, J, J + 1}):
            for extra in (1, 2, 129, 259, 2 * J + 7):
                n = e + extra
                cols["strike"].append(70.0 + (J + e + extra) % 30)
                cols["maturity"].append((n + 0.5) * dt)
                cols["length"].append((e + 0.5) * dt)
                cols["term_unit"].append(365)
                cols["term_steps"].append(36)
                cols["a"].append(a)
                cols["sigma"].append(0.01)
                cols["type"].append((J + e + extra) % 2)
    o = pkg.Options(**cols)
    assert {oracle.constants(o, i).width for i in range(o.N)} >= {255, 257, 259, 511, 1027, 1327}
    ref = oracle.price(o, y)
    assert_parity(pkg.capi.price(o, y), ref, "window phase boundaries")
    assert_parity(pkg.capi.price(o, y, pkg.capi.config(flags=pkg.capi.FLAG_NO_GRAPH, sort="-")), ref,
                  "window phase boundaries, unsorted")


def test_too_wide_is_unsupported_and_others_priced(pkg, oracle):
    y = pkg.datasets.synthetic_yield()
    dt = 1 / 36
    a_bad = -math.log(1 - 0.184 / 680.5) / dt
    o = pkg.Options.concat([_one(pkg), _one(pkg, a=a_bad, term_steps=36, maturity=2.0, length=1.0)])
    assert oracle.constants(o, 1).width > 1344
    rc, got = pkg.capi.price_status(o, y)
    assert rc == pkg.capi.HWT_EUNSUPPORTED
    assert np.isnan(got[1])
    assert_parity(got[:1], oracle.price(o.take([0]), y))


def test_invalid_inputs(pkg, oracle):
    y = pkg.datasets.synthetic_yield()
    good = _one(pkg)
    for kw in (dict(a=0.0), dict(a=-0.1), dict(term_steps=0), dict(term_unit=0),
               dict(maturity=-1.0), dict(sigma=-0.01)):
        o = pkg.Options.concat([good, _one(pkg, **kw)])
        rc, got = pkg.capi.price_status(o, y)
        assert rc == pkg.capi.HWT_EINVAL, kw
        assert np.isnan(got[1])
        assert_parity(got[:1], oracle.price(good, y))


def test_empty_batch(pkg):
    y = pkg.datasets.synthetic_yield()
    o = pkg.Options(strike=[], maturity=[], length=[], term_unit=[], term_steps=[], a=[], sigma=[], type=[])
    assert pkg.capi.price(o, y).shape == (0,)


def test_alphas_match_oracle(pkg, oracle, golden_yield):
    opts = pkg.Options.concat([pkg.datasets.book([1, 5, 12, 36]),
                               pkg.formats.read_options(GOLDEN / "1_RAND_1000.in").take(np.arange(0, 1000, 97))])
    got = pkg.capi.alphas(opts, golden_yield)
    for i in range(opts.N):
        _, ref = oracle.price_single(opts, i, golden_yield, with_alphas=True)
        assert got[i].shape == ref.shape
        np.testing.assert_allclose(got[i], ref, rtol=1e-9, atol=1e-12)


def test_deterministic_and_order_independent(pkg):
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed(seed=9, n=3000)
    a = pkg.capi.price(opts, y)
    b = pkg.capi.price(opts, y)
    assert np.array_equal(a, b)
    for sort in "-wWhH":
        assert np.array_equal(a, pkg.capi.price(opts, y, pkg.capi.config(sort=sort)))
    perm = np.random.default_rng(1).permutation(opts.N)
    c = pkg.capi.price(opts.take(perm), y)
    assert np.array_equal(c, a[perm])


def test_plan_graph_replays(pkg, oracle):
    """A plan captured into a HIP graph and replayed: every replay, with other
    device work in between, equals the first and the oracle."""
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed(seed=4, n=2000)
    dev = pkg.capi.to_device(opts, y)
    plan = pkg.capi.Plan(opts.N, y.N)
    plan.execute(dev)
    plan.sync()
    first = dev["out"].to_host()
    assert_parity(first, oracle.price(opts, y))
    other = pkg.capi.to_device(pkg.datasets.uniform_1k(), y)
    plan2 = pkg.capi.Plan(1000, y.N, pkg.capi.config(flags=pkg.capi.FLAG_NO_GRAPH))
    for _ in range(4):
        dev["out"].fill_zero()
        plan2.execute(other)
        plan2.sync()
        plan.execute(dev)
        st = plan.sync()
        assert np.array_equal(dev["out"].to_host(), first)
        assert st["tasks"] == opts.N and st["unsupported"] == 0
    plan.close()
    plan2.close()


@pytest.mark.parametrize("flags", [0, 1, 32], ids=["graph", "eager", "direct"])
def test_back_to_back_executions_leave_counters_clear(pkg, oracle, flags):
    """One plan, executions queued back to back with no sync between them,
    then an empty batch, a batch with an invalid option and a good batch
    again: each execution starts from clear counters (the scan kernel clears
    them for the next one), so prices and the tallies read at sync are the
    last execution's own."""
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed(seed=6, n=1500)
    dev = pkg.capi.to_device(opts, y)
    ref = oracle.price(opts, y)
    plan = pkg.capi.Plan(opts.N + 1, y.N, pkg.capi.config(flags=flags))
    plan.execute(dev)
    st1 = plan.sync()
    assert_parity(dev["out"].to_host(), ref, "first execution")
    keys = ("tasks", "wide_tasks", "node_updates", "wide_node_updates", "seq_node_updates", "unsupported",
            "wide_step_ops", "wide_unit_slots", "wide_plain_slots")
    assert st1["tasks"] == opts.N and st1["wide_tasks"] > 0
    for _ in range(3):
        plan.execute(dev)
    st3 = plan.sync()
    assert_parity(dev["out"].to_host(), ref, "after three back-to-back executions")
    assert {k: st3[k] for k in keys} == {k: st1[k] for k in keys}
    plan.execute(dev, n=0)
    st0 = plan.sync()
    assert st0["tasks"] == 0 and st0["node_updates"] == 0 and st0["wide_tasks"] == 0
    bad = pkg.Options.concat([opts, _one(pkg, a=0.0)])
    dbad = pkg.capi.to_device(bad, y)
    plan.execute(dbad)
    with pytest.raises(pkg.capi.HwtError) as err:
        plan.sync()
    assert err.value.code == pkg.capi.HWT_EINVAL
    assert_parity(dbad["out"].to_host()[:opts.N], ref, "good options beside an invalid one")
    dev["out"].fill_zero()
    plan.execute(dev)
    st = plan.sync()  # no invalid count left over from the previous execution
    assert_parity(dev["out"].to_host(), ref, "after the invalid batch")
    assert {k: st[k] for k in keys} == {k: st1[k] for k in keys}
    plan.close()


@pytest.mark.parametrize("flags", [0, 1, 4, 8, 32], ids=["graph", "eager", "concurrent", "fused", "direct"])
def test_execution_time_stamps(pkg, flags):
    """hwt_stats.total_ms comes from the kernels' wall-clock stamps (the
    set-up kernel's start to the last pricing wavefront's end), not from
    events around the execution.  For back-to-back executions it must lie
    between 40% and 105% of their period measured by events around all of
    them, and an empty batch reports 0."""
    import ctypes as C
    capi, dm = pkg.capi, pkg.devmem
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed(seed=6, n=6000)
    dev = capi.to_device(opts, y)
    plan = capi.Plan(opts.N, y.N, capi.config(flags=flags))
    for _ in range(5):
        plan.execute(dev)
    plan.sync()
    stream = C.c_void_p()
    dm._ck(dm.hip().hipStreamCreate(C.byref(stream)))
    reps = 20
    e0, e1 = dm.Event(), dm.Event()
    e0.record(stream)
    for _ in range(reps):
        plan.execute(dev, stream=stream.value)
    e1.record(stream)
    e1.synchronize()
    period = e0.elapsed_ms(e1) / reps
    st = plan.sync()
    assert 0.4 * period <= st["total_ms"] <= 1.05 * period, (st["total_ms"], period)
    if flags & capi.FLAG_NO_GRAPH:
        assert 0.0 < st["kernel_ms"] <= 1.05 * period
    plan.execute(dev, n=0)
    assert plan.sync()["total_ms"] == 0.0
    plan.close()
    dm._ck(dm.hip().hipStreamDestroy(stream))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wide_step_counts(pkg, precision):
    """The wide kernel's own count of what its stencil steps executed.  Every
    window slot it stepped is counted once, in the form it ran in, so the
    slots cover the active tree region (the work estimate) with padding on
    top; the operations follow from the forms: 4K+1 per lane and step in the
    unit-diagonal form (between 4 and 5 per slot), 5 per slot in the plain
    form.  FP64 steps every window of these trees in the unit form; mixed
    precision steps the same windows, in the unit form wherever the setup
    kernel's FP32 scale check passes (here: every window), else plain."""
    capi = pkg.capi
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed(seed=3, n=3000)
    prec = capi.PREC_FP64 if precision == "fp64" else capi.PREC_MIXED
    _, st = capi.price(opts, y, capi.config(form=capi.FORM_WAVE, precision=prec), stats=True)
    unit, plain, ops = st["wide_unit_slots"], st["wide_plain_slots"], st["wide_step_ops"]
    assert st["wide_tasks"] > 0
    assert st["wide_node_updates"] <= unit + plain <= 1.6 * st["wide_node_updates"]
    assert 4 * unit + 5 * plain <= ops <= 5 * (unit + plain)
    if precision == "fp64":
        assert plain == 0 and unit > 0
    else:
        _, s64 = capi.price(opts, y, capi.config(form=capi.FORM_WAVE), stats=True)
        assert unit + plain == s64["wide_unit_slots"]  # the same windows and steps
        assert unit > 0 and plain == 0


def test_plan_capacity_enforced(pkg):
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.uniform_1k()
    dev = pkg.capi.to_device(opts, y)
    plan = pkg.capi.Plan(10, y.N)
    assert plan.execute_status(dev) == pkg.capi.HWT_ENOMEM
    assert plan.execute_status(dev, n=10) == pkg.capi.HWT_OK
    plan.sync()
    plan.close()


@pytest.mark.parametrize("curve", ["unsorted", "long", "single"])
def test_yield_curve_shapes(pkg, oracle, curve):
    """The setup kernel reads curves that fit in LDS by binary search when
    their days never decrease, and keeps the reference's linear scan
    (common/Domain.hpp:11-46) otherwise: every shape prices as the oracle."""
    base = pkg.datasets.synthetic_yield()
    if curve == "unsorted":
        perm = np.arange(base.N)
        perm[[10, 11, 60, 61]] = perm[[11, 10, 61, 60]]
        y = pkg.formats.Yield(base.rates[perm], base.days[perm])
    elif curve == "long":  # more knots than the setup kernel stages
        k = np.arange(1, 3001)
        y = pkg.formats.Yield(0.01 + 0.03 * (k - 1) / 2999.0, np.rint(365.0 * 40 * k / 3000).astype(np.int32))
    else:
        y = pkg.formats.Yield(np.array([0.025]), np.array([365], np.int32))
    opts = pkg.Options.concat([pkg.datasets.book(), pkg.datasets.skewed(seed=9, n=200)])
    assert_parity(pkg.capi.price(opts, y), oracle.price(opts, y), curve)


def test_long_curve_on_a_multi_block_batch(pkg, oracle):
    """A 3,650-knot curve (more knots than the setup kernel stages in LDS) on
    the 60,000-option batch, 235 setup blocks: the same piecewise-linear curve
    as the 120-knot one, so every price, the task count and the work tallies
    must equal the short curve's, and a stratified subset the oracle's.  The
    setup kernel's per-block histogram and tallies are cleared in LDS before
    any thread adds to them, for curves of every length."""
    short = pkg.datasets.synthetic_yield()
    days = np.arange(1, 3651, dtype=np.int32)  # every day to the short curve's last knot
    dense = pkg.formats.Yield(np.interp(days, short.days, short.rates), days)
    # knots of the short curve lie on the dense one: same interpolated rates
    assert np.all(np.isin(short.days, days))
    opts = pkg.datasets.skewed_60k()
    got, st = pkg.capi.price(opts, dense, stats=True)
    want, st0 = pkg.capi.price(opts, short, stats=True)
    for k in ("tasks", "wide_tasks", "node_updates", "wide_node_updates", "seq_node_updates", "unsupported"):
        assert st[k] == st0[k], k
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11)
    sub = np.arange(0, opts.N, 97)
    assert_parity(got[sub], oracle.price(opts.take(sub), dense), "3,650-knot curve")


@pytest.mark.parametrize("name", ["0_UNIFORM", "1_RAND", "2_RANDCONSTHEIGHT", "3_RANDCONSTWIDTH", "4_SKEWED",
                                  "5_SKEWEDCONSTHEIGHT", "6_SKEWEDCONSTWIDTH"])
def test_reference_distributions_65k_stratified(pkg, oracle, golden_yield, name):
    """The reference's seven 65,536-option benchmark distributions, drawn by
    the seeded generator and priced whole on the reference's 15-knot curve:
    a stratified subset (every tree width and height class, the skewed part
    included) against the oracle, and the whole batch against the
    one-tree-per-wave arrangement."""
    opts = pkg.datasets.reference_set(name, 65536, seed=5)
    got = pkg.capi.price(opts, golden_yield)
    assert np.all(np.isfinite(got))
    w = 2 * ((-0.184 / (np.exp(-opts.a / 12.0) - 1.0)).astype(int) + 1) + 1
    order = np.lexsort((opts.maturity, w))
    sub = np.unique(np.concatenate([order[::331], order[-40:], np.nonzero(w >= 411)[0][:40]]))
    assert_parity(got[sub], oracle.price(opts.take(sub), golden_yield), name)
    other = pkg.capi.price(opts, golden_yield, pkg.capi.config(flags=pkg.capi.FLAG_NO_ROWS | pkg.capi.FLAG_FUSED))
    assert_parity(other, got, name + " (one tree per wave, fused)")


@pytest.mark.parametrize("flags", [0, 2, 4, 8, 16])
def test_flat_form_tiny_trees(pkg, oracle, golden_yield, flags):
    """Trees 3-21 nodes wide, packed 64 / w to a wavefront one node per lane
    (the flattened form; flags 16 turns it off, 2 puts the narrow trees one
    per wavefront, 4 and 8 run the narrow kernel beside or fused with the
    wide one), with the extreme shapes: jmax = 1, exercise at step 0, a
    one-step tree, a mix of heights in every group."""
    y = golden_yield
    opts = pkg.datasets.tiny(61, 5000)
    extra = pkg.Options(strike=[63.0, 70.0, 80.0, 90.0], maturity=[1.0 / 12, 0.25, 9.0, 30.0],
                        length=[0.0, 0.0, 3.0, 29.9], term_unit=[365] * 4, term_steps=[12] * 4,
                        a=[3.0, 0.6, 3.0, 0.3], sigma=[0.01, 0.02, 0.01, 0.005], type=[1, 0, 1, 0])
    opts = pkg.Options.concat([extra, opts])
    got = pkg.capi.price(opts, y, pkg.capi.config(flags=flags))
    sub = np.concatenate([np.arange(4), np.arange(4, opts.N, 7)])
    assert_parity(got[sub], oracle.price(opts.take(sub), y), f"flags {flags}")
    assert_parity(got, pkg.capi.price(opts, y), f"flags {flags} against the default")


def test_flat_form_mixed_precision_and_packing(pkg, oracle, golden_yield):
    """Mixed precision through the flattened form, within the mixed bar of the
    oracle; prices do not depend on which trees share a wavefront (the same
    options shuffled price the same, bit for bit)."""
    y = golden_yield
    opts = pkg.datasets.tiny(67, 3000)
    mixed = pkg.capi.price(opts, y, pkg.capi.config(precision=pkg.capi.PREC_MIXED))
    sub = np.arange(0, opts.N, 11)
    assert np.abs(mixed[sub] - oracle.price(opts.take(sub), y)).max() <= 1e-3
    a = pkg.capi.price(opts, y)
    perm = np.random.default_rng(2).permutation(opts.N)
    b = pkg.capi.price(opts.take(perm), y)
    assert np.array_equal(a[perm], b)


@pytest.mark.parametrize("setting", [{"flags": 2}, {"flags": 4}, {"flags": 6}, {"flags": 8}, {"flags": 10},
                                     {"flags": 9}, {"flags": 16}, {"flags": 18}, {"flags": 24},
                                     {"flags": 32}, {"flags": 48}, {"flags": 36},
                                     {"wide_per_cu": 2, "narrow_per_cu": 4},
                                     {"sort": "H"}, {"sort": "W"}])
def test_form2_settings_price_the_same(pkg, oracle, setting):
    """Every setting the autotuner can choose prices within the parity bar."""
    y = pkg.datasets.synthetic_yield()
    opts = pkg.Options.concat([pkg.datasets.book(), pkg.datasets.skewed(seed=17, n=1500)])
    kw = dict(setting)
    cfg = pkg.capi.config(sort=kw.pop("sort", "L"), flags=kw.pop("flags", 0))
    for k, v in kw.items():
        setattr(cfg, k, v)
    got = pkg.capi.price(opts, y, cfg)
    assert_parity(got, oracle.price(opts, y), str(setting))


def test_autotune_chooses_and_caches(pkg, oracle):
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.skewed(seed=13, n=4000)
    dev = pkg.capi.to_device(opts, y)
    cfg, rep = pkg.capi.autotune(dev)
    assert rep["candidates"] >= 2 and rep["cached"] == 0
    assert 0 < rep["ms"] <= rep["default_ms"]
    tuned = pkg.capi.price(opts, y, cfg)
    sub = np.arange(0, opts.N, 40)
    assert_parity(tuned[sub], oracle.price(opts.take(sub), y), "tuned configuration")
    # the same batch shape again: the cached choice, no search
    cfg2, rep2 = pkg.capi.autotune(dev)
    assert rep2["cached"] == 1 and rep2["candidates"] == 1
    assert cfg2.as_dict() == cfg.as_dict()


def test_fresh_plan_syncs_clean(pkg):
    """A plan that has never executed reports success and empty stats."""
    plan = pkg.capi.Plan(1000, 10)
    st = plan.sync()
    assert st["tasks"] == 0 and st["node_updates"] == 0 and st["unsupported"] == 0
    plan.close()
