"""The surrogate on the GPU against its CPU reference and the FP64 tree.

Bars:
- inputs: |x_gpu - x_cpu| <= 1e-5 (FP32 of the same FP64 formulas); targets
  within 1e-6 of (tree price - closed form) / 0.01;
- forward: network outputs within 2e-4 * max(1, |y|) of the BF16-emulating
  CPU forward, for every member (same rounding points; the accumulation
  order differs, which now and then moves a BF16 activation by one unit);
- gradients: per tensor, ||g_gpu - g_cpu|| <= 1e-2 ||g_cpu|| (the GPU also
  rounds back-propagated values to BF16); a transposed or misplaced operand
  gives errors of order one;
- inference kernel: ensemble mean and spread within 0.01 * 4e-4 * max(1, |y|)
  of the CPU members' BF16-emulating forward (other accumulation order);
- Adam: the closed-form first and second steps to FP32 rounding;
- gate: with threshold 0 every option is re-priced by the tree and the output
  equals the tree's bit for bit; options outside the box always go to the tree.
"""
import numpy as np
import pytest

import surrogate_ref as R

pytestmark = pytest.mark.gpu

NTRAIN = 16384


@pytest.fixture(scope="module")
def setup(pkg):
    y = pkg.datasets.synthetic_yield()
    opts, dev, prices = pkg.surrogate.training_set(21, NTRAIN, y)
    cfg = pkg.surrogate.config(batch=1024, seed=3)
    s = pkg.surrogate.Surrogate(NTRAIN, 4096, y.N, cfg)
    s.set_training_data(dev, prices)
    return {"y": y, "opts": opts, "dev": dev, "tree": prices.to_host(), "s": s}


def test_inputs_and_targets_match_cpu(setup):
    s, opts, y, tree = setup["s"], setup["opts"], setup["y"], setup["tree"]
    rows = np.arange(0, NTRAIN, 97)
    x, t = s.get_inputs(rows)
    for k, i in enumerate(rows):
        xr, base, ok = R.inputs(opts, i, y)
        assert ok
        np.testing.assert_allclose(x[k], xr, rtol=0, atol=1e-5)
        assert abs(t[k] - (tree[i] - base) / R.RES) <= 1e-6 * max(1.0, abs(t[k])) + 1e-5
    # the closed form is the tree's continuous-time limit: residuals are small
    assert np.max(np.abs(t)) * R.RES < 0.1


@pytest.mark.parametrize("member", [0, 1, 3])
def test_forward_matches_bf16_cpu_reference(setup, member):
    s = setup["s"]
    P = s.get_params()
    rows = np.arange(256) * 61 % NTRAIN
    y_gpu, loss, _ = s.eval_rows(member, rows)
    x, t = s.get_inputs(rows)
    y_cpu, _ = R.forward_bf16(P[member].astype(np.float64), x)
    err = np.abs(y_gpu - y_cpu)
    assert np.all(err <= 2e-4 * np.maximum(1.0, np.abs(y_cpu))), err.max()
    assert abs(loss - np.mean((y_cpu - t) ** 2)) <= 1e-3 * max(1e-6, loss)


def test_gradients_match_cpu_reference(setup):
    s = setup["s"]
    P = s.get_params()
    rows = (np.arange(512) * 37 + 5) % NTRAIN
    _, _, g_gpu = s.eval_rows(2, rows)
    x, t = s.get_inputs(rows)
    y_cpu, cache = R.forward_bf16(P[2].astype(np.float64), x)
    g_cpu = R.backward(cache, y_cpu.astype(np.float64), t.astype(np.float64))
    for name, o in R.OFF.items():
        size = int(np.prod(R.SHAPE[name]))
        a, b = g_gpu[o:o + size].astype(np.float64), g_cpu[o:o + size]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b) + 1e-12, \
            (name, np.linalg.norm(a - b), np.linalg.norm(b))


def test_adam_closed_form(pkg, setup):
    y = setup["y"]
    s = pkg.surrogate.Surrogate(NTRAIN, 128, y.N, pkg.surrogate.config(batch=1024, seed=9))
    P0 = s.get_params().astype(np.float64)
    rng = np.random.default_rng(0)
    g1 = rng.normal(size=P0.shape).astype(np.float32)
    g2 = rng.normal(size=P0.shape).astype(np.float32)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    s.adam_step(g1, lr)
    P1 = s.get_params().astype(np.float64)
    m = (1 - b1) * g1
    v = (1 - b2) * g1.astype(np.float64) ** 2
    want1 = P0 - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    np.testing.assert_allclose(P1, want1, rtol=0, atol=2e-6)
    s.adam_step(g2, lr)
    P2 = s.get_params().astype(np.float64)
    m = b1 * m + (1 - b1) * g2
    v = b2 * v + (1 - b2) * g2.astype(np.float64) ** 2
    want2 = P1 - lr * (m / (1 - b1 ** 2)) / (np.sqrt(v / (1 - b2 ** 2)) + eps)
    np.testing.assert_allclose(P2, want2, rtol=0, atol=2e-6)
    s.close()


def test_training_reduces_loss_and_beats_the_closed_form(setup):
    """The learned part's gain: after a short training, the network's
    residual correction cuts the error of the closed form against the tree."""
    s, tree = setup["s"], setup["tree"]
    first = s.train(1, 12)
    last = s.train(11, 12)
    assert last["loss"] < 0.3 * first["loss"], (first, last)
    mean, spread, inbox = s.predict(setup["dev"])
    assert inbox.all()
    rows = np.arange(NTRAIN)
    x, t = s.get_inputs(rows)
    base = tree - t.astype(np.float64) * R.RES
    err_base = np.abs(base - tree)
    err_surr = np.abs(mean - tree)
    print(f"\nclosed form alone: mean err {err_base.mean():.2e}, max {err_base.max():.2e}; "
          f"with network: mean {err_surr.mean():.2e}, max {err_surr.max():.2e}; "
          f"spread median {np.median(spread):.2e}")
    assert err_surr.mean() < 0.5 * err_base.mean()


def test_gate_threshold_zero_reprices_everything_on_the_tree(pkg, setup):
    y = setup["y"]
    opts = pkg.datasets.box_mix(31, 3000)
    dev = pkg.capi.to_device(opts, y)
    tree_plan = pkg.capi.Plan(4096, y.N, pkg.capi.config(flags=pkg.capi.FLAG_NO_GRAPH))
    tree_plan.execute(dev)
    tree_plan.sync()
    ref = dev["out"].to_host()
    s0 = pkg.surrogate.Surrogate(NTRAIN, 4096, y.N, pkg.surrogate.config(batch=1024, spread_threshold=0.0))
    s0.set_params(setup["s"].get_params())
    dev["out"].fill_zero()
    st = s0.price(tree_plan, dev)
    got = dev["out"].to_host()
    # a spread of exactly zero is possible only if all members agree to the bit
    assert st["fallback"] >= opts.N - 5
    same = got == ref
    assert same.sum() >= opts.N - 5
    s0.close()
    tree_plan.close()


def test_out_of_box_always_falls_back(pkg, setup):
    y = setup["y"]
    o = pkg.datasets.box(41, 1000)
    o.a[:] = 0.3  # outside the box
    dev = pkg.capi.to_device(o, y)
    plan = pkg.capi.Plan(4096, y.N, pkg.capi.config(flags=pkg.capi.FLAG_NO_GRAPH))
    plan.execute(dev)
    plan.sync()
    ref = dev["out"].to_host()
    dev["out"].fill_zero()
    st = setup["s"].price(plan, dev)
    assert st["fallback"] == o.N and st["out_of_box"] == o.N
    assert np.array_equal(dev["out"].to_host(), ref)
    mean, spread, inbox = setup["s"].predict(dev)
    assert not inbox.any() and np.isnan(mean).all()
    plan.close()


def test_pass_graph_replays_and_recaptures(pkg, setup):
    """A pricing pass is one graph: captured on the first call, replayed
    after, captured again when an argument changes (output buffer, tree
    plan).  Every replay, with the tree plan's own executions and host
    synchronisations between them, gives the first call's prices and
    fallback counts bit for bit."""
    capi, y, s = pkg.capi, setup["y"], setup["s"]
    opts = pkg.datasets.box_mix(51, 3000)
    dev = capi.to_device(opts, y)
    other = pkg.datasets.box(53, 500)
    dother = capi.to_device(other, y)
    plan = capi.Plan(4096, y.N)
    plan.execute(dother)
    plan.sync()
    ref_other = dother["out"].to_host()
    st1 = s.price(plan, dev)
    got1 = dev["out"].to_host()
    assert st1["fallback"] >= st1["out_of_box"] > 0 and np.all(np.isfinite(got1))
    for _ in range(3):
        dother["out"].fill_zero()
        plan.execute(dother)  # the tree plan's own graph between the passes
        plan.sync()
        np.testing.assert_array_equal(dother["out"].to_host(), ref_other)
        dev["out"].fill_zero()
        st = s.price(plan, dev)
        np.testing.assert_array_equal(dev["out"].to_host(), got1)
        assert (st["fallback"], st["out_of_box"]) == (st1["fallback"], st1["out_of_box"])
    out2 = pkg.devmem.DeviceArray.empty((opts.N,), "float64")
    s.price(plan, dev, out=out2)  # another output buffer: captured again
    np.testing.assert_array_equal(out2.to_host(), got1)
    plan2 = capi.Plan(4096, y.N, capi.config(flags=capi.FLAG_NO_GRAPH))
    dev["out"].fill_zero()
    s.price(plan2, dev)  # another tree plan: captured again
    np.testing.assert_array_equal(dev["out"].to_host(), got1)
    plan2.close()
    plan.close()


def test_huge_threshold_takes_every_in_box_price_from_the_network(pkg, setup):
    y = setup["y"]
    o = pkg.datasets.box(43, 2000)
    dev = pkg.capi.to_device(o, y)
    s = pkg.surrogate.Surrogate(NTRAIN, 4096, y.N, pkg.surrogate.config(batch=1024, spread_threshold=1e9))
    s.set_params(setup["s"].get_params())
    plan = pkg.capi.Plan(4096, y.N, pkg.capi.config(flags=pkg.capi.FLAG_NO_GRAPH))
    st = s.price(plan, dev)
    assert st["fallback"] == 0
    mean, _, inbox = s.predict(dev)
    assert inbox.all()
    np.testing.assert_array_equal(dev["out"].to_host(), mean)
    s.close()
    plan.close()


def test_checkpoint_resume_is_exact(pkg, setup, tmp_path):
    y = setup["y"]
    cfg = pkg.surrogate.config(batch=1024, seed=17, members=2)
    a = pkg.surrogate.Surrogate(NTRAIN, 128, y.N, cfg)
    a.set_training_data(setup["dev"], pkg.devmem.DeviceArray.from_host(setup["tree"]))
    a.train(1, 3)
    path = tmp_path / "surr.ckpt"
    a.save(path)
    b = pkg.surrogate.Surrogate(NTRAIN, 128, y.N, pkg.surrogate.config(batch=1024, seed=99, members=2))
    b.set_training_data(setup["dev"], pkg.devmem.DeviceArray.from_host(setup["tree"]))
    b.load(path)
    np.testing.assert_array_equal(a.get_params(), b.get_params())
    a.train(1, 3)
    b.train(1, 3)
    np.testing.assert_array_equal(a.get_params(), b.get_params())
    a.close()
    b.close()


def test_inference_kernel_matches_cpu_reference(setup):
    """The fused inference kernel (all members, ensemble mean and spread)
    against the BF16-emulating CPU forward of each member, on training rows:
    mean = closed form + 0.01 * mean of the member outputs, spread = 0.01 *
    (max - min).  The closed form is recovered from the stored residual
    target (FP32), so the bar carries its rounding too."""
    s, tree = setup["s"], setup["tree"]
    P = s.get_params().astype(np.float64)
    mean, spread, inbox = s.predict(setup["dev"])
    rows = np.arange(3, NTRAIN, 257)
    x, t = s.get_inputs(rows)
    ys = np.stack([R.forward_bf16(P[m], x)[0] for m in range(s.members)])
    base = tree[rows] - t.astype(np.float64) * R.RES
    want_mean = base + R.RES * ys.mean(axis=0)
    want_spread = R.RES * (ys.max(axis=0) - ys.min(axis=0))
    assert inbox[rows].all()
    tol = R.RES * 4e-4 * np.maximum(1.0, np.abs(ys).max(axis=0)) + 1e-6 * np.abs(t) * R.RES + 1e-12
    np.testing.assert_array_less(np.abs(mean[rows] - want_mean), tol)
    np.testing.assert_array_less(np.abs(spread[rows] - want_spread), 2 * tol)


@pytest.mark.parametrize("members", [1, 2])
def test_smaller_ensembles_price_through_the_gate(pkg, setup, members):
    """Ensembles of 1 and 2 members.  A single member has no spread (always
    0), so it gates only with an infinite threshold, where every in-box option
    takes the network's price; a finite threshold is refused.  Out-of-box
    options still go to the tree, and the network's prices equal predict()'s
    mean."""
    y = setup["y"]
    o = pkg.datasets.box_mix(51, 3000)
    dev = pkg.capi.to_device(o, y)
    thr = float("inf") if members == 1 else 1e-3
    s = pkg.surrogate.Surrogate(0, 4096, y.N, pkg.surrogate.config(members=members, spread_threshold=thr))
    s.set_params(setup["s"].get_params()[:members])
    plan = pkg.surrogate.fallback_plan(4096, y.N)
    st = s.price(plan, dev)
    got = dev["out"].to_host()
    mean, spread, inbox = s.predict(dev)
    if members == 1:
        assert np.all(spread[inbox] == 0.0)
        assert st["fallback"] == st["out_of_box"] == int((~inbox).sum())
    accepted = inbox & (spread <= s.cfg.spread_threshold)
    np.testing.assert_array_equal(got[accepted], mean[accepted])
    plan.execute(dev)
    plan.sync()
    tree = dev["out"].to_host()
    np.testing.assert_array_equal(got[~accepted], tree[~accepted])
    s.close()
    if members == 1:
        one = pkg.surrogate.Surrogate(0, 4096, y.N, pkg.surrogate.config(members=1, spread_threshold=1e-3))
        one.set_params(setup["s"].get_params()[:1])
        with pytest.raises(pkg.capi.HwtError, match="one-member"):
            one.price(plan, dev)
        one.close()
    plan.close()


def test_empty_batch_prices_nothing(pkg, setup):
    y = setup["y"]
    o = pkg.datasets.box(3, 0)
    dev = pkg.capi.to_device(o, y)
    plan = pkg.surrogate.fallback_plan(128, y.N)
    st = setup["s"].price(plan, dev)
    assert st["fallback"] == 0 and st["out_of_box"] == 0
    plan.close()
