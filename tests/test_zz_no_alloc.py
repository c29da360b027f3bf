"""The allocation-free entry point (hwt_price_device_noalloc) against the
plan-based one it varies (hwt_plan_execute, same kernels): prices equal bit for
bit, over two calls in a row on the same static scratch, the second with a
smaller batch (the array keeps the first call's contents), on a caller stream;
a batch above the fixed capacity is refused with HWT_ENOMEM."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(pkg, dev, n, stream=None):
    capi = pkg.capi
    o = capi.OptionsC(int(n), *[dev[f].value for f in capi.Plan.FIELDS])
    y = capi.YieldC(int(dev["ny"]), dev["rates"].value, dev["days"].value)
    return capi.lib().hwt_price_device_noalloc(C.byref(o), C.byref(y), dev["out"].value,
                                               C.c_void_p(stream) if stream else None)


def test_noalloc_matches_plan_twice_second_smaller(pkg, oracle):
    capi, dm = pkg.capi, pkg.devmem
    y = pkg.datasets.synthetic_yield()
    big = pkg.datasets.skewed(seed=41, n=20000)
    small = pkg.Options.concat([pkg.datasets.book(), pkg.datasets.skewed(seed=43, n=3000)])
    stream = C.c_void_p()
    dm._ck(dm.hip().hipStreamCreate(C.byref(stream)))
    for opts in (big, small):
        dev = capi.to_device(opts, y)
        plan = capi.Plan(opts.N, y.N)
        plan.execute(dev)
        plan.sync()
        want = dev["out"].to_host()
        plan.close()
        dev["out"].fill_zero()
        dm.synchronize()
        print(f"noalloc call: {opts.N} options", flush=True)
        assert _noalloc(pkg, dev, opts.N, stream.value) == capi.HWT_OK
        dm._ck(dm.hip().hipStreamSynchronize(stream))
        got = dev["out"].to_host()
        assert np.array_equal(got, want), np.nonzero(got != want)[0][:5]
        sub = np.arange(0, opts.N, 53)
        from conftest import parity_tol
        ref = oracle.price(opts.take(sub), y)
        assert np.all(np.abs(got[sub] - ref) <= parity_tol(ref))
        for k in list(dev):
            if hasattr(dev[k], "free"):
                dev[k].free()
    dm._ck(dm.hip().hipStreamDestroy(stream))


def test_noalloc_capacity_and_empty_batch(pkg):
    capi = pkg.capi
    y = pkg.datasets.synthetic_yield()
    opts = pkg.datasets.uniform_1k(n=16)
    dev = capi.to_device(opts, y)
    assert _noalloc(pkg, dev, 65537) == capi.HWT_ENOMEM  # refused on the host, nothing launched
    assert b"65536" in capi.lib().hwt_last_error()
    assert _noalloc(pkg, dev, 0) == capi.HWT_OK
    pkg.devmem.synchronize()
