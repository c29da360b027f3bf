"""Host-side checks of the C ABI library (no GPU needed, no compute calls):
the library loads and exports every function include/hwtree.h declares, and
the host-only entry points behave."""
import ctypes as C
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared_functions(header="hwtree.h"):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hwt_\w+)\s*\(", text)))


def test_header_declares_the_api():
    names = _declared_functions()
    for must in ("hwt_price_f64", "hwt_plan_create", "hwt_plan_execute", "hwt_plan_sync",
                 "hwt_plan_destroy", "hwt_alphas_f64", "hwt_strerror"):
        assert must in names


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.capi.lib()
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    assert set(pkg.capi.EXPORTS) == set(_declared_functions())
    surr = _declared_functions("hwtsurr.h")
    missing = [n for n in surr if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    assert set(pkg.surrogate.EXPORTS) == set(surr)


def test_host_only_entry_points(pkg):
    lib = pkg.capi.lib()
    assert lib.hwt_abi_version() == 2
    assert lib.hwt_max_width(pkg.capi.FORM_WAVE) == 1344
    assert lib.hwt_strerror(0) == b"success"
    assert lib.hwt_strerror(-1) == b"invalid argument"
    assert lib.hwt_strerror(-2).startswith(b"unsupported")
    cfg = pkg.capi.config()
    assert cfg.form == 0 and cfg.precision == 0 and cfg.sort == ord("L") and cfg.device == -1


def test_struct_layouts_match_header(pkg):
    # 8-byte aligned pointers after an int32 count
    assert C.sizeof(pkg.capi.OptionsC) == 8 + 8 * 8
    assert C.sizeof(pkg.capi.YieldC) == 24
    assert C.sizeof(pkg.capi.ConfigC) == 9 * 4
    assert C.sizeof(pkg.capi.ConstantsC) == 5 * 8 + 3 * 4 + 2 + 1 + 1  # trailing pad
    assert C.sizeof(pkg.capi.ConstantsF32C) == 5 * 4 + 3 * 4 + 2 + 1 + 1
    assert C.sizeof(pkg.capi.TuneReportC) == 3 * 8 + 2 * 4
    text = (ROOT / "include" / "hwtree.h").read_text()
    for name, _ in pkg.capi.ConfigC._fields_:
        assert re.search(rf"int32_t\s+{name};", text), name
    assert C.sizeof(pkg.capi.StatsC) == 8 * 8 + 4 * 4 + 3 * 8
    assert C.sizeof(pkg.surrogate.SurrConfigC) == 10 * 4 + 8 + 6 * 4
    assert C.sizeof(pkg.surrogate.SurrStatsC) == 3 * 8 + 2 * 4 + 2 * 8 + 4 * 4


def test_surrogate_host_only_entry_points(pkg):
    lib = pkg.surrogate._lib()
    cfg = pkg.surrogate.config()
    assert cfg.members == 4 and cfg.batch == 8192 and abs(cfg.spread_threshold - 1e-3) < 1e-9
    assert lib.hwt_surr_create(None, 10, 10, 10, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_surr_train(None, 1, 1, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_surr_price(None, None, None, None, None, None) == pkg.capi.HWT_EINVAL
    lib.hwt_surr_destroy(None)  # no-op
    # parameter offsets of the header
    text = (ROOT / "include" / "hwtsurr.h").read_text()
    for name, off in pkg.surrogate.OFFSETS.items():
        assert re.search(rf"#define HWT_SURR_{name.upper()}\s+{off}\b", text), name
    sizes = {k: int(np.prod(v)) for k, v in pkg.surrogate.SHAPES.items()}
    assert sum(sizes.values()) == pkg.surrogate.NPARAMS == 199937


def test_null_arguments_rejected_without_gpu(pkg):
    lib = pkg.capi.lib()
    assert lib.hwt_price_f64(None, None, None, None, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_plan_create(None, 10, 10, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_price_f32(None, None, None, None, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_price_constants_f64(None, 1, None, None, None, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_price_constants_f32(None, 1, None, None, None, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_batch_create(None, None, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_batch_create_f32(None, None, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_batch_price(None, None, None, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_batch_device_memory(None) == 0
    lib.hwt_batch_destroy(None)  # no-op
    assert lib.hwt_plan_sync(None, None) == pkg.capi.HWT_EINVAL
    assert lib.hwt_autotune(None, None, None, None, None) == pkg.capi.HWT_EINVAL
    lib.hwt_plan_destroy(None)  # no-op
    assert b"hwt_" in lib.hwt_last_error()
