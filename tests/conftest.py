import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _load_pkg():
    if "optionspricing_amd" in sys.modules:
        return sys.modules["optionspricing_amd"]
    pkg_dir = ROOT / "diku.optionspricing_amd"
    spec = importlib.util.spec_from_file_location(
        "optionspricing_amd", pkg_dir / "__init__.py", submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["optionspricing_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return _load_pkg()


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    return orc


@pytest.fixture(scope="session")
def golden_yield(pkg):
    return pkg.formats.read_yield(GOLDEN / "yield.in")


GOLDEN_SETS = ["book", "options-1000", "options-60000_s60", "0_UNIFORM_1000", "1_RAND_1000",
               "2_RANDCONSTHEIGHT_1000", "3_RANDCONSTWIDTH_1000", "4_SKEWED_s64",
               "5_SKEWEDCONSTHEIGHT_s64", "6_SKEWEDCONSTWIDTH_s64"]


def parity_tol(ref):
    """FP64 parity bar: 1e-10 relative, with an absolute floor of 1e-10 per
    100 notional for prices below 1 (deep out-of-the-money options)."""
    import numpy as np
    return 1e-10 * np.maximum(np.abs(ref), 1.0)
