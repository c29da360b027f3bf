"""The C++ drop-ins (hwtree/pricer.hpp) and CLI drivers (reference flags and
text output)."""
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, parity_tol

BIN = ROOT / "diku.optionspricing_amd" / "bin"
TEST_BOOK = ROOT / "tests" / "cpp" / "test_book"
TEST_BOOK32 = ROOT / "tests" / "cpp" / "test_book32"
FLT_EPS = float(np.finfo(np.float32).eps)


def catch_approx_ok(got, want):
    """The reference's FP32 comparison (test/Compare.hpp:8-16): Catch's Approx
    with margin 1000 * FLT_EPSILON, which also passes a difference within
    100 * FLT_EPSILON of (1 + the larger magnitude)."""
    d = np.abs(got - want)
    return (d <= 1000 * FLT_EPS) | (d <= 100 * FLT_EPS * (1.0 + np.maximum(np.abs(got), np.abs(want))))


def _need(p):
    if not p.exists():
        pytest.skip(f"{p} not built (run __graft_entry__.build())")


def test_cli_missing_file_is_an_error():
    _need(BIN / "hwt-seq")
    r = subprocess.run([str(BIN / "hwt-seq"), "-o", "/nonexistent.in", "-y", str(GOLDEN / "yield.in")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "does not exist" in r.stderr


def test_cli_bad_flag_is_an_error():
    _need(BIN / "hwt-multi")
    r = subprocess.run([str(BIN / "hwt-multi"), "-q"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown flag" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("exe", [TEST_BOOK, TEST_BOOK32])
def test_cpp_book_options_all_runners(exe):
    """Every C++ drop-in (both widths of `real`), including
    seq::computeSingleOption on OptionConstants and the run(CudaOptions&)
    overload, against the oracle."""
    _need(exe)
    r = subprocess.run([str(exe), str(GOLDEN / "yield.in")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("exe,extra", [("hwt-seq", []), ("hwt-option", ["-v", "2"]),
                                       ("hwt-multi", ["-v", "3", "-s", "H"])])
def test_cli_matches_golden_file(pkg, exe, extra):
    """The reference's compare.sh flow: price a data file with the CLI and
    compare the printed array with the golden output."""
    _need(BIN / exe)
    r = subprocess.run([str(BIN / exe), "-o", str(GOLDEN / "4_SKEWED_s64.in"), "-y", str(GOLDEN / "yield.in"), *extra],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([float(t) for t in r.stdout.strip()[1:-1].split(",")])
    gold = pkg.formats.read_results(GOLDEN / "4_SKEWED_s64.out")
    assert np.all(np.abs(got - gold) <= parity_tol(gold) + 1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("exe,extra", [("hwt-seq32", []), ("hwt-multi32", ["-v", "3", "-s", "H"])])
@pytest.mark.parametrize("name", ["book", "1_RAND_1000", "4_SKEWED_s64"])
def test_cli32_matches_fp32_golden_file(pkg, exe, extra, name):
    """The FP32 builds of the drivers against the reference's FP32 seq output
    (data/out32), at the reference's own FP32 bar; the printed array has 9
    decimals, as the reference prints floats."""
    _need(BIN / exe)
    r = subprocess.run([str(BIN / exe), "-o", str(GOLDEN / f"{name}.in"), "-y", str(GOLDEN / "yield.in"), *extra],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    toks = r.stdout.strip()[1:-1].split(",")
    assert all(len(t.strip().split(".")[1]) == 9 for t in toks)
    got = np.array([float(t) for t in toks])
    gold = pkg.formats.read_results(GOLDEN / f"{name}.out32")
    ok = catch_approx_ok(got, gold)
    assert ok.all(), (np.nonzero(~ok)[0][:5], got[~ok][:5], gold[~ok][:5])


@pytest.mark.gpu
def test_cli_runs_csv():
    _need(BIN / "hwt-multi")
    r = subprocess.run([str(BIN / "hwt-multi"), "-o", str(GOLDEN / "book.in"), "-y", str(GOLDEN / "yield.in"),
                        "-r", "2", "-s", "-", "W", "-v", "1", "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split(",") for l in r.stdout.strip().splitlines()]
    assert len(rows) == 4 and all(len(x) == 6 for x in rows)
    assert [x[0] for x in rows] == ["1", "1", "3", "3"] and [x[2] for x in rows] == ["-", "W", "-", "W"]
