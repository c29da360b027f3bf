// test_book.cpp -- the reference's only GPU test (test/TestCuda.cu:16-99),
// restated against this pricer's C++ drop-ins: Hull's book option (3-year put
// on a 9-year zero-coupon bond, strike 63, a = 0.1, sigma = 0.01) with 1..100
// steps per year, priced by every runner and compared with the sequential
// reference (the test-only CPU oracle).
//
// Built twice, like the reference: with USE_DOUBLE (real = double) the bar is
// the FP64 one, |gpu - seq| <= 1e-10 * max(|seq|, 1); without it (real =
// float, the reference's default) the oracle prices the FP32 build's trees
// (constants and exercise step derived in float: at 15, 30, 60, 63, 75 and
// 77 steps a year it exercises a step earlier than FP64) and the bar is the
// reference's own, Catch's Approx with margin 1000 * FLT_EPSILON
// (test/Compare.hpp:8-16), which also accepts a relative difference of
// 100 * FLT_EPSILON.
//
// Also covers seq::computeSingleOption on OptionConstants, the
// run(CudaOptions&, ...) overload, and the runners' error behaviour: a
// cuda::multi runner given a block smaller than the widest tree throws
// std::invalid_argument, like the reference (cuda-multi/Kernel.cuh:395-400).
#include "hwtree/pricer.hpp"
#include "seq_oracle.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>

static int g_failures = 0;

#define CHECK(cond, ...)                                    \
    do {                                                    \
        if (!(cond)) {                                      \
            ++g_failures;                                   \
            std::fprintf(stderr, "FAILED %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);              \
            std::fprintf(stderr, "\n");                     \
        }                                                   \
    } while (0)

using trinom::real;

// The sequential reference in the caller's width of `real`: with USE_DOUBLE
// the FP64 seq restatement; without it the FP32 build's trees (constants and
// exercise step derived in float) evaluated in double.
static std::vector<double> oracle_prices(const trinom::Options &o, const trinom::Yield &y)
{
    if (!hwtree::kDouble) {
        const std::vector<double> r(y.Prices.begin(), y.Prices.end());
        std::vector<double> out(o.N);
        for (int i = 0; i < o.N; ++i) {
            orc_constants c;
            orc_option_constants_f32(o.StrikePrices[i], o.Maturities[i], o.Lengths[i], o.TermUnits[i],
                                     o.TermStepCounts[i], o.ReversionRates[i], o.Volatilities[i],
                                     static_cast<int8_t>(o.Types[i]), &c);
            out[i] = orc_price_single_f32step(&c, r.data(), y.TimeSteps.data(), y.N);
        }
        return out;
    }
    const std::vector<double> s(o.StrikePrices.begin(), o.StrikePrices.end()),
        m(o.Maturities.begin(), o.Maturities.end()), l(o.Lengths.begin(), o.Lengths.end()),
        a(o.ReversionRates.begin(), o.ReversionRates.end()), v(o.Volatilities.begin(), o.Volatilities.end()),
        r(y.Prices.begin(), y.Prices.end());
    std::vector<double> out(o.N);
    orc_price_batch(o.N, s.data(), m.data(), l.data(), o.TermUnits.data(), o.TermStepCounts.data(), a.data(),
                    v.data(), reinterpret_cast<const int8_t *>(o.Types.data()), r.data(), y.TimeSteps.data(), y.N,
                    out.data(), 1);
    return out;
}

static bool within_bar(double got, double want)
{
    if (hwtree::kDouble) return std::fabs(got - want) <= 1e-10 * std::fmax(std::fabs(want), 1.0);
    const double eps = FLT_EPSILON, d = std::fabs(got - want);
    return d <= 1000 * eps || d <= 100 * eps * (1.0 + std::fmax(std::fabs(got), std::fabs(want)));
}

static void compare(const std::string &what, const std::vector<real> &got, const std::vector<double> &gold)
{
    CHECK(got.size() == gold.size(), "%s: size", what.c_str());
    for (size_t i = 0; i < gold.size(); ++i)
        CHECK(within_bar(got[i], gold[i]), "%s: option %zu got %.17g want %.17g", what.c_str(), i, (double)got[i],
              gold[i]);
}

int main(int argc, char **argv)
{
    const std::string yield_path = argc > 1 ? argv[1] : "tests/golden/yield.in";
    const trinom::Yield yield(yield_path);

    trinom::Options options(100);
    for (int i = 1; i <= options.N; ++i) {
        options.Lengths.push_back(3);
        options.Maturities.push_back(9);
        options.StrikePrices.push_back(63);
        options.TermUnits.push_back(365);
        options.TermStepCounts.push_back(static_cast<uint16_t>(i));
        options.ReversionRates.push_back(static_cast<real>(0.1));
        options.Volatilities.push_back(static_cast<real>(0.01));
        options.Types.push_back(trinom::OptionType::PUT);
    }
    const std::vector<double> gold = oracle_prices(options, yield);

    using Runner = std::function<void(std::vector<real> &)>;
    const std::pair<std::string, Runner> runners[] = {
        {"seq::computeOptions", [&](std::vector<real> &r) { seq::computeOptions(options, yield, r); }},
        {"option version 1", [&](std::vector<real> &r) { cuda::option::KernelRunNaive k; k.run(options, yield, r, 64); }},
        {"option version 2", [&](std::vector<real> &r) { cuda::option::KernelRunCoalesced k; k.run(options, yield, r, 64); }},
        {"option version 3", [&](std::vector<real> &r) { cuda::option::KernelRunCoalescedGranular k(64); k.run(options, yield, r, 64); }},
        {"option version 4", [&](std::vector<real> &r) { cuda::option::KernelRunCoalescedGranular k(32); k.run(options, yield, r, 64); }},
        {"option, lane form", [&](std::vector<real> &r) {
             hwtree::option_runner_form() = HWT_FORM_LANE;
             cuda::option::KernelRunCoalesced k;
             k.run(options, yield, r, 128);
             hwtree::option_runner_form() = HWT_FORM_AUTO;
         }},
        {"multi version 1", [&](std::vector<real> &r) { cuda::multi::KernelRunNaive k; k.run(options, yield, r, -1); }},
        {"multi version 2", [&](std::vector<real> &r) { cuda::multi::KernelRunCoalesced k; k.run(options, yield, r, -1); }},
        {"multi version 3", [&](std::vector<real> &r) { cuda::multi::KernelRunCoalescedBlock k; k.run(options, yield, r, -1, trinom::SortType::HEIGHT_DESC); }},
        {"run(CudaOptions&)", [&](std::vector<real> &r) {
             cuda::CudaOptions co(options, yield);
             const auto t0 = std::chrono::steady_clock::now();
             co.initialize();
             cuda::multi::KernelRunCoalesced k;
             k.run(co, r, t0, -1, trinom::SortType::WIDTH_DESC);
             CHECK(k.runtime.DeviceMemory > 0 && k.runtime.TotalRuntime >= 0, "CudaOptions runtime");
             std::vector<real> again(options.N);
             cuda::option::KernelRunNaive k2;  // a second runner on the same uploaded batch
             k2.run(co, again, t0);
             CHECK(again == r, "second run on the same CudaOptions differs");
         }},
        {"seq::computeSingleOption", [&](std::vector<real> &r) {
             for (int i = 0; i < options.N; ++i) r[i] = seq::computeSingleOption(trinom::OptionConstants(options, i), yield);
         }},
    };
    for (const auto &r : runners) {
        std::vector<real> results(options.N);
        try {
            r.second(results);
            compare(r.first, results, gold);
        } catch (const std::exception &e) {
            CHECK(false, "%s threw: %s", r.first.c_str(), e.what());
        }
    }

    // the reference rejects a block narrower than the widest tree
    bool threw = false;
    try {
        std::vector<real> results(options.N);
        cuda::multi::KernelRunNaive k;
        k.run(options, yield, results, 256);  // 100 steps/yr gives width 371
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw, "multi runner with a too-small block did not throw");

    // undersized results vector is an argument error
    threw = false;
    try {
        std::vector<real> small(3);
        seq::computeOptions(options, yield, small);
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw, "undersized results did not throw");

    // OptionConstants bounds-checks its index like the reference (.at)
    threw = false;
    try {
        trinom::OptionConstants c(options, options.N);
        (void)c;
    } catch (const std::out_of_range &) {
        threw = true;
    }
    CHECK(threw, "OptionConstants past the end did not throw");

    std::printf("%s: %d failure(s) (%s)\n", g_failures ? "FAIL" : "OK", g_failures,
                hwtree::kDouble ? "real = double" : "real = float");
    return g_failures ? 1 : 0;
}
