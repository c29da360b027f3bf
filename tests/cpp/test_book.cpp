// test_book.cpp -- the reference's only GPU test (test/TestCuda.cu:16-99),
// restated against this pricer's C++ drop-ins: Hull's book option (3-year put
// on a 9-year zero-coupon bond, strike 63, a = 0.1, sigma = 0.01) with 1..100
// steps per year, priced by every runner and compared with the sequential
// reference (the test-only CPU oracle).  The reference's own bar is
// Catch's Approx with margin 1000*eps (test/Compare.hpp:8-16); this test
// demands the FP64 bar |gpu - seq| <= 1e-10 * max(|seq|, 1).
//
// Also checks the runners' error behaviour: a cuda::multi runner given a
// block smaller than the widest tree throws std::invalid_argument, like the
// reference (cuda-multi/Kernel.cuh:395-400).
#include "hwtree/pricer.hpp"
#include "seq_oracle.h"

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

static std::vector<double> oracle_prices(const trinom::Options &o, const trinom::Yield &y)
{
    std::vector<double> out(o.N);
    orc_price_batch(o.N, o.StrikePrices.data(), o.Maturities.data(), o.Lengths.data(), o.TermUnits.data(),
                    o.TermStepCounts.data(), o.ReversionRates.data(), o.Volatilities.data(),
                    reinterpret_cast<const int8_t *>(o.Types.data()), y.Prices.data(), y.TimeSteps.data(), y.N,
                    out.data(), 1);
    return out;
}

static void compare(const std::string &what, const std::vector<double> &got, const std::vector<double> &gold)
{
    CHECK(got.size() == gold.size(), "%s: size", what.c_str());
    for (size_t i = 0; i < gold.size(); ++i) {
        const double tol = 1e-10 * std::fmax(std::fabs(gold[i]), 1.0);
        CHECK(std::fabs(got[i] - gold[i]) <= tol, "%s: option %zu got %.17g want %.17g", what.c_str(), i, got[i],
              gold[i]);
    }
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
        options.ReversionRates.push_back(0.1);
        options.Volatilities.push_back(0.01);
        options.Types.push_back(trinom::OptionType::PUT);
    }
    const std::vector<double> gold = oracle_prices(options, yield);

    using Runner = std::function<void(std::vector<double> &)>;
    const std::pair<std::string, Runner> runners[] = {
        {"seq::computeOptions", [&](std::vector<double> &r) { seq::computeOptions(options, yield, r); }},
        {"option version 1", [&](std::vector<double> &r) { cuda::option::KernelRunNaive k; k.run(options, yield, r, 64); }},
        {"option version 2", [&](std::vector<double> &r) { cuda::option::KernelRunCoalesced k; k.run(options, yield, r, 64); }},
        {"option version 3", [&](std::vector<double> &r) { cuda::option::KernelRunCoalescedGranular k(64); k.run(options, yield, r, 64); }},
        {"option version 4", [&](std::vector<double> &r) { cuda::option::KernelRunCoalescedGranular k(32); k.run(options, yield, r, 64); }},
        {"multi version 1", [&](std::vector<double> &r) { cuda::multi::KernelRunNaive k; k.run(options, yield, r, -1); }},
        {"multi version 2", [&](std::vector<double> &r) { cuda::multi::KernelRunCoalesced k; k.run(options, yield, r, -1); }},
        {"multi version 3", [&](std::vector<double> &r) { cuda::multi::KernelRunCoalescedBlock k; k.run(options, yield, r, -1, trinom::SortType::HEIGHT_DESC); }},
    };
    for (const auto &r : runners) {
        std::vector<double> results(options.N);
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
        std::vector<double> results(options.N);
        cuda::multi::KernelRunNaive k;
        k.run(options, yield, results, 256);  // 100 steps/yr gives width 371
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw, "multi runner with a too-small block did not throw");

    // undersized results vector is an argument error
    threw = false;
    try {
        std::vector<double> small(3);
        seq::computeOptions(options, yield, small);
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw, "undersized results did not throw");

    std::printf("%s: %d failure(s)\n", g_failures ? "FAIL" : "OK", g_failures);
    return g_failures ? 1 : 0;
}
