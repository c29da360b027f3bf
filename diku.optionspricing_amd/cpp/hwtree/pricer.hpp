// hwtree/pricer.hpp -- drop-in replacements for the reference's price-vector
// entry points, running on the MI355X through the C ABI (include/hwtree.h):
//
//   seq::computeOptions(options, yield, results)          reference seq/Seq.hpp:181-190
//   seq::computeSingleOption(constants, yield)             reference seq/Seq.hpp:24
//   cuda::CudaOptions (device-resident batch)              reference cuda/CudaDomain.cuh:88-221
//   cuda::option::KernelRun{Naive,Coalesced,CoalescedGranular}::run(...)
//                                                          reference cuda-option/Kernel.cuh:303-325
//   cuda::multi::KernelRun{Naive,Coalesced,CoalescedBlock}::run(...)
//                                                          reference cuda-multi/Kernel.cuh:374-413
//
// Same signatures and argument meaning: the caller pre-sizes `results` to N
// and gets prices in input order; blockSize <= 0 picks the default; SortType
// sets the processing order only.  `real` follows USE_DOUBLE as in the
// reference (trinom.hpp): FP32 callers go through hwt_price_f32.  The
// cuda::multi runners price with the on-chip wave form; the cuda::option
// runners with the fastest form for the batch unless the lane-per-option form
// is asked for (hwtree::option_runner_form()).  Failures throw:
// std::invalid_argument for bad inputs (as the reference does for missing
// files or a too-small block), std::runtime_error for device errors.
#pragma once

#include "../../../include/hwtree.h"
#include "trinom.hpp"

#include <chrono>
#include <cmath>
#include <algorithm>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

namespace hwtree {

using real_t = trinom::real;
constexpr bool kDouble = std::is_same<real_t, double>::value;
static_assert(kDouble || std::is_same<real_t, float>::value, "real is float or double");
static_assert(sizeof(trinom::OptionType) == 1, "OptionType is one byte");

template <class R>
struct Views;
template <>
struct Views<double> {
    typedef hwt_options_f64 opts;
    typedef hwt_yield_f64 yield;
};
template <>
struct Views<float> {
    typedef hwt_options_f32 opts;
    typedef hwt_yield_f32 yield;
};

inline typename Views<real_t>::opts view(const trinom::Options &o)
{
    return typename Views<real_t>::opts{o.N, o.StrikePrices.data(), o.Maturities.data(), o.Lengths.data(),
                                        o.TermUnits.data(), o.TermStepCounts.data(), o.ReversionRates.data(),
                                        o.Volatilities.data(), reinterpret_cast<const int8_t *>(o.Types.data())};
}

inline typename Views<real_t>::yield view(const trinom::Yield &y)
{
    return typename Views<real_t>::yield{y.N, y.Prices.data(), y.TimeSteps.data()};
}

inline void check_sizes(const trinom::Options &o, const trinom::Yield &y, const std::vector<trinom::real> &r)
{
    const size_t n = static_cast<size_t>(o.N);
    if (o.StrikePrices.size() != n || o.Maturities.size() != n || o.Lengths.size() != n || o.TermUnits.size() != n ||
        o.TermStepCounts.size() != n || o.ReversionRates.size() != n || o.Volatilities.size() != n ||
        o.Types.size() != n)
        throw std::invalid_argument("Options arrays do not all hold N elements.");
    if (y.N < 1 || y.Prices.size() != static_cast<size_t>(y.N) || y.TimeSteps.size() != static_cast<size_t>(y.N))
        throw std::invalid_argument("Yield curve arrays are empty or of unequal length.");
    if (r.size() < n) throw std::invalid_argument("results must be pre-sized to options.N.");
}

inline void raise(int rc)
{
    if (rc == HWT_OK) return;
    const std::string msg = std::string(hwt_strerror(rc)) + ": " + hwt_last_error();
    if (rc == HWT_EINVAL || rc == HWT_EUNSUPPORTED) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// HIP device the drop-ins price on (the reference selects it with
// cudaSetDevice before calling, CudaOptionMain.cu:104); -1 = the current one.
inline int &default_device()
{
    static int device = -1;
    return device;
}

// Kernel form of the cuda::option runners: HWT_FORM_AUTO (the default, the
// fastest form for the batch) or HWT_FORM_LANE (one option per lane, tree
// columns in HBM: the reference's cuda-option layout, far slower on batches
// with wide trees).
inline int &option_runner_form()
{
    static int form = HWT_FORM_AUTO;
    return form;
}

inline hwt_config make_config(int form, int blockSize, trinom::SortType sort)
{
    hwt_config cfg;
    hwt_config_init(&cfg);
    cfg.form = form;
    cfg.device = default_device();
    cfg.block_size = blockSize;
    cfg.sort = sort == trinom::SortType::NONE ? HWT_SORT_WORK_DESC : static_cast<int>(static_cast<char>(sort));
    return cfg;
}

inline int price_views(const hwt_options_f64 &o, const hwt_yield_f64 &y, double *out, const hwt_config &c,
                       hwt_stats &st)
{
    return hwt_price_f64(&o, &y, out, &c, &st);
}

inline int price_views(const hwt_options_f32 &o, const hwt_yield_f32 &y, float *out, const hwt_config &c,
                       hwt_stats &st)
{
    return hwt_price_f32(&o, &y, out, &c, &st);
}

inline hwt_stats price(const trinom::Options &o, const trinom::Yield &y, std::vector<trinom::real> &results,
                       int form, int blockSize, trinom::SortType sort)
{
    check_sizes(o, y, results);
    const hwt_config cfg = make_config(form, blockSize, sort);
    hwt_stats st{};
    raise(price_views(view(o), view(y), results.data(), cfg, st));
    return st;
}

inline std::vector<double> widen(const std::vector<trinom::real> &v) { return std::vector<double>(v.begin(), v.end()); }

// One option from its constants; with R = float the exercise step (int)(t / dt)
// is taken in float, as the FP32 reference takes it.
template <class R>
R price_single(const trinom::OptionConstants &c, const trinom::Yield &yield)
{
    const hwt_config cfg = make_config(HWT_FORM_AUTO, -1, trinom::SortType::NONE);
    const R t = static_cast<R>(c.t), dt = static_cast<R>(c.dt), dr = static_cast<R>(c.dr), X = static_cast<R>(c.X),
            M = static_cast<R>(c.M);
    const R *rates = reinterpret_cast<const R *>(yield.Prices.data());
    R price = 0;
    if constexpr (std::is_same<R, double>::value) {
        const hwt_option_constants_f64 k{t, dt, dr, X, M, c.jmax, c.n, c.width, c.termUnit, static_cast<int8_t>(c.type)};
        const hwt_yield_f64 y{yield.N, rates, yield.TimeSteps.data()};
        raise(hwt_price_constants_f64(&k, 1, &y, &price, &cfg, nullptr));
    } else {
        const hwt_option_constants_f32 k{t, dt, dr, X, M, c.jmax, c.n, c.width, c.termUnit, static_cast<int8_t>(c.type)};
        const hwt_yield_f32 y{yield.N, rates, yield.TimeSteps.data()};
        raise(hwt_price_constants_f32(&k, 1, &y, &price, &cfg, nullptr));
    }
    return price;
}

} // namespace hwtree

namespace seq {

// Prices every option of `options` into results[0..N) (pre-sized by the caller).
inline void computeOptions(const trinom::Options &options, const trinom::Yield &yield,
                           std::vector<trinom::real> &results)
{
    hwtree::price(options, yield, results, HWT_FORM_AUTO, -1, trinom::SortType::NONE);
}

// Prices one option given by its tree constants (seq/Seq.hpp:24): the tree is
// c.width nodes wide and c.n steps high, with the exercise step (int)(c.t/c.dt).
inline trinom::real computeSingleOption(const trinom::OptionConstants &c, const trinom::Yield &yield)
{
    if (yield.N < 1 || yield.Prices.size() != static_cast<size_t>(yield.N) ||
        yield.TimeSteps.size() != static_cast<size_t>(yield.N))
        throw std::invalid_argument("Yield curve arrays are empty or of unequal length.");
    return hwtree::price_single<trinom::real>(c, yield);
}

} // namespace seq

namespace cuda {

// Microseconds, as the reference's runners report them.
struct CudaRuntime {
    long KernelRuntime = std::numeric_limits<long>::max();
    long TotalRuntime = std::numeric_limits<long>::max();
    long DeviceMemory = 0;
};

inline bool operator<(const CudaRuntime &x, const CudaRuntime &y)
{
    return std::tie(x.KernelRuntime, x.TotalRuntime) < std::tie(y.KernelRuntime, y.TotalRuntime);
}

// A batch uploaded to the device once (cuda/CudaDomain.cuh:88-221).  The
// reference sorts the device copy and unsorts the results; here the sort
// order is a processing hint and results always come back in input order, so
// sortResult has nothing to do.
class CudaOptions {
public:
    const int N;
    const int YieldSize;
    long DeviceMemory = 0;

    CudaOptions(const trinom::Options &options, const trinom::Yield &yield)
        : N(options.N), YieldSize(yield.N)
    {
        std::vector<trinom::real> probe(static_cast<size_t>(options.N));
        hwtree::check_sizes(options, yield, probe);
        hwtree::raise(create(hwtree::view(options), hwtree::view(yield)));
    }
    CudaOptions(const CudaOptions &) = delete;
    CudaOptions &operator=(const CudaOptions &) = delete;
    ~CudaOptions() { hwt_batch_destroy(batch_); }

    // The batch is complete once constructed: the device computes every
    // tree's width and height when it prices.
    void initialize() { DeviceMemory = static_cast<long>(hwt_batch_device_memory(batch_)); }
    void sortOptions(const trinom::SortType sort, const bool isTest)
    {
        sort_ = sort;
        if (isTest && sort != trinom::SortType::NONE)
            std::cout << "Processing order " << static_cast<char>(sort) << " (results stay in input order)"
                      << std::endl;
    }
    void sortResult() {}

    trinom::SortType sortType() const { return sort_; }
    hwt_batch *handle() { return batch_; }

private:
    int create(const hwt_options_f64 &o, const hwt_yield_f64 &y) { return hwt_batch_create(&batch_, &o, &y); }
    int create(const hwt_options_f32 &o, const hwt_yield_f32 &y) { return hwt_batch_create_f32(&batch_, &o, &y); }

    hwt_batch *batch_ = nullptr;
    trinom::SortType sort_ = trinom::SortType::NONE;
};

namespace detail {

class RunnerBase {
public:
    CudaRuntime runtime;

    void run(const trinom::Options &options, const trinom::Yield &yield, std::vector<trinom::real> &results,
             int blockSize = -1, trinom::SortType sortType = trinom::SortType::NONE, bool isTest = false)
    {
        preflight(options, blockSize);
        const auto t0 = std::chrono::steady_clock::now();
        const hwt_stats st = hwtree::price(options, yield, results, form(), blockSize, sortType);
        const auto t1 = std::chrono::steady_clock::now();
        report(st, options.N, t0, t1, isTest);
    }

    // The overload over an uploaded batch (cuda-option/Kernel.cuh:316-325):
    // timing starts at timeBegin, as in the reference.
    void run(CudaOptions &cudaOptions, std::vector<trinom::real> &results,
             const std::chrono::time_point<std::chrono::steady_clock> timeBegin, int blockSize = -1,
             trinom::SortType sortType = trinom::SortType::NONE, bool isTest = false)
    {
        if (results.size() < static_cast<size_t>(cudaOptions.N))
            throw std::invalid_argument("results must be pre-sized to options.N.");
        cudaOptions.sortOptions(sortType, isTest);
        const hwt_config cfg = hwtree::make_config(form(), blockSize, sortType);
        hwt_stats st{};
        std::vector<double> out(static_cast<size_t>(cudaOptions.N));
        hwtree::raise(hwt_batch_price(cudaOptions.handle(), out.data(), &cfg, &st));
        std::copy(out.begin(), out.end(), results.begin());
        report(st, cudaOptions.N, timeBegin, std::chrono::steady_clock::now(), isTest);
        runtime.DeviceMemory += static_cast<long>(hwt_batch_device_memory(cudaOptions.handle()));
    }

protected:
    explicit RunnerBase(int form) : form_(form) {}
    virtual ~RunnerBase() = default;
    virtual void preflight(const trinom::Options &, int) {}
    virtual int form() const { return form_; }
    int form_;

private:
    void report(const hwt_stats &st, int n, std::chrono::steady_clock::time_point t0,
                std::chrono::steady_clock::time_point t1, bool isTest)
    {
        runtime.KernelRuntime = static_cast<long>(st.kernel_ms * 1000.0);
        runtime.TotalRuntime = static_cast<long>(st.total_ms * 1000.0);
        runtime.DeviceMemory = static_cast<long>(st.device_memory);
        if (isTest) {
            std::cout << "Priced " << n << (hwtree::kDouble ? " double" : " float") << " options (" << st.tasks
                      << " tree tasks, " << st.node_updates << " node updates) in "
                      << std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count()
                      << " microsec including transfers; kernels " << runtime.KernelRuntime << " microsec, total "
                      << runtime.TotalRuntime << " microsec" << std::endl;
        }
    }
};

} // namespace detail

namespace option {

// The reference's cuda-option runners.  They price with the fastest form for
// the batch unless hwtree::option_runner_form() asks for one option per lane;
// the reference's memory layouts (packed, globally padded, block/warp-granular
// padding) all map to the same lane-interleaved kernel, and granularity is
// accepted for compatibility.
class KernelRunBase : public detail::RunnerBase {
protected:
    KernelRunBase() : RunnerBase(HWT_FORM_LANE) {}
    int form() const override { return hwtree::option_runner_form(); }
};
class KernelRunNaive : public KernelRunBase {};
class KernelRunCoalesced : public KernelRunBase {};
class KernelRunCoalescedGranular : public KernelRunBase {
public:
    explicit KernelRunCoalescedGranular(int granularity) : granularity_(granularity) {}

private:
    int granularity_;
};

} // namespace option

namespace multi {

// Many options per workgroup with on-chip tree columns (the reference's
// cuda-multi runners).  Like the reference, an explicit blockSize smaller than
// the widest tree is rejected (cuda-multi/Kernel.cuh:386-400); blockSize <= 0
// accepts any width this pricer supports.
class KernelRunBase : public detail::RunnerBase {
protected:
    KernelRunBase() : RunnerBase(HWT_FORM_WAVE) {}
    void preflight(const trinom::Options &o, int blockSize) override
    {
        if (blockSize <= 0) return;
        int maxWidth = 0;
        for (int i = 0; i < o.N; ++i) {
            const int tuy = static_cast<int>(std::ceil(365.0 / o.TermUnits[i]));
            const double dt = tuy / static_cast<double>(o.TermStepCounts[i]);
            const double M = std::exp(-o.ReversionRates[i] * dt) - 1.0;
            const int w = 2 * (static_cast<int>(-0.184 / M) + 1) + 1;
            maxWidth = w > maxWidth ? w : maxWidth;
        }
        if (maxWidth > blockSize)
            throw std::invalid_argument("Block size (" + std::to_string(blockSize) +
                                        ") cannot be smaller than max option width (" + std::to_string(maxWidth) +
                                        ").");
    }
};
class KernelRunNaive : public KernelRunBase {};
class KernelRunCoalesced : public KernelRunBase {};
class KernelRunCoalescedBlock : public KernelRunBase {};

} // namespace multi

} // namespace cuda
