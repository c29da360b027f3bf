// hwtree/pricer.hpp -- drop-in replacements for the reference's price-vector
// entry points, running on the MI355X through the C ABI (include/hwtree.h):
//
//   seq::computeOptions(options, yield, results)          reference seq/Seq.hpp:181-190
//   cuda::option::KernelRun{Naive,Coalesced,CoalescedGranular}::run(...)
//                                                          reference cuda-option/Kernel.cuh:303-325
//   cuda::multi::KernelRun{Naive,Coalesced,CoalescedBlock}::run(...)
//                                                          reference cuda-multi/Kernel.cuh:374-413
//
// Same signatures and argument meaning: the caller pre-sizes `results` to N
// and gets prices in input order; blockSize <= 0 picks the default; SortType
// sets the processing order only.  The cuda::option runners price with the
// lane-per-option kernel form, the cuda::multi runners with the on-chip
// wave form.  Failures throw: std::invalid_argument for bad inputs (as the
// reference does for missing files or a too-small block), std::runtime_error
// for device errors.
#pragma once

#include "../../../include/hwtree.h"
#include "trinom.hpp"

#include <chrono>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace hwtree {

inline hwt_options_f64 view(const trinom::Options &o)
{
    static_assert(sizeof(trinom::OptionType) == 1, "OptionType is one byte");
    return hwt_options_f64{o.N, o.StrikePrices.data(), o.Maturities.data(), o.Lengths.data(), o.TermUnits.data(),
                           o.TermStepCounts.data(), o.ReversionRates.data(), o.Volatilities.data(),
                           reinterpret_cast<const int8_t *>(o.Types.data())};
}

inline hwt_yield_f64 view(const trinom::Yield &y) { return hwt_yield_f64{y.N, y.Prices.data(), y.TimeSteps.data()}; }

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

inline hwt_stats price(const trinom::Options &o, const trinom::Yield &y, std::vector<trinom::real> &results,
                       int form, int blockSize, trinom::SortType sort)
{
    check_sizes(o, y, results);
    hwt_config cfg;
    hwt_config_init(&cfg);
    cfg.form = form;
    cfg.device = default_device();
    cfg.block_size = blockSize;
    cfg.sort = sort == trinom::SortType::NONE ? HWT_SORT_WORK_DESC : static_cast<int>(static_cast<char>(sort));
    hwt_stats st{};
    const hwt_options_f64 ov = view(o);
    const hwt_yield_f64 yv = view(y);
    raise(hwt_price_f64(&ov, &yv, results.data(), &cfg, &st));
    return st;
}

} // namespace hwtree

namespace seq {

// Prices every option of `options` into results[0..N) (pre-sized by the caller).
inline void computeOptions(const trinom::Options &options, const trinom::Yield &yield,
                           std::vector<trinom::real> &results)
{
    hwtree::price(options, yield, results, HWT_FORM_AUTO, -1, trinom::SortType::NONE);
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

namespace detail {

class RunnerBase {
public:
    CudaRuntime runtime;

    void run(const trinom::Options &options, const trinom::Yield &yield, std::vector<trinom::real> &results,
             int blockSize = -1, trinom::SortType sortType = trinom::SortType::NONE, bool isTest = false)
    {
        preflight(options, blockSize);
        const auto t0 = std::chrono::steady_clock::now();
        const hwt_stats st = hwtree::price(options, yield, results, form_, blockSize, sortType);
        const auto t1 = std::chrono::steady_clock::now();
        runtime.KernelRuntime = static_cast<long>(st.kernel_ms * 1000.0);
        runtime.TotalRuntime = static_cast<long>(st.total_ms * 1000.0);
        runtime.DeviceMemory = static_cast<long>(st.device_memory);
        if (isTest) {
            std::cout << "Priced " << options.N << " double options (" << st.tasks << " tree tasks, "
                      << st.node_updates << " node updates) in "
                      << std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count()
                      << " microsec including transfers; kernels " << runtime.KernelRuntime << " microsec, total "
                      << runtime.TotalRuntime << " microsec" << std::endl;
        }
    }

protected:
    explicit RunnerBase(int form) : form_(form) {}
    virtual ~RunnerBase() = default;
    virtual void preflight(const trinom::Options &, int) {}
    int form_;
};

} // namespace detail

namespace option {

// One option per lane (the reference's cuda-option runners).  The reference's
// memory layouts (packed, globally padded, block/warp-granular padding) map to
// the same lane-interleaved kernel; granularity is accepted for compatibility.
class KernelRunBase : public detail::RunnerBase {
protected:
    KernelRunBase() : RunnerBase(HWT_FORM_LANE) {}
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
