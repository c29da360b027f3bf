// cli.cpp -- command-line drivers with the reference's flags and output, so
// its benchmark and golden-file scripts can run against this pricer:
//
//   hwt-seq     replaces Seq          (seq/SeqMain.cpp:12-81)
//   hwt-option  replaces CudaOption   (cuda-option/CudaOptionMain.cu:10-119)
//   hwt-multi   replaces CudaMulti    (cuda-multi/CudaMultiMain.cu:10-113)
// each built with USE_DOUBLE (REAL=64) and without it (REAL=32, the "32"
// binaries, which read and print floats with 9 decimals like the reference).
//
// Flags (common/Args.hpp:38-66): -o/--options FILE, -y/--yield FILE,
// -s/--sort {-,w,W,h,H}..., -v/--version N..., -b/--block N..., -r/--runs N,
// -d/--device N, -t/--test.  Multi-valued flags take every following token
// that is not another flag ("-s - W H").
//
// Output: without -r, the prices as one "[p0, p1, ...]" line (17 decimals);
// with -r, the best of the runs as CSV: "-,-,-,-,<total us>,-" for hwt-seq,
// "<version>,<block>,<sort>,<kernel us>,<total us>,<device bytes>" otherwise;
// -t switches to human-readable timing lines.
#include "hwtree/pricer.hpp"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#ifndef HWT_CLI_MODE
#define HWT_CLI_MODE 0  // 0 seq, 1 option, 2 multi
#endif

namespace {

struct Args {
    std::string options, yield;
    std::vector<trinom::SortType> sorts;
    std::vector<int> blocks, versions;
    int runs = 0, device = 0;
    bool test = false;
};

bool is_flag(const std::string &s)
{
    if (s.size() < 2 || s[0] != '-') return false;
    if (s[1] == '-') return s.size() > 2;
    return std::isalpha(static_cast<unsigned char>(s[1])) != 0;
}

Args parse(int argc, char **argv)
{
    static const std::map<std::string, char> longs = {{"options", 'o'}, {"yield", 'y'}, {"sort", 's'},
                                                      {"version", 'v'}, {"runs", 'r'},  {"block", 'b'},
                                                      {"device", 'd'},  {"test", 't'}};
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string tok = argv[i];
        char f = 0;
        if (tok.rfind("--", 0) == 0) {
            auto it = longs.find(tok.substr(2));
            if (it == longs.end()) throw std::invalid_argument("unknown flag " + tok);
            f = it->second;
        } else if (is_flag(tok) && tok.size() == 2) {
            f = tok[1];
        } else {
            throw std::invalid_argument("unexpected argument " + tok);
        }
        std::vector<std::string> vals;
        while (i + 1 < argc && !is_flag(argv[i + 1])) vals.push_back(argv[++i]);
        auto one = [&]() -> std::string {
            if (vals.empty()) throw std::invalid_argument(std::string("flag -") + f + " needs a value");
            return vals.front();
        };
        switch (f) {
        case 'o': a.options = one(); break;
        case 'y': a.yield = one(); break;
        case 'r': a.runs = std::atoi(one().c_str()); break;
        case 'd': a.device = std::atoi(one().c_str()); break;
        case 't': a.test = true; break;
        case 'v': for (auto &v : vals) a.versions.push_back(std::atoi(v.c_str())); break;
        case 'b': for (auto &v : vals) a.blocks.push_back(std::atoi(v.c_str())); break;
        case 's':
            for (auto &v : vals)
                if (v.size() == 1 && std::strchr("-wWhH", v[0])) a.sorts.push_back(static_cast<trinom::SortType>(v[0]));
            break;
        default: throw std::invalid_argument("unknown flag " + tok);
        }
    }
    if (a.sorts.empty()) a.sorts.push_back(trinom::SortType::NONE);
    if (a.blocks.empty()) a.blocks.push_back(-1);
    if (a.versions.empty()) a.versions.push_back(1);
    return a;
}

cuda::CudaRuntime run_once(const trinom::Options &o, const trinom::Yield &y, std::vector<trinom::real> &res,
                           int version, int block, trinom::SortType sort, bool test)
{
#if HWT_CLI_MODE == 1
    switch (version) {
    case 1: { cuda::option::KernelRunNaive k; k.run(o, y, res, block, sort, test); return k.runtime; }
    case 2: { cuda::option::KernelRunCoalesced k; k.run(o, y, res, block, sort, test); return k.runtime; }
    case 3: { cuda::option::KernelRunCoalescedGranular k(block); k.run(o, y, res, block, sort, test); return k.runtime; }
    case 4: { cuda::option::KernelRunCoalescedGranular k(32); k.run(o, y, res, block, sort, test); return k.runtime; }
    }
#elif HWT_CLI_MODE == 2
    switch (version) {
    case 1: { cuda::multi::KernelRunNaive k; k.run(o, y, res, block, sort, test); return k.runtime; }
    case 2: { cuda::multi::KernelRunCoalesced k; k.run(o, y, res, block, sort, test); return k.runtime; }
    case 3: { cuda::multi::KernelRunCoalescedBlock k; k.run(o, y, res, block, sort, test); return k.runtime; }
    }
#else
    (void)version; (void)block; (void)sort; (void)test;
    const auto t0 = std::chrono::steady_clock::now();
    seq::computeOptions(o, y, res);
    cuda::CudaRuntime r;
    r.TotalRuntime = std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
    r.KernelRuntime = r.TotalRuntime;
    return r;
#endif
    return cuda::CudaRuntime();
}

} // namespace

int main(int argc, char **argv)
{
    try {
        const Args args = parse(argc, argv);
        if (args.test) std::cout << "Loading options " << args.options << ", yield " << args.yield << std::endl;
        const trinom::Options options(args.options);
        const trinom::Yield yield(args.yield);
        hwtree::default_device() = args.device;
#if HWT_CLI_MODE == 1
        // this driver times the one-option-per-lane kernel, as CudaOption
        // times the reference's one-option-per-thread kernel
        hwtree::option_runner_form() = HWT_FORM_LANE;
#endif
        for (int version : args.versions)
            for (int block : args.blocks)
                for (trinom::SortType sort : args.sorts) {
                    if (args.runs > 0) {
                        cuda::CudaRuntime best;
                        for (int r = 0; r < args.runs; ++r) {
                            std::vector<trinom::real> res(options.N);
                            const auto rt = run_once(options, yield, res, version, block, sort, args.test);
                            if (rt < best) best = rt;
                        }
                        if (args.test)
                            std::cout << "Best times: kernel " << best.KernelRuntime << " microsec, total "
                                      << best.TotalRuntime << " microsec." << std::endl;
                        else if (HWT_CLI_MODE == 0)
                            std::cout << "-,-,-,-," << best.TotalRuntime << ",-" << std::endl;
                        else
                            std::cout << version << ',' << block << ',' << static_cast<char>(sort) << ','
                                      << best.KernelRuntime << ',' << best.TotalRuntime << ',' << best.DeviceMemory
                                      << std::endl;
                    } else {
                        std::vector<trinom::real> res(options.N);
                        run_once(options, yield, res, version, block, sort, args.test);
                        if (!args.test) trinom::Arrays::write_array(std::cout, res);
                    }
                }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
