// hwtree/trinom.hpp -- the reference's option and yield containers and their
// text format, for C++ callers of the MI355X pricer.
//
// Same public names, field order and element types as the reference's
// trinom::Options (common/Options.hpp:48-125), trinom::Yield
// (common/Yield.hpp:14-40), OptionType/SortType (common/Options.hpp:15-46)
// and Arrays (common/Arrays.hpp:12-47), so existing callers compile unchanged.
// `real` is float unless USE_DOUBLE is defined, as in the reference
// (common/Real.hpp:13-21); the pricer computes in FP64 (or FP32 tree states
// with FP64 sums) whichever width the caller holds its data in.
#pragma once

#include <cmath>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace trinom {

#ifdef USE_DOUBLE
typedef double real;
#else
typedef float real;
#endif

enum class SortType : char { WIDTH_DESC = 'W', WIDTH_ASC = 'w', HEIGHT_DESC = 'H', HEIGHT_ASC = 'h', NONE = '-' };

enum class OptionType : int8_t { CALL = 0, PUT = 1 };

inline std::ostream &operator<<(std::ostream &os, OptionType t) { return os << static_cast<int>(t); }

inline std::istream &operator>>(std::istream &is, OptionType &t)
{
    int v;
    is >> v;
    if (v != 0 && v != 1) throw std::out_of_range("Invalid OptionType read from stream.");
    t = static_cast<OptionType>(v);
    return is;
}

// Arrays of the form "[v0, v1, ...]", one per line.
struct Arrays {
    template <class T>
    static void read_array(std::istream &in, std::vector<T> &out)
    {
        char c = 0;
        if (!(in >> c) || c != '[') return;
        if (in >> std::ws && in.peek() == ']') {
            in.get();
            return;
        }
        while (in) {
            T v;
            if (!(in >> v)) break;
            out.push_back(v);
            if (!(in >> c) || c == ']') break;
        }
    }

    // std::fixed with max_digits10 decimals (17 for double, 0 for integers).
    template <class T>
    static void write_array(std::ostream &out, const std::vector<T> &a)
    {
        const auto old_prec = out.precision(std::numeric_limits<T>::max_digits10);
        const auto old_flags = out.flags();
        out << '[' << std::fixed;
        for (size_t i = 0; i < a.size(); ++i) {
            if (i) out << ", ";
            write_one(out, a[i]);
        }
        out << ']' << std::endl;
        out.precision(old_prec);
        out.flags(old_flags);
    }

private:
    template <class T>
    static void write_one(std::ostream &out, const T &v) { out << v; }
    static void write_one(std::ostream &out, const uint16_t &v) { out << static_cast<unsigned>(v); }
};

// uint16 fields are read as integers, not characters.
template <>
inline void Arrays::read_array<uint16_t>(std::istream &in, std::vector<uint16_t> &out)
{
    std::vector<unsigned> tmp;
    read_array(in, tmp);
    out.insert(out.end(), tmp.begin(), tmp.end());
}

struct Options {
    int N = 0;
    std::vector<real> StrikePrices;
    std::vector<real> Maturities;      // bond maturity in years: sets the tree height
    std::vector<real> Lengths;         // option expiry in years: sets the exercise step
    std::vector<uint16_t> TermUnits;
    std::vector<uint16_t> TermStepCounts;
    std::vector<real> ReversionRates;
    std::vector<real> Volatilities;
    std::vector<OptionType> Types;

    explicit Options(int count) : N(count)
    {
        StrikePrices.reserve(count);
        Maturities.reserve(count);
        Lengths.reserve(count);
        TermUnits.reserve(count);
        TermStepCounts.reserve(count);
        ReversionRates.reserve(count);
        Volatilities.reserve(count);
        Types.reserve(count);
    }

    explicit Options(const std::string &filename)
    {
        if (filename.empty()) throw std::invalid_argument("File not specified.");
        std::ifstream in(filename);
        if (!in) throw std::invalid_argument("File '" + filename + "' does not exist.");
        Arrays::read_array(in, StrikePrices);
        Arrays::read_array(in, Maturities);
        Arrays::read_array(in, Lengths);
        Arrays::read_array(in, TermUnits);
        Arrays::read_array(in, TermStepCounts);
        Arrays::read_array(in, ReversionRates);
        Arrays::read_array(in, Volatilities);
        Arrays::read_array(in, Types);
        N = static_cast<int>(StrikePrices.size());
    }

    void writeToFile(const std::string &filename) const
    {
        if (filename.empty()) throw std::invalid_argument("File not specified.");
        std::ofstream out(filename);
        if (!out) throw std::invalid_argument("File does not exist.");
        Arrays::write_array(out, StrikePrices);
        Arrays::write_array(out, Maturities);
        Arrays::write_array(out, Lengths);
        Arrays::write_array(out, TermUnits);
        Arrays::write_array(out, TermStepCounts);
        Arrays::write_array(out, ReversionRates);
        Arrays::write_array(out, Volatilities);
        Arrays::write_array(out, Types);
    }
};

struct Yield {
    int N = 0;
    std::vector<real> Prices;      // continuously compounded zero rates
    std::vector<int32_t> TimeSteps;  // day offsets

    Yield() = default;

    explicit Yield(const std::string &filename)
    {
        if (filename.empty()) throw std::invalid_argument("File not specified.");
        std::ifstream in(filename);
        if (!in) throw std::invalid_argument("File does not exist.");
        Arrays::read_array(in, Prices);
        Arrays::read_array(in, TimeSteps);
        N = static_cast<int>(Prices.size());
    }
};

// An option's tree constants (common/OptionConstants.hpp:14-53), derived as
// the reference derives them in either width of `real`, so a caller sees the
// same height, jmax and width: `real` arithmetic, with the C library's double
// exp, sqrt and ceil (what the reference's unqualified calls resolve to)
// rounded to `real` where the fields store them.
struct OptionConstants {
    real t;
    real dt;  // [years]
    real dr;
    real X;
    real M;
    int32_t jmax;
    int32_t n;
    int32_t width;
    uint16_t termUnit;
    OptionType type;

    OptionConstants() {}

    OptionConstants(const Options &options, const int ind)
    {
        termUnit = options.TermUnits.at(ind);
        const real T = options.Maturities.at(ind);
        const int termUnitsInYearCount =
            static_cast<int>(std::ceil(static_cast<double>(static_cast<real>(365.0) / termUnit)));
        const uint16_t termStepCount = options.TermStepCounts.at(ind);
        n = static_cast<int32_t>(termStepCount * termUnitsInYearCount * T);
        t = options.Lengths.at(ind);
        dt = termUnitsInYearCount / static_cast<real>(termStepCount);
        type = options.Types.at(ind);
        const real a = options.ReversionRates.at(ind);
        X = options.StrikePrices.at(ind);
        const real sigma = options.Volatilities.at(ind);
        const double V = sigma * sigma *
                         (static_cast<real>(1) - std::exp(static_cast<double>(static_cast<real>(-2) * a * dt))) /
                         (static_cast<real>(2) * a);
        dr = static_cast<real>(std::sqrt(static_cast<real>(3) * V));
        M = static_cast<real>(std::exp(static_cast<double>(-a * dt)) - static_cast<real>(1));
        jmax = static_cast<int32_t>(static_cast<real>(-0.184) / M) + 1;
        width = 2 * jmax + 1;
    }
};

} // namespace trinom
