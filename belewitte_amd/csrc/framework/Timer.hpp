// Wall-clock timers grouped in epochs: the Timer / Epoch API of Timer.hpp
// (merijn/Belewitte).  An Epoch collects every Timer created while it is the
// newest epoch and, when destroyed, writes one line per timer in the format
// the experiment tooling parses (Timer.cpp:287-291 of the reference):
//     name:min avg max stddev        (nanoseconds, C locale)
#ifndef BW_FRAMEWORK_TIMER_HPP
#define BW_FRAMEWORK_TIMER_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

// Summary statistics with the index conventions of
// utils/StatisticalSummary.hpp: median and quartiles average the two middle
// elements of even-sized halves, stdDev is the sample standard deviation.
template<typename T>
struct StatisticalSummary
{
    template<typename V>
    explicit StatisticalSummary(std::vector<V> values)
    {
        if (values.empty()) throw std::domain_error("Empty vector has no summary!");
        auto mid = [](size_t n) -> std::pair<size_t, size_t> {
            return n % 2 == 0 ? std::make_pair(n / 2 - 1, n / 2) : std::make_pair(n / 2, n / 2);
        };
        const size_t n = values.size(), half = n / 2;
        std::pair<size_t, size_t> med{0, 0}, lo{0, 0}, hi{0, 0};
        if (n > 1) {
            med = mid(n);
            lo = mid(half);
            hi = {lo.first + half + n % 2, lo.second + half + n % 2};
        }
        std::sort(values.begin(), values.end());
        double total = 0, M = 0, S = 0;
        long k = 1;
        for (const auto &x : values) {
            const double v = static_cast<double>(x);
            total += v;
            const double prevM = M;
            M += (v - prevM) / static_cast<double>(k);
            S += (v - prevM) * (v - M);
            k++;
        }
        auto at = [&](size_t i) { return static_cast<double>(values[i]); };
        min = T(at(0));
        lowerQuantile = T((at(lo.first) + at(lo.second)) / 2.0);
        median = T((at(med.first) + at(med.second)) / 2.0);
        mean = T(total / static_cast<double>(n));
        upperQuantile = T((at(hi.first) + at(hi.second)) / 2.0);
        max = T(at(n - 1));
        stdDev = T(k - 2 == 0 ? 0.0 : std::sqrt(S / static_cast<double>(k - 2)));
    }

    T min, lowerQuantile, median, mean, upperQuantile, max, stdDev;
};

class Epoch
{
  public:
    Epoch() = delete;
    Epoch(const Epoch &) = delete;
    Epoch(Epoch &&) = delete;

    explicit Epoch(bool humanReadable = false, const std::string &name = std::string());
    Epoch(std::ofstream &&out, const std::string &name = std::string(), bool humanReadable = false);
    Epoch(std::ofstream &&out, bool humanReadable, const std::string &name = std::string());
    Epoch(const std::string &file, bool humanReadable, const std::string &name = std::string());
    Epoch(const std::string &file, const std::string &name = std::string(), bool humanReadable = false);
    ~Epoch();

    void set_output(const std::string &file);
    void set_output(std::ofstream &&out);
    void print_results(std::ostream &out, bool humanReadable = false);

  private:
    bool direct_printed;
    bool human_readable;
    std::ofstream output;
    size_t index;
};

class Timer
{
  public:
    explicit Timer(const std::string &name, size_t reserve = 10);
    Timer(const Timer &) = delete;
    Timer(Timer &&) = default;

    void start();
    void stop();
    void reserve(size_t count);
    // Records an externally measured duration (device events).
    void add(double nanoseconds);

  private:
    std::chrono::steady_clock::time_point begin;
    std::shared_ptr<std::vector<double>> timings;
};

#endif
