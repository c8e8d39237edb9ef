#include "Timer.hpp"

#include <iomanip>
#include <iostream>
#include <locale>

namespace {

struct TimerSlot {
    std::string name;
    std::shared_ptr<std::vector<double>> timings;
};

struct EpochState {
    std::string name;
    std::vector<TimerSlot> timers;
};

std::vector<EpochState> &epochs()
{
    static std::vector<EpochState> e;
    return e;
}

size_t register_epoch(const std::string &name)
{
    epochs().push_back({name, {}});
    return epochs().size() - 1;
}

std::string human(double ns)
{
    std::ostringstream s;
    long long t = static_cast<long long>(ns);
    const long long h = t / 3600000000000LL;
    t -= h * 3600000000000LL;
    const long long m = t / 60000000000LL;
    t -= m * 60000000000LL;
    const long long sec = t / 1000000000LL;
    t -= sec * 1000000000LL;
    const long long ms = t / 1000000LL;
    t -= ms * 1000000LL;
    const long long us = t / 1000LL;
    t -= us * 1000LL;
    if (h) s << std::setw(3) << h << " h   ";
    if (m) s << std::setw(3) << m << " m   ";
    if (sec) s << std::setw(3) << sec << " s   ";
    if (ms) s << std::setw(3) << ms << " ms   ";
    if (us) s << std::setw(3) << us << " us   ";
    s << std::setw(3) << t << " ns";
    return s.str();
}

}  // namespace

Epoch::Epoch(bool humanReadable, const std::string &name)
  : direct_printed(false), human_readable(humanReadable), output("/dev/null"), index(register_epoch(name))
{}

Epoch::Epoch(std::ofstream &&out, bool humanReadable, const std::string &name)
  : direct_printed(true), human_readable(humanReadable), output(std::move(out)), index(register_epoch(name))
{}

Epoch::Epoch(std::ofstream &&out, const std::string &name, bool humanReadable)
  : Epoch(std::move(out), humanReadable, name)
{}

Epoch::Epoch(const std::string &file, const std::string &name, bool humanReadable)
  : Epoch(std::ofstream(file), humanReadable, name)
{}

Epoch::Epoch(const std::string &file, bool humanReadable, const std::string &name)
  : Epoch(std::ofstream(file), humanReadable, name)
{}

Epoch::~Epoch()
{
    if (direct_printed) print_results(output, human_readable);
}

void Epoch::set_output(const std::string &file) { output = std::ofstream(file); }

void Epoch::set_output(std::ofstream &&out) { output = std::move(out); }

void Epoch::print_results(std::ostream &out, bool humanReadable)
{
    auto &epoch = epochs()[index];
    if (humanReadable) out << "Timer results with precision: 1000000000" << std::endl << std::endl;
    for (auto &t : epoch.timers) {
        if (t.timings->empty()) continue;
        StatisticalSummary<double> r(*t.timings);
        if (!epoch.name.empty()) out << epoch.name << ":";
        if (humanReadable) {
            out << t.name << " (" << t.timings->size() << "): " << std::endl;
            out << "Min: " << human(r.min) << std::endl << "Avg: " << human(r.mean) << std::endl
                << "Max: " << human(r.max) << std::endl << "Std: " << human(r.stdDev) << std::endl
                << std::endl;
        } else {
            // Doubles in the stream's default notation (6 significant
            // digits), as the reference prints Timing::total_time.
            out.imbue(std::locale::classic());
            out << t.name << ":" << r.min << " " << r.mean << " " << r.max << " " << r.stdDev << std::endl;
        }
    }
}

Timer::Timer(const std::string &name, size_t reserve)
  : timings(std::make_shared<std::vector<double>>())
{
    timings->reserve(reserve);
    if (epochs().empty()) register_epoch("");
    epochs().back().timers.push_back({name, timings});
}

void Timer::start() { begin = std::chrono::steady_clock::now(); }

void Timer::stop()
{
    const auto d = std::chrono::steady_clock::now() - begin;
    timings->push_back(std::chrono::duration<double, std::nano>(d).count());
}

void Timer::reserve(size_t count)
{
    if (timings->capacity() - timings->size() < count)
        timings->reserve(timings->size() + std::max<size_t>(count, 100));
}

void Timer::add(double ns) { timings->push_back(ns); }
