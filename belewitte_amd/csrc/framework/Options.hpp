// Command-line flags with per-run reset: the Options class of
// options/Options.hpp (merijn/Belewitte), same method names and semantics
// (short/long flags, global versus nested reservation, parseArgs /
// parseArgsNoUsage / parseArgsFinal, usage, reset), without boost.
#ifndef BW_FRAMEWORK_OPTIONS_HPP
#define BW_FRAMEWORK_OPTIONS_HPP

#include <functional>
#include <iostream>
#include <map>
#include <set>
#include <string>
#include <type_traits>
#include <vector>

class Options
{
    struct OptionResult {
        std::vector<std::string> remainingArgs;
        bool usageRequested;
    };

    struct Option {
        char shortOption = '\0';
        std::string longOption;
        std::function<void(const std::string &)> action;
        std::function<void()> resetAction;
        std::string argName;
        std::string helpString;
        std::string defaultVal;
        bool hasArg = false;
        bool multiFlag = false;

        void reset() const
        { if (resetAction) resetAction(); }
    };

  public:
    Options() : parent(nullptr), usageOutput(std::cerr), hasUsage(false), isGlobal(false) {}

    Options(char c, const char *l, std::ostream &out, std::function<void(std::ostream &)> preamble)
      : parent(nullptr), usageOutput(out), usagePreamble(std::move(preamble)), hasUsage(true), isGlobal(true)
    {
        usageFlag.shortOption = c;
        usageFlag.longOption = l;
        usageFlag.helpString = "This help.";
    }

    explicit Options(Options &o) : parent(&o), usageOutput(std::cerr), hasUsage(false), isGlobal(false)
    { parent->children.insert(this); }

    Options(const Options &) = delete;
    Options &operator=(const Options &) = delete;

    ~Options()
    { if (parent) parent->children.erase(this); }

    Options &add(const Option &o);

    // String-valued flag.
    Options &add(char so, const char *lo, std::string arg, std::string &var, std::string help);
    Options &add(const char *lo, std::string arg, std::string &var, std::string help)
    { return add('\0', lo, std::move(arg), var, std::move(help)); }

    // Double-valued flag.
    Options &add(char so, const char *lo, std::string arg, double &var, std::string help);
    Options &add(const char *lo, std::string arg, double &var, std::string help)
    { return add('\0', lo, std::move(arg), var, std::move(help)); }

    // Switch: sets var = val when present.
    template<typename T>
    Options &add(char so, const char *lo, T &var, T val, std::string help)
    {
        Option o;
        o.shortOption = so;
        o.longOption = lo;
        o.action = [&var, val](const std::string &) { var = val; };
        o.resetAction = [&var, initial = T(var)]() { var = initial; };
        o.helpString = std::move(help);
        return add(o);
    }

    template<typename T>
    Options &add(const char *lo, T &var, T val, std::string help)
    { return add('\0', lo, var, val, std::move(help)); }

    // Integral flag with an argument.
    template<typename T, typename = std::enable_if_t<std::is_integral<T>::value>>
    Options &add(char so, const char *lo, std::string arg, T &var, std::string help)
    {
        Option o;
        o.shortOption = so;
        o.longOption = lo;
        o.action = [&var](const std::string &s) { var = static_cast<T>(std::stoll(s)); };
        o.resetAction = [&var, initial = T(var)]() { var = initial; };
        o.argName = std::move(arg);
        o.helpString = std::move(help);
        o.defaultVal = std::to_string(var);
        o.hasArg = true;
        return add(o);
    }

    template<typename T, typename = std::enable_if_t<std::is_integral<T>::value>>
    Options &add(const char *lo, std::string arg, T &var, std::string help)
    { return add('\0', lo, std::move(arg), var, std::move(help)); }

    // Repeatable flag collecting into a vector.
    template<typename T>
    Options &add(char so, const char *lo, std::string arg, std::vector<T> &var, std::string def,
                 std::string help, std::function<T(std::string)> fun = [](std::string s) { return s; })
    {
        Option o;
        o.shortOption = so;
        o.longOption = lo;
        o.action = [&var, fun](const std::string &s) { var.push_back(fun(s)); };
        o.resetAction = [&var]() { var.clear(); };
        o.argName = std::move(arg);
        o.helpString = std::move(help);
        o.defaultVal = std::move(def);
        o.hasArg = true;
        o.multiFlag = true;
        return add(o);
    }

    std::vector<std::string> parseArgs(const std::vector<std::string> &args);
    std::vector<std::string> parseArgs(int argc, char *const *argv);
    OptionResult parseArgsNoUsage(const std::vector<std::string> &args);
    OptionResult parseArgsNoUsage(int argc, char *const *argv);
    std::vector<std::string> parseArgsFinal(const std::vector<std::string> &args);
    std::vector<std::string> parseArgsFinal(int argc, char *const *argv);

    void usage(std::ostream &out, std::string prefix = "");
    void reset();

  private:
    OptionResult parse(const std::vector<std::string> &args, bool exitUnknown, bool exitUsage);
    static std::vector<std::string> toVector(int argc, char *const *argv);

    static std::set<char> &globalShort();
    static std::set<std::string> &globalLong();

    Options *parent;
    std::set<Options *> children;
    std::set<char> reservedShort;
    std::set<std::string> reservedLong;
    std::ostream &usageOutput;
    std::function<void(std::ostream &)> usagePreamble;
    Option usageFlag;
    bool hasUsage, isGlobal;
    std::map<std::string, Option> options;  // keyed by long name, render order
};

#endif
