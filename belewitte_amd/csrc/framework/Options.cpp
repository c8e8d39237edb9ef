#include "Options.hpp"

#include <cassert>
#include <cstdlib>

#include "Util.hpp"

std::set<char> &Options::globalShort()
{
    static std::set<char> s;
    return s;
}

std::set<std::string> &Options::globalLong()
{
    static std::set<std::string> s;
    return s;
}

Options &Options::add(const Option &o)
{
    const bool shortTaken = o.shortOption != '\0' &&
                            (globalShort().count(o.shortOption) || reservedShort.count(o.shortOption));
    if (shortTaken) reportError("Flag '-", o.shortOption, "' is already reserved.");
    if (globalLong().count(o.longOption) || reservedLong.count(o.longOption) || options.count(o.longOption))
        reportError("Flag '--", o.longOption, "' is already reserved.");

    options.emplace(o.longOption, o);
    auto &shorts = isGlobal ? globalShort() : reservedShort;
    auto &longs = isGlobal ? globalLong() : reservedLong;
    if (o.shortOption != '\0') shorts.insert(o.shortOption);
    longs.insert(o.longOption);
    if (parent) parent->add(o);
    return *this;
}

Options &Options::add(char so, const char *lo, std::string arg, std::string &var, std::string help)
{
    Option o;
    o.shortOption = so;
    o.longOption = lo;
    o.action = [&var](const std::string &s) { var = s; };
    o.resetAction = [&var, initial = var]() { var = initial; };
    o.argName = std::move(arg);
    o.helpString = std::move(help);
    o.defaultVal = var;
    o.hasArg = true;
    return add(o);
}

Options &Options::add(char so, const char *lo, std::string arg, double &var, std::string help)
{
    Option o;
    o.shortOption = so;
    o.longOption = lo;
    o.action = [&var](const std::string &s) { var = std::stod(s); };
    o.resetAction = [&var, initial = var]() { var = initial; };
    o.argName = std::move(arg);
    o.helpString = std::move(help);
    o.defaultVal = std::to_string(var);
    o.hasArg = true;
    return add(o);
}

std::vector<std::string> Options::toVector(int argc, char *const *argv)
{
    std::vector<std::string> v;
    for (int i = 1; i < argc; i++) v.emplace_back(argv[i]);
    return v;
}

std::vector<std::string> Options::parseArgs(const std::vector<std::string> &args)
{ return parse(args, false, true).remainingArgs; }

std::vector<std::string> Options::parseArgs(int argc, char *const *argv)
{ return parseArgs(toVector(argc, argv)); }

Options::OptionResult Options::parseArgsNoUsage(const std::vector<std::string> &args)
{ return parse(args, false, false); }

Options::OptionResult Options::parseArgsNoUsage(int argc, char *const *argv)
{ return parseArgsNoUsage(toVector(argc, argv)); }

std::vector<std::string> Options::parseArgsFinal(const std::vector<std::string> &args)
{ return parse(args, true, true).remainingArgs; }

std::vector<std::string> Options::parseArgsFinal(int argc, char *const *argv)
{ return parseArgsFinal(toVector(argc, argv)); }

Options::OptionResult Options::parse(const std::vector<std::string> &args, bool exitUnknown, bool exitUsage)
{
    bool usageRequested = false;
    std::map<std::string, const Option *> table;
    Option usageOpt = usageFlag;
    auto enter = [&table](const std::string &flag, const Option *o) {
        if (!table.emplace(flag, o).second) reportError("Option ", flag, " is already in use!");
    };
    if (hasUsage) {
        if (exitUsage)
            usageOpt.action = [this](const std::string &) {
                usage(usageOutput);
                std::exit(EXIT_SUCCESS);
            };
        else
            usageOpt.action = [&usageRequested](const std::string &) { usageRequested = true; };
        if (usageOpt.shortOption != '\0') enter(std::string("-") + usageOpt.shortOption, &usageOpt);
        enter("--" + usageOpt.longOption, &usageOpt);
    }
    for (auto &kv : options) {
        const Option &o = kv.second;
        if (o.shortOption != '\0') enter(std::string("-") + o.shortOption, &o);
        enter("--" + o.longOption, &o);
    }

    std::vector<std::string> rest;
    for (size_t i = 0; i < args.size(); i++) {
        auto it = table.find(args[i]);
        if (it == table.end()) {
            if (args[i] == "--") {
                rest.insert(rest.end(), args.begin() + static_cast<long>(i) + 1, args.end());
                break;
            }
            if (exitUnknown && !args[i].empty() && args[i][0] == '-') {
                std::cerr << "Unknown option '" << args[i] << "'!" << std::endl << std::endl;
                std::cerr << "Possible options:" << std::endl;
                usage(usageOutput);
                std::exit(EXIT_FAILURE);
            }
            rest.push_back(args[i]);
            continue;
        }
        const Option &o = *it->second;
        if (o.hasArg) {
            if (i + 1 >= args.size()) reportError("Option ", args[i], " doesn't have an argument!");
            o.action(args[++i]);
        } else {
            o.action("");
        }
    }
    return {rest, usageRequested};
}

void Options::usage(std::ostream &out, std::string prefix)
{
    auto render = [&](const Option &o) {
        out << prefix;
        if (o.shortOption != '\0') {
            out << '-' << o.shortOption;
            if (o.hasArg) out << ' ' << o.argName;
            out << " | ";
        }
        out << "--" << o.longOption;
        if (o.hasArg) out << ' ' << o.argName;
        out << std::endl << prefix << '\t' << o.helpString << std::endl;
        if (!o.defaultVal.empty()) out << prefix << "\t(default: " << o.defaultVal << ')' << std::endl;
        out << std::endl;
    };
    if (hasUsage) {
        if (usagePreamble) usagePreamble(out);
        render(usageFlag);
    }
    for (auto &kv : options) render(kv.second);
}

void Options::reset()
{
    for (Options *c : children) c->reset();
    for (auto &kv : options) kv.second.reset();
}
