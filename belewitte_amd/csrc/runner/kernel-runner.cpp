// kernel-runner: loads algorithm libraries (lib<algo>kernel.so) and runs
// one implementation on one or more graphs.  Command line, plugin
// discovery, stdin batch mode and output files follow kernel-runner.cpp of
// merijn/Belewitte, with std::filesystem/std::regex in place of boost and
// the HIP backend (register symbol `registerHIP`) in place of CUDA.
//
//   kernel-runner list platforms|devices|algorithms|implementations [-v]
//   kernel-runner query platform|device|algorithm-version
//   kernel-runner -a ALGORITHM -k KERNEL [OPTIONS] <graph file(s)>
//   kernel-runner -S        (jobs "label -a ALGO -k KERNEL [opts] graph" on stdin)
//
// For each graph, <output-dir>/<label>.timings receives the Timer lines and
// <output-dir>/<label>.output the result; stdout gets "<commit>:<label>".

#include <dlfcn.h>
#include <pthread.h>
#include <sched.h>
#include <sys/resource.h>
#include <wordexp.h>

#include <filesystem>
#include <iostream>
#include <locale>
#include <map>
#include <regex>
#include <string>
#include <vector>

#include "Algorithm.hpp"
#include "HIPBackend.hpp"
#include "Options.hpp"
#include "Timer.hpp"
#include "Util.hpp"

namespace fs = std::filesystem;

#define TO_STRING(a) XSTR(a)
#define XSTR(a) #a
#ifndef VERSION
#define VERSION
#endif

namespace {

std::map<std::string, Algorithm> algorithms;
bool debug = false, verbose = false, warnings = false, noOutput = false, printStdOut = false,
     fromStdin = false, useOpenCL = false;
int device = 0;
size_t platform = 0;
std::string outputDir(".");
std::string algorithmName, kernelName;
std::vector<std::string> libPaths;

const char *exeName = "kernel-runner";

Options options('h', "help", std::cout, [](std::ostream &out) {
    out << "Usage:" << std::endl;
    out << exeName << " list platforms [-v | --verbose]" << std::endl;
    out << exeName << " list devices [-p NUM | --platform NUM] [-v | --verbose]" << std::endl;
    out << exeName << " list algorithms [-v | --verbose] [-L PATH | --lib PATH]" << std::endl;
    out << exeName << " list implementations [-a NAME | --algorithm NAME] [-v | --verbose] [-L PATH | --lib PATH]"
        << std::endl;
    out << exeName << " query platform [-p NUM | --platform NUM] [-v | --verbose]" << std::endl;
    out << exeName << " query device [-p NUM | --platform NUM] [-d NUM | --device NUM] [-v | --verbose]"
        << std::endl;
    out << exeName << " query algorithm-version [-a NAME | --algorithm NAME]" << std::endl;
    out << exeName << " -S" << std::endl;
    out << exeName << " -a ALGORITHM -k KERNEL [OPTIONS] <graph file(s)>" << std::endl;
    out << std::endl << "Options:" << std::endl;
});

[[noreturn]] void usage(int code = EXIT_FAILURE)
{
    options.usage(std::cout, "    ");
    std::exit(code);
}

// Directory of this executable, to find the in-tree kernel libraries.
fs::path selfDir()
{
    std::error_code ec;
    auto p = fs::read_symlink("/proc/self/exe", ec);
    return ec ? fs::path(".") : p.parent_path();
}

std::map<std::string, Algorithm> loadAlgorithms(const char *sym, std::vector<std::string> paths)
{
    std::map<std::string, std::string> libs;
    std::map<std::string, Algorithm> result;
    if (fs::is_directory("./.build/kernels/")) paths.insert(paths.begin(), "./.build/kernels/");
    paths.push_back((selfDir() / ".." / "lib" / "kernels").string());

    const std::regex pattern(debug ? "lib(.*)kerneldebug" TO_STRING(VERSION) "\\.so"
                                   : "lib(.*)kernel" TO_STRING(VERSION) "\\.so");
    for (const auto &dir : paths) {
        std::error_code ec;
        if (!fs::is_directory(dir, ec)) continue;
        for (const auto &entry : fs::directory_iterator(dir, ec)) {
            if (!entry.is_regular_file()) continue;
            const std::string file = entry.path().filename().string();
            std::smatch m;
            if (std::regex_match(file, m, pattern)) libs.emplace(m[1].str(), entry.path().string());
        }
    }
    for (const auto &[name, path] : libs) {
        void *h = dlopen(path.c_str(), RTLD_NOW);
        if (!h) {
            if (warnings) std::cerr << "dlopen() failed: " << path << std::endl << dlerror() << std::endl;
            continue;
        }
        auto *reg = reinterpret_cast<register_algorithm_t *>(dlsym(h, sym));
        if (reg) reg(result[name]);
        else if (warnings)
            std::cerr << "dlsym() failed: " << sym << " (" << path << ") " << std::endl << dlerror() << std::endl;
    }
    return result;
}

Algorithm &getAlgorithm(const std::string &name)
{
    auto it = algorithms.find(name);
    if (it != algorithms.end()) return it->second;
    std::string names;
    for (auto &kv : algorithms) names += "    " + kv.first + "\n";
    reportError(name.empty() ? "Algorithm name not specified!" : "No algorithm named \"" + name + "\"!",
                "\n\nSupported algorithms:\n", names);
}

void handle_subcommands(Backend &backend, const std::vector<std::string> &args)
{
    if (args.empty() || (args[0] != "list" && args[0] != "query")) return;
    if (args.size() < 2) usage();
    const std::string &what = args[1];
    if (args[0] == "list" && what == "platforms") {
        backend.listPlatforms(verbose);
        std::exit(static_cast<int>(backend.platformCount()));
    } else if (args[0] == "list" && what == "devices") {
        backend.listDevices(platform, verbose);
        std::exit(backend.deviceCount(platform));
    } else if (args[0] == "list" && what == "algorithms") {
        for (auto &[name, algo] : algorithms) {
            std::cout << name << std::endl;
            if (verbose)
                for (auto &[kname, impl] : algo) {
                    std::cout << "    " << kname << std::endl;
                    impl->help(std::cout, "\t");
                }
        }
        std::exit(EXIT_SUCCESS);
    } else if (args[0] == "list" && what == "implementations") {
        for (auto &[kname, impl] : getAlgorithm(algorithmName)) {
            std::cout << kname << std::endl;
            if (verbose) impl->help(std::cout, "    ");
        }
        std::exit(EXIT_SUCCESS);
    } else if (args[0] == "query" && what == "platform") {
        backend.queryPlatform(platform, verbose);
        std::exit(backend.deviceCount(platform));
    } else if (args[0] == "query" && what == "device") {
        backend.queryDevice(platform, device, verbose);
        std::exit(EXIT_SUCCESS);
    } else if (args[0] == "query" && what == "algorithm-version") {
        std::cout << getAlgorithm(algorithmName).commit() << std::endl;
        std::exit(EXIT_SUCCESS);
    }
    usage();
}

void runJob(const std::string &algo, const std::string &kern, std::vector<std::string> args,
            const std::string &tag = std::string())
{
    auto &algorithm = getAlgorithm(algo);
    algorithm.selectKernel(kern);
    auto graphs = algorithm.setup(args);
    if (!tag.empty() && graphs.size() > 1) reportError("Tagged output with more than 1 graph!");
    for (const auto &graph : graphs) {
        const std::string label = tag.empty() ? fs::path(graph).stem().string() : tag;
        std::string timeFile = (fs::path(outputDir) / (label + ".timings")).string();
        std::string outputFile = (fs::path(outputDir) / (label + ".output")).string();
        if (noOutput) timeFile = outputFile = "/dev/null";
        {
            Epoch epoch(printStdOut ? std::string("/dev/stdout") : timeFile, verbose);
            algorithm(graph, outputFile);
        }
        std::cout << algorithm.commit() << ":" << label << std::endl;
    }
}

void pin_cpu()
{
    cpu_set_t set;
    CPU_ZERO(&set);
    CPU_SET(0, &set);
    if (pthread_setaffinity_np(pthread_self(), sizeof set, &set) != 0 && warnings)
        std::cerr << "pthread_setaffinity_np failed" << std::endl;
}

}  // namespace

int main(int argc, char *const *argv)
{
    options.add('d', "device", "NUM", device, "Device to use.")
        .add('f', "framework", useOpenCL, true, "Use OpenCL (not part of this build).")
        .add('L', "lib", "PATH", libPaths, "\".\"", "Search path for algorithm libraries.")
        .add('o', "output-dir", "DIR", outputDir, "Location to use for writing algorithm output.")
        .add('O', "output", printStdOut, true, "Print timings to stdout, inhibits timings file creation.")
        .add('p', "platform", "NUM", platform, "Platform to use.")
        .add('g', "debug", debug, true, "Use debug kernels.")
        .add('v', "verbose", verbose, true, "Verbose output.")
        .add('W', "warn", warnings, true, "Verbose/debug warnings.")
        .add('S', "stdin", fromStdin, true, "Read work from stdin.")
        .add('q', "quiet", noOutput, true, "Inhibit creation of output and timing files.");

    Options kernelParser(options);
    kernelParser.add('a', "algorithm", "NAME", algorithmName, "Algorithm to use.")
        .add('k', "kernel", "NAME", kernelName, "Which algorithm implementation to use.");

    std::set_new_handler(out_of_memory);
    try {
        std::locale::global(std::locale(""));
    } catch (const std::runtime_error &) {
        std::locale::global(std::locale::classic());
    }
    std::cout.imbue(std::locale());

    struct rlimit limits = {RLIM_INFINITY, RLIM_INFINITY};
    setrlimit(RLIMIT_CORE, &limits);

    auto parsed = options.parseArgsNoUsage(argc, argv);
    pin_cpu();
    if (useOpenCL) reportError("The OpenCL framework is not part of this build; the HIP backend is the default.");
    if (libPaths.empty()) libPaths.push_back(".");
    algorithms = loadAlgorithms("registerHIP", libPaths);

    Backend &backend = HIPBackend::get();
    handle_subcommands(backend, parsed.remainingArgs);

    const bool missingArgs = !fromStdin && parsed.remainingArgs.empty();
    if (parsed.usageRequested || missingArgs) {
        options.usage(std::cout, "    ");
        if (!algorithmName.empty()) getAlgorithm(algorithmName).help(std::cout, "    ");
        std::exit(EXIT_SUCCESS);
    }

    if (!backend.initialised) {
        std::cerr << "Backend not initialised: No devices?" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    backend.setDevice(platform, device);

    if (fromStdin) {
        std::string line;
        while (std::getline(std::cin, line)) {
            wordexp_t words;
            if (wordexp(line.c_str(), &words, WRDE_NOCMD | WRDE_UNDEF))
                reportError("Failed to expand commandline with wordexp(3)!");
            if (words.we_wordc == 0) {
                wordfree(&words);
                continue;
            }
            auto rest = kernelParser.parseArgs(static_cast<int>(words.we_wordc), words.we_wordv);
            const std::string tag = words.we_wordv[0];
            runJob(algorithmName, kernelName, rest, tag);
            kernelParser.reset();
            wordfree(&words);
        }
    } else {
        runJob(algorithmName, kernelName, parsed.remainingArgs);
    }
    return 0;
}
