#include "Util.hpp"

#include <cxxabi.h>
#include <execinfo.h>

#include <cstdio>

void out_of_memory() { dump_stack_trace(EXIT_FAILURE); }

void dump_stack_trace(int exit_code)
{
    void *frames[128];
    const int n = backtrace(frames, 128);
    char **names = backtrace_symbols(frames, n);
    for (int i = 1; i < n; i++) {
        Dl_info info;
        const char *shown = names ? names[i] : "?";
        char *demangled = nullptr;
        if (dladdr(frames[i], &info) && info.dli_sname) {
            int status = -1;
            if (info.dli_sname[0] == '_')
                demangled = abi::__cxa_demangle(info.dli_sname, nullptr, nullptr, &status);
            shown = (demangled && status == 0) ? demangled : info.dli_sname;
        }
        std::fprintf(stderr, "%-3d %s\n", i, shown);
        std::free(demangled);
    }
    std::free(names);
    if (n == 128) std::fprintf(stderr, "[truncated]\n");
    std::exit(exit_code | 128);
}
