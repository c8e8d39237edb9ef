// Error reporting and small utilities of the plugin framework.
// API of utils/Util.hpp (merijn/Belewitte): reportError/checkError exit
// with status 1, dump_stack_trace prints the call stack and exits with the
// code | 128, safe_dlsym reports a missing symbol, refmap is a map of
// references.
#ifndef BW_FRAMEWORK_UTIL_HPP
#define BW_FRAMEWORK_UTIL_HPP

#include <cstdlib>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <string>

#include <dlfcn.h>

class Algorithm;
typedef void register_algorithm_t(Algorithm &);

[[noreturn]] void out_of_memory();
[[noreturn]] void dump_stack_trace(int exit_code);

inline void printVals() {}

template<typename T, typename... Args>
inline void printVals(const T &v, const Args &...rest)
{
    std::cerr << v;
    printVals(rest...);
}

template<typename... Args>
[[noreturn]] inline void reportError(const Args &...args)
{
    printVals(args...);
    std::cerr << std::endl;
    std::exit(EXIT_FAILURE);
}

template<typename... Args>
inline void checkError(bool cond, const Args &...args)
{
    if (!cond) reportError(args...);
}

template<typename T>
T *safe_dlsym(void *handle, const std::string &name)
{
    void *sym = dlsym(handle, name.c_str());
    if (!sym) {
        const char *why = dlerror();
        reportError("dlsym failed to find '", name, "': ", why ? why : "unknown");
    }
    return reinterpret_cast<T *>(sym);
}

template<typename Key, typename Val>
class refmap : public std::map<Key, std::reference_wrapper<Val>>
{
  public:
    Val &operator[](const Key &key) { return this->at(key).get(); }
};

#endif
