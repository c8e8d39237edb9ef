// Named set of implementations exported by one kernel library: Algorithm of
// Algorithm.hpp (merijn/Belewitte).  A library's register function runs
// INITIALISE_ALGORITHM(result) and addImplementation for each name.
#ifndef BW_FRAMEWORK_ALGORITHM_HPP
#define BW_FRAMEWORK_ALGORITHM_HPP

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ImplementationBase.hpp"

#ifndef KERNEL_COMMIT
#define KERNEL_COMMIT "unknown"
#endif

#define INITIALISE_ALGORITHM(x) x = Algorithm(KERNEL_COMMIT)

class Algorithm
{
    using ImplMap = std::map<std::string, std::unique_ptr<ImplementationBase>>;

  public:
    typedef ImplMap::iterator iterator;
    typedef ImplMap::const_iterator const_iterator;

    Algorithm();
    explicit Algorithm(const std::string &commit);
    Algorithm(Algorithm &&other);
    Algorithm &operator=(Algorithm &&other);

    void selectKernel(const std::string &kernelName);
    void addImplementation(std::string name, std::unique_ptr<ImplementationBase> &&impl);

    void operator()(const std::string &graphFile, std::ofstream &&output);
    void operator()(const std::string &graphFile, const std::string &output);
    void help(std::ostream &out, std::string prefix);
    std::vector<std::string> setup(std::vector<std::string> args);
    std::string commit();

    iterator begin() { return implementations.begin(); }
    iterator end() { return implementations.end(); }
    const_iterator begin() const { return implementations.begin(); }
    const_iterator end() const { return implementations.end(); }

  private:
    std::string kernelNames();

    std::string commitHash;
    ImplementationBase *selectedKernel;
    ImplMap implementations;
};

#endif
