// One runnable implementation of an algorithm: ImplementationBase of
// ImplementationBase.hpp (merijn/Belewitte).  Lifecycle per graph:
// prepareRun -> loadGraph -> runImplementation -> freeGraph -> cleanupRun
// -> options.reset().  Flags -n/--count and --validate.
#ifndef BW_FRAMEWORK_IMPLEMENTATIONBASE_HPP
#define BW_FRAMEWORK_IMPLEMENTATIONBASE_HPP

#include <fstream>
#include <string>
#include <vector>

#include "Options.hpp"

struct ImplementationBase {
  private:
    bool validate_;

  protected:
    ImplementationBase() : validate_(false), run_count(1), validate(validate_)
    {
        options.add('n', "count", "NUM", run_count, "Number of times to run algorithm.");
        options.add("validate", validate_, true, "Output validation results.");
    }

    ImplementationBase(const ImplementationBase &) = delete;
    void operator=(const ImplementationBase &) = delete;

    Options options;
    size_t run_count;

    virtual void loadGraph(const std::string) = 0;
    virtual void prepareRun() {}
    virtual void runImplementation(std::ofstream &) = 0;
    virtual void cleanupRun() {}
    virtual void freeGraph() = 0;

  public:
    const bool &validate;
    void operator()(const std::string &graphFile, std::ofstream &&output);
    void operator()(const std::string &graphFile, const std::string &output);
    void help(std::ostream &out, std::string prefix);
    std::vector<std::string> setup(std::vector<std::string> args);

    virtual ~ImplementationBase();
};

#endif
