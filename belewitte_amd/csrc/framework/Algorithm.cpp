#include "Algorithm.hpp"

#include <sstream>

#include "Util.hpp"

Algorithm::Algorithm() : selectedKernel(nullptr) {}

Algorithm::Algorithm(const std::string &commit) : commitHash(commit), selectedKernel(nullptr) {}

Algorithm::Algorithm(Algorithm &&o)
  : commitHash(std::move(o.commitHash)), selectedKernel(o.selectedKernel),
    implementations(std::move(o.implementations))
{}

Algorithm &Algorithm::operator=(Algorithm &&o)
{
    commitHash = std::move(o.commitHash);
    selectedKernel = o.selectedKernel;
    implementations = std::move(o.implementations);
    return *this;
}

void Algorithm::selectKernel(const std::string &kernelName)
{
    auto it = implementations.find(kernelName);
    selectedKernel = it == implementations.end() ? nullptr : it->second.get();
    if (commitHash.empty()) reportError("Algorithm version not specified!");
    if (!selectedKernel) {
        const std::string msg =
            kernelName.empty() ? "Kernel name not specified!" : "No kernel named \"" + kernelName + "\" exists!";
        reportError(msg, "\n\nSupported kernels:\n", kernelNames());
    }
}

void Algorithm::addImplementation(std::string name, std::unique_ptr<ImplementationBase> &&impl)
{
    if (!implementations.emplace(name, std::move(impl)).second)
        reportError("Implementation with name \"" + name + "\" already exists!");
}

void Algorithm::operator()(const std::string &graphFile, std::ofstream &&output)
{
    if (!selectedKernel) reportError("No kernel selected to run!");
    (*selectedKernel)(graphFile, std::move(output));
}

void Algorithm::operator()(const std::string &graphFile, const std::string &output)
{
    if (!selectedKernel) reportError("No kernel selected to run!");
    (*selectedKernel)(graphFile, output);
}

void Algorithm::help(std::ostream &out, std::string prefix)
{
    if (!selectedKernel) reportError("No kernel selected!");
    selectedKernel->help(out, prefix);
}

std::vector<std::string> Algorithm::setup(std::vector<std::string> args)
{
    if (!selectedKernel) reportError("No kernel selected!");
    return selectedKernel->setup(args);
}

std::string Algorithm::commit() { return commitHash; }

std::string Algorithm::kernelNames()
{
    std::ostringstream names;
    for (auto &kv : implementations) names << "    " << kv.first << std::endl;
    return names.str();
}
