#include "ImplementationBase.hpp"

ImplementationBase::~ImplementationBase() {}

void ImplementationBase::operator()(const std::string &graphFile, std::ofstream &&output)
{
    prepareRun();
    loadGraph(graphFile);
    runImplementation(output);
    freeGraph();
    cleanupRun();
    options.reset();
}

void ImplementationBase::operator()(const std::string &graphFile, const std::string &output)
{ (*this)(graphFile, std::ofstream(output)); }

void ImplementationBase::help(std::ostream &out, std::string prefix)
{ options.usage(out, prefix); }

std::vector<std::string> ImplementationBase::setup(std::vector<std::string> args)
{ return options.parseArgsFinal(args); }
