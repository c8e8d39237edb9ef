// Per-level implementation switch (SwitchImplementation of
// ImplementationTemplate.hpp:545-908, merijn/Belewitte).
//
// Every registered implementation shares the algorithm's result buffers,
// so the active kernel tuple can be replaced between levels.  A predictor
// library (-m FILE, dlopen'ed) exports the reference's ABI:
//     int32_t lookup();                         index into implNames or -1
//     const std::vector<std::tuple<std::string,int64_t,size_t,size_t,size_t>>
//         implNames;                            (name, id, index, warp, chunk)
//     const std::map<std::string,std::reference_wrapper<double>> propNames;
//     int32_t default_impl;
// Without a model the switch runs "edge-list".  -l FILE logs the property
// values and predictions in the grammar the experiment tooling parses:
//     graph:<name>:<value>   step:<n>:<name>:<value>   prediction:<n>:<id>
#ifndef BW_FRAMEWORK_SWITCHIMPLEMENTATION_HPP
#define BW_FRAMEWORK_SWITCHIMPLEMENTATION_HPP

template<typename AlgorithmBase, typename... Kernels>
struct SwitchImplementation : public AlgorithmBase {
    using typename AlgorithmBase::Vertex;
    using typename AlgorithmBase::Edge;
    using AlgorithmBase::algorithmProperties;
    using AlgorithmBase::graphProperties;
    using AlgorithmBase::loader;
    using AlgorithmBase::options;

    using prop_set = std::set<std::string>;
    using Map = KernelMap<std::string, std::tuple<Kernels...>>;

    std::tuple<Kernels...> kernels;

    class graph_prop
    {
        prop_ref absProp, inProp, outProp;

      public:
        graph_prop(const std::string &prefix, const std::string &suffix, SwitchImplementation &cfg)
          : absProp(prefix + "abs" + suffix, cfg, true), inProp(prefix + "in" + suffix, cfg, true),
            outProp(prefix + "out" + suffix, cfg, true)
        {}
        graph_prop(const graph_prop &) = delete;

        prop_ref &operator[](Degrees d)
        {
            if (d == Degrees::abs) return absProp;
            return d == Degrees::in ? inProp : outProp;
        }
    };

    struct impl_id {
        int64_t id = -1;
        std::string name;
        impl_id() = default;
        impl_id(int64_t i, std::string s) : id(i), name(std::move(s)) {}
    };

    explicit SwitchImplementation(Map ks) : SwitchImplementation(ks, std::index_sequence_for<Kernels...>()) {}

    template<size_t... I>
    SwitchImplementation(Map ks, std::index_sequence<I...>)
      : AlgorithmBase(std::get<I>(kernels)...), kernels((static_cast<void>(I), nullptr)...), modelHandle(nullptr),
        kernelMap(ks), vertices("vertex count", *this, true), edges("edge count", *this, true),
        min("min ", " degree", *this), lowerQuantile("lower quantile ", " degree", *this),
        mean("mean ", " degree", *this), median("median ", " degree", *this),
        upperQuantile("upper quantile ", " degree", *this), max("max ", " degree", *this),
        stdDev("stddev ", " degree", *this)
    {
        options.add('m', "model", "FILE", model, "Prediction model to use.");
        options.add('l', "log", "FILE", logFile, "Where to log properties.");
    }

    void loadGraph(const Graph<Vertex, Edge> &graph) override final
    {
        vertices = static_cast<double>(graph.vertex_count);
        edges = static_cast<double>(graph.edge_count);
        for (Degrees d : {Degrees::abs, Degrees::in, Degrees::out}) {
            auto s = graph.degreeStatistics(d);
            min[d] = s.min;
            lowerQuantile[d] = s.lowerQuantile;
            mean[d] = s.mean;
            median[d] = s.median;
            upperQuantile[d] = s.upperQuantile;
            max[d] = s.max;
            stdDev[d] = s.stdDev;
        }
        std::vector<GraphRep> reps;
        for (auto &impl : implementations)
            mapKernels([&reps](auto &&k) { reps.push_back(k->representation); }, impl);
        loader.loadGraph(graph, reps);
    }

    void transferGraph() override final
    {
        for (auto &impl : implementations)
            mapKernels([this](auto &&k) { loader.transferGraph(k->representation); }, impl);
    }

  protected:
    void predictInitial() override final
    {
        stepNum = 0;
        logGraphProps();
        logAlgorithmProps();
        lastKernel = lookup();
        if (lastKernel == -1) lastKernel = defaultKernel;
        kernels = implementations[static_cast<size_t>(lastKernel)];
    }

    bool predict() override final
    {
        ++stepNum;
        logAlgorithmProps();
        const int32_t result = lookup();
        if (result != -1 && result != lastKernel) {
            kernels = implementations[static_cast<size_t>(result)];
            lastKernel = result;
            return true;
        }
        return false;
    }

    void prepareRun() override final
    {
        prop_set missingGraphProps, missingAlgoProps;
        for (auto &kv : graphProperties) missingGraphProps.insert(kv.first);
        for (auto &kv : algorithmProperties) missingAlgoProps.insert(kv.first);

        lastKernel = -1;
        if (!model.empty()) {
            setupPredictor(model.c_str(), missingGraphProps, missingAlgoProps);
        } else {
            lookup = []() { return -1; };
            auto it = kernelMap.find("edge-list");
            if (it != kernelMap.end()) kernels = it->second;
            defaultKernel = 0;
            implIds.emplace_back(-1, "edge-list");
            implementations.emplace_back(kernels);
        }

        if (!logFile.empty()) {
            setupLogging(missingGraphProps, missingAlgoProps);
        } else {
            logGraphProps = []() {};
            logAlgorithmProps = []() {};
        }

        bool uninitialised = false;
        mapKernels([&uninitialised](auto &&k) { if (!k) uninitialised = true; }, kernels);
        if (uninitialised) reportError("No edge list implementation found!");
    }

    void cleanupRun() override final
    {
        defaultKernel = -1;
        lastKernel = -1;
        implIds.clear();
        implementations.clear();
        if (modelHandle) {
            if (dlclose(modelHandle)) reportError("dlclose() failed!\n", dlerror());
            modelHandle = nullptr;
        }
        if (!logFile.empty()) propLog = std::ofstream();
        for (auto &kv : graphProperties) kv.second.get().reset();
        for (auto &kv : algorithmProperties) kv.second.get().reset();
        ownedProps.clear();
    }

  private:
    void setupPredictor(const char *lib, prop_set &graphProps, prop_set &algoProps)
    {
        typedef std::tuple<std::string, int64_t, size_t, size_t, size_t> impl_tuple;
        typedef const std::vector<impl_tuple> implementations_t;
        typedef const std::map<std::string, std::reference_wrapper<double>> properties;

        modelHandle = dlopen(lib, RTLD_NOW);
        if (!modelHandle) reportError("dlopen() failed: ", lib, "\n", dlerror());

        auto *fn = safe_dlsym<int32_t()>(modelHandle, "lookup");
        lookup = [fn]() { return fn(); };
        const auto &impls = *safe_dlsym<implementations_t>(modelHandle, "implNames");
        const auto &params = *safe_dlsym<properties>(modelHandle, "propNames");
        defaultKernel = *safe_dlsym<int32_t>(modelHandle, "default_impl");

        bool missing = false;
        for (auto &kv : params) {
            auto g = graphProperties.find(kv.first);
            if (g != graphProperties.end()) {
                g->second.get() = kv.second;
                graphProps.erase(kv.first);
                continue;
            }
            auto a = algorithmProperties.find(kv.first);
            if (a != algorithmProperties.end()) {
                a->second.get() = kv.second;
                algoProps.erase(kv.first);
                continue;
            }
            std::cerr << "Missing property: " << kv.first << std::endl;
            missing = true;
        }

        implIds.resize(impls.size());
        implementations.resize(impls.size());
        for (auto &t : impls) {
            const auto &[name, implId, idx, warp, chunk] = t;
            auto it = kernelMap.find(name);
            if (it == kernelMap.end() || idx >= impls.size()) {
                std::cerr << "Missing implementation: " << name << std::endl;
                missing = true;
                continue;
            }
            implIds[idx] = impl_id(implId, name);
            implementations[idx] = it->second;
            const size_t w = warp, c = chunk;
            mapKernels([w, c](auto &&k) { if (k->isWarp) k->setWarpConfig(w, c); }, implementations[idx]);
        }
        if (missing) reportError("Missing properties/implementations!");
        if (defaultKernel < 0 || static_cast<size_t>(defaultKernel) >= implementations.size())
            reportError("Predictor default_impl out of range!");
        kernels = implementations[static_cast<size_t>(defaultKernel)];
    }

    void setupLogging(prop_set &missingGraphProps, prop_set &missingAlgoProps)
    {
        propLog = std::ofstream(logFile);
        propLog.imbue(std::locale::classic());
        lookup = [this, inner = lookup]() {
            int32_t r = inner();
            if (r == -1) r = lastKernel == -1 ? defaultKernel : lastKernel;
            propLog << "prediction:" << stepNum << ":" << implIds[static_cast<size_t>(r)].id << std::endl;
            return r;
        };
        // Properties no predictor claimed are logged from owned storage.
        for (auto &name : missingGraphProps) {
            ownedProps.push_back(std::make_unique<double>(0.0));
            graphProperties.find(name)->second.get() = std::ref(*ownedProps.back());
        }
        for (auto &name : missingAlgoProps) {
            ownedProps.push_back(std::make_unique<double>(0.0));
            algorithmProperties.find(name)->second.get() = std::ref(*ownedProps.back());
        }
        logGraphProps = [this]() {
            for (auto &kv : graphProperties)
                propLog << "graph:" << kv.first << ":" << kv.second.get().get() << std::endl;
        };
        logAlgorithmProps = [this]() {
            for (auto &kv : algorithmProperties)
                propLog << "step:" << stepNum << ":" << kv.first << ":" << kv.second.get().get() << std::endl;
        };
    }

    template<typename F>
    void mapKernels(F f, std::tuple<Kernels...> &t)
    { mapKernels(f, t, std::index_sequence_for<Kernels...>()); }

    template<typename F, size_t... I>
    void mapKernels(F f, std::tuple<Kernels...> &t, std::index_sequence<I...>)
    { (f(std::get<I>(t)), ...); }

    std::string logFile, model;
    void *modelHandle;
    std::function<int32_t()> lookup;
    std::ofstream propLog;
    std::function<void()> logGraphProps, logAlgorithmProps;
    int32_t lastKernel = -1, defaultKernel = -1;
    std::vector<impl_id> implIds;
    std::vector<std::tuple<Kernels...>> implementations;
    std::map<std::string, std::tuple<Kernels...>> kernelMap;
    std::vector<std::unique_ptr<double>> ownedProps;
    prop_ref vertices, edges;
    graph_prop min, lowerQuantile, mean, median, upperQuantile, max, stdDev;
    size_t stepNum = 0;
};

template<template<typename, typename, typename, bool> class AlgorithmBase, typename Platform, typename Vertex,
         typename Edge, typename... KernelArgs, typename... Kernels,
         typename Impl = AlgorithmBase<Platform, Vertex, Edge, true>>
std::unique_ptr<ImplementationBase>
make_switch_implementation(KernelMap<std::string, std::tuple<GraphKernel<Platform, Vertex, Edge, KernelArgs...>, Kernels...>> ks)
{
    using Kernel = GraphKernel<Platform, Vertex, Edge, KernelArgs...>;
    return std::make_unique<SwitchImplementation<Impl, Kernel, Kernels...>>(ks);
}

#endif
