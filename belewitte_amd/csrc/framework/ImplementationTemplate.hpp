// Kernel wrappers, implementation templates and the per-level switch: the
// plugin API of ImplementationTemplate.hpp (merijn/Belewitte), with the same
// names and template signatures so an algorithm library written against
// the reference API compiles here (KernelBuilder, GraphKernel, KernelMap,
// ImplementationTemplate, SimpleImplementation, SwitchImplementation,
// make_implementation, make_switch_implementation, prop_ref, tag_t,
// work_division).
#ifndef BW_FRAMEWORK_IMPLEMENTATIONTEMPLATE_HPP
#define BW_FRAMEWORK_IMPLEMENTATIONTEMPLATE_HPP

#include <dlfcn.h>

#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <tuple>
#include <utility>
#include <vector>

#include "GraphLoader.hpp"
#include "ImplementationBase.hpp"
#include "Timer.hpp"
#include "Util.hpp"

#define tag_t(v) (tag_t<decltype(v), v>{})

template<typename T, T value>
struct tag_t {};

enum class work_division { vertex, edge };

template<typename Platform>
struct BaseKernel {
    template<typename, typename, typename, bool>
    friend struct ImplementationTemplate;

    using KernelType = typename Platform::kernel_type;

    BaseKernel(KernelType k, GraphRep rep, work_division w, bool warp)
      : isWarp(warp), representation(rep), workDivision(w), kernel(k), backend(Platform::get())
    {}
    virtual ~BaseKernel() {}

    const bool isWarp;
    const GraphRep representation;
    const work_division workDivision;

    virtual void setWarpConfig(size_t, size_t) {}

  protected:
    explicit BaseKernel(std::nullptr_t)
      : isWarp(false), representation({Rep::VertexCount, Dir::Forward}), workDivision(work_division::vertex),
        kernel(nullptr), backend(Platform::get())
    {}

    virtual size_t getSharedMemSize(size_t) { return 0; }

    KernelType kernel;
    Platform &backend;
};

namespace internals {

template<typename Platform, typename V, typename E, typename... Args>
struct GraphKernel : public BaseKernel<Platform> {
    using typename BaseKernel<Platform>::KernelType;
    template<typename T>
    using DevToHost = const typename Platform::template DevToHost<T>::type &;

    virtual explicit operator bool() const = 0;
    virtual void run(GraphLoader<Platform, V, E> &, DevToHost<Args>...) = 0;

    explicit GraphKernel(std::nullptr_t p) : BaseKernel<Platform>(p) {}
    GraphKernel(KernelType k, GraphRep rep, work_division w, bool warp) : BaseKernel<Platform>(k, rep, w, warp) {}
};

template<typename Platform, typename V, typename E, Rep rep, Dir dir, typename... Args>
struct GraphKernelImpl : GraphKernel<Platform, V, E, Args...> {
    using Graph = typename LoaderRep<rep, Platform, V, E>::GraphType;
    using Kernel = typename Platform::template kernel<Graph, Args...>::type;
    using typename GraphKernel<Platform, V, E, Args...>::KernelType;
    using GraphKernel<Platform, V, E, Args...>::backend;
    using GraphKernel<Platform, V, E, Args...>::kernel;
    template<typename T>
    using DevToHost = typename GraphKernel<Platform, V, E, Args...>::template DevToHost<T>;

    GraphKernelImpl(Kernel k, work_division w)
      : GraphKernel<Platform, V, E, Args...>(reinterpret_cast<KernelType>(k), {rep, dir}, w, false)
    {}

    explicit operator bool() const override { return true; }

    void run(GraphLoader<Platform, V, E> &loader, DevToHost<Args>... args) override
    {
        auto &graph = loader.template getGraph<rep, dir>();
        backend.runKernel(reinterpret_cast<Kernel>(kernel), graph, args...);
    }
};

template<typename Platform, typename V, typename E, Rep rep, Dir dir, typename... Args>
struct WarpKernel : GraphKernel<Platform, V, E, Args...> {
    using Graph = typename LoaderRep<rep, Platform, V, E>::GraphType;
    using Kernel = typename Platform::template kernel<size_t, size_t, Graph, Args...>::type;
    using typename GraphKernel<Platform, V, E, Args...>::KernelType;
    using GraphKernel<Platform, V, E, Args...>::backend;
    using GraphKernel<Platform, V, E, Args...>::kernel;
    template<typename T>
    using DevToHost = typename GraphKernel<Platform, V, E, Args...>::template DevToHost<T>;

    WarpKernel(Kernel k, work_division w, std::function<size_t(size_t)> mem)
      : GraphKernel<Platform, V, E, Args...>(reinterpret_cast<KernelType>(k), {rep, dir}, w, true),
        chunkMemory(std::move(mem)), warp_size(32), chunk_size(32)
    {}

    explicit operator bool() const override { return true; }

    void run(GraphLoader<Platform, V, E> &loader, DevToHost<Args>... args) override
    {
        auto &graph = loader.template getGraph<rep, dir>();
        backend.runKernel(reinterpret_cast<Kernel>(kernel), warp_size, chunk_size, graph, args...);
    }

    void setWarpConfig(size_t warp, size_t chunk) final
    {
        warp_size = warp;
        chunk_size = chunk;
    }

  protected:
    size_t getSharedMemSize(size_t blockSize) override
    { return (blockSize / warp_size) * chunkMemory(chunk_size); }

    std::function<size_t(size_t)> chunkMemory;
    size_t warp_size, chunk_size;
};

template<typename Platform, typename V, typename E, typename... Args>
struct NullKernel : GraphKernel<Platform, V, E, Args...> {
    template<typename T>
    using DevToHost = typename GraphKernel<Platform, V, E, Args...>::template DevToHost<T>;
    NullKernel() : GraphKernel<Platform, V, E, Args...>(nullptr) {}
    explicit operator bool() const override { return false; }
    void run(GraphLoader<Platform, V, E> &, DevToHost<Args>...) override {}
};

}  // namespace internals

template<typename P, typename V, typename E, typename... Args>
struct GraphKernel : std::shared_ptr<internals::GraphKernel<P, V, E, Args...>> {
    using Null = internals::NullKernel<P, V, E, Args...>;
    using Base = internals::GraphKernel<P, V, E, Args...>;
    template<typename T>
    using DevToHost = typename Base::template DevToHost<T>;

    GraphKernel() : GraphKernel(std::make_shared<Null>()) {}
    GraphKernel(std::nullptr_t) : GraphKernel(std::make_shared<Null>()) {}
    GraphKernel(const std::shared_ptr<Base> &f) : std::shared_ptr<Base>(f) {}

    explicit operator bool() const
    { return std::shared_ptr<Base>::operator bool() && static_cast<bool>(*this->get()); }

    GraphKernel &operator=(std::nullptr_t)
    {
        std::shared_ptr<Base>::operator=(std::make_shared<Null>());
        return *this;
    }

    void operator()(GraphLoader<P, V, E> &l, DevToHost<Args>... args) { this->get()->run(l, args...); }
};

template<typename P, typename V, typename E, Rep r, Dir d, typename... Args>
GraphKernel(const std::shared_ptr<internals::GraphKernelImpl<P, V, E, r, d, Args...>> &)
    -> GraphKernel<P, V, E, Args...>;

template<typename P, typename V, typename E, Rep r, Dir d, typename... Args>
GraphKernel(const std::shared_ptr<internals::WarpKernel<P, V, E, r, d, Args...>> &)
    -> GraphKernel<P, V, E, Args...>;

template<typename Platform, typename V, typename E>
struct KernelBuilder {
    template<Rep rep, Dir dir = Dir::Forward, typename... Args,
             typename Graph = typename LoaderRep<rep, Platform, V, E>::GraphType,
             typename KernelImpl = internals::GraphKernelImpl<Platform, V, E, rep, dir, Args...>>
    auto operator()(void (*k)(Graph, Args...), work_division w, tag_t<Rep, rep>, tag_t<Dir, dir> = {})
    {
        using K = typename KernelImpl::Kernel;
        auto impl = std::make_shared<KernelImpl>(reinterpret_cast<K>(k), w);
        return GraphKernel<Platform, V, E, Args...>(
            std::static_pointer_cast<internals::GraphKernel<Platform, V, E, Args...>>(impl));
    }

    template<Rep rep, Dir dir = Dir::Forward, typename... Args,
             typename Graph = typename LoaderRep<rep, Platform, V, E>::GraphType,
             typename KernelImpl = internals::WarpKernel<Platform, V, E, rep, dir, Args...>>
    auto operator()(void (*k)(size_t, size_t, Graph, Args...), work_division w, std::function<size_t(size_t)> mem,
                    tag_t<Rep, rep>, tag_t<Dir, dir> = {})
    {
        using K = typename KernelImpl::Kernel;
        auto impl = std::make_shared<KernelImpl>(reinterpret_cast<K>(k), w, std::move(mem));
        return GraphKernel<Platform, V, E, Args...>(
            std::static_pointer_cast<internals::GraphKernel<Platform, V, E, Args...>>(impl));
    }
};

template<typename Key, typename T>
class KernelMap : public std::map<Key, T>
{
  public:
    KernelMap() {}
    KernelMap(std::initializer_list<std::pair<Key, T>> l) : std::map<Key, T>(l.begin(), l.end()) {}

    KernelMap &operator+=(const KernelMap &other)
    {
        for (auto &kv : other)
            if (!this->insert(kv).second) throw std::domain_error("Key already exists!");
        return *this;
    }
};

template<typename T>
KernelMap(std::initializer_list<std::pair<const char *, T>>) -> KernelMap<std::string, T>;

template<typename T>
KernelMap(std::initializer_list<std::pair<std::string, T>>) -> KernelMap<std::string, T>;

template<bool switching>
struct ImplementationTemplateBase;

template<>
struct ImplementationTemplateBase<false> : public ImplementationBase {
    static constexpr bool isSwitching = false;
    virtual ~ImplementationTemplateBase() {}
};

class prop_ref;

template<>
struct ImplementationTemplateBase<true> : public ImplementationBase {
    static constexpr bool isSwitching = true;
    virtual ~ImplementationTemplateBase() {}
    virtual void predictInitial() = 0;
    virtual bool predict() = 0;

    refmap<std::string, prop_ref> graphProperties;
    refmap<std::string, prop_ref> algorithmProperties;
};

template<typename Platform, typename V, typename E, bool switching>
struct ImplementationTemplate : public ImplementationTemplateBase<switching> {
    typedef V Vertex;
    typedef E Edge;
    using ImplementationTemplateBase<switching>::options;
    using ImplementationTemplateBase<switching>::run_count;

    template<typename T>
    using alloc_t = typename Platform::template alloc_t<T>;

    template<typename... Args>
    using GraphKernel = ::GraphKernel<Platform, V, E, Args...> &;

  protected:
    Platform &backend;
    GraphLoader<Platform, V, E> loader;
    size_t vertex_count, edge_count;
    size_t warp_size, chunk_size;

    ImplementationTemplate()
      : backend(Platform::get()), vertex_count(0), edge_count(0), warp_size(32), chunk_size(32)
    {}

    void setKernelConfig(std::shared_ptr<BaseKernel<Platform>> k)
    {
        const auto &div = getWorkDivision(k->workDivision);
        const size_t shared = k->getSharedMemSize(div.first);
        backend.setWorkSizes(1, {div.first}, {div.second}, shared);
    }

    void setKernelConfig(work_division w, size_t sharedMem = 0)
    {
        const auto &div = getWorkDivision(w);
        backend.setWorkSizes(1, {div.first}, {div.second}, sharedMem);
    }

    virtual void loadGraph(const Graph<V, E> &) = 0;
    virtual void transferGraph() = 0;

    void loadGraph(const std::string filename) override final
    {
        Timer graphTransfer("graphTransfer", run_count);
        Graph<V, E> graph(filename);
        loadGraph(graph);
        vertex_count = graph.vertex_count;
        edge_count = graph.edge_count;
        vertexDivision = backend.computeDivision(vertex_count);
        edgeDivision = backend.computeDivision(edge_count);
        graphTransfer.start();
        transferGraph();
        graphTransfer.stop();
    }

    void freeGraph() override final { loader.freeGraph(); }

  private:
    const std::pair<size_t, size_t> &getWorkDivision(work_division w)
    { return w == work_division::edge ? edgeDivision : vertexDivision; }

    std::pair<size_t, size_t> vertexDivision, edgeDivision;
};

template<typename AlgorithmBase, typename... Kernels>
struct SimpleImplementation : public AlgorithmBase {
    using typename AlgorithmBase::Vertex;
    using typename AlgorithmBase::Edge;
    using AlgorithmBase::options;
    using AlgorithmBase::loader;
    using AlgorithmBase::warp_size;
    using AlgorithmBase::chunk_size;

    std::tuple<Kernels...> kernels;

    explicit SimpleImplementation(std::tuple<Kernels...> ks)
      : SimpleImplementation(ks, std::index_sequence_for<Kernels...>())
    {}

    template<size_t... I>
    SimpleImplementation(std::tuple<Kernels...> ks, std::index_sequence<I...>)
      : AlgorithmBase(std::get<I>(kernels)...), kernels(ks)
    {
        if ((std::get<I>(ks)->isWarp || ...)) {
            options.add('w', "warp", "NUM", warp_size, "Virtual warp size for warp variants.")
                .add('c', "chunk", "NUM", chunk_size, "Work chunk size for warp variants.");
        }
    }

  protected:
    void loadGraph(const Graph<Vertex, Edge> &graph) override final
    {
        std::vector<GraphRep> reps;
        mapKernels([&reps](auto &&k) { reps.push_back(k->representation); });
        loader.loadGraph(graph, reps);
    }

    void transferGraph() override final
    { mapKernels([this](auto &&k) { loader.transferGraph(k->representation); }); }

  private:
    void prepareRun() override
    {
        // wave64: a virtual warp may span the whole 64-lane wave.
        if (warp_size == 0) reportError("Warp size should be positive!");
        else if (warp_size > 64) reportError("Warp size should be at most 64!");
        else if (warp_size & (warp_size - 1)) reportError("Warp size should be a power of 2!");
        else if (chunk_size % warp_size != 0) reportError("Chunk size should be a multiple of warp size!");
        mapKernels([this](auto &&k) {
            if (k->isWarp) k->setWarpConfig(warp_size, chunk_size);
        });
    }

    template<typename F>
    void mapKernels(F f)
    { mapKernels(f, std::index_sequence_for<Kernels...>()); }

    template<typename F, size_t... I>
    void mapKernels(F f, std::index_sequence<I...>)
    { (f(std::get<I>(kernels)), ...); }
};

template<template<typename, typename, typename, bool> class BaseImpl, typename Platform, typename Vertex,
         typename Edge, typename... KernelArgs, typename... Kernels,
         typename Impl = BaseImpl<Platform, Vertex, Edge, false>>
std::unique_ptr<ImplementationBase>
make_implementation(std::tuple<GraphKernel<Platform, Vertex, Edge, KernelArgs...>, Kernels...> kernels)
{
    using Kernel = GraphKernel<Platform, Vertex, Edge, KernelArgs...>;
    return std::make_unique<SimpleImplementation<Impl, Kernel, Kernels...>>(kernels);
}

// A double that an implementation publishes for the per-level predictor;
// rebinding it (operator= with a reference_wrapper) points it at a
// predictor's own variable.
class prop_ref : public std::reference_wrapper<double>
{
    inline static double dummyProp = 0;

  public:
    prop_ref(prop_ref &&) = delete;
    prop_ref(const prop_ref &) = delete;

    prop_ref(const std::string &, ImplementationTemplateBase<false> &) : std::reference_wrapper<double>(dummyProp) {}

    prop_ref(const std::string &name, ImplementationTemplateBase<true> &cfg, bool graphProp = false)
      : std::reference_wrapper<double>(dummyProp)
    {
        if (graphProp) cfg.graphProperties.emplace(name, std::ref(*this));
        else cfg.algorithmProperties.emplace(name, std::ref(*this));
    }

    prop_ref &operator=(const std::reference_wrapper<double> &val)
    {
        std::reference_wrapper<double>::operator=(val);
        return *this;
    }

    double &operator=(const double &val) { return this->get() = val; }

    void reset() { operator=(std::ref(dummyProp)); }
};

#include "SwitchImplementation.hpp"

#endif
