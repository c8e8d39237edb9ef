// Host-to-device graph representations: GraphLoader<Platform,V,E> with the
// API of GraphLoader.hpp (merijn/Belewitte): Rep, Dir, GraphRep, LoaderRep,
// loadGraph(graph, reps), transferGraph(rep), getGraph<rep,dir>(),
// freeGraph().  Unlike the reference (which builds every representation for
// both directions before keeping the requested ones), only the requested
// representations are built, straight from the bulk CSR arrays.
#ifndef BW_FRAMEWORK_GRAPHLOADER_HPP
#define BW_FRAMEWORK_GRAPHLOADER_HPP

#include <ostream>
#include <tuple>
#include <vector>

#include "Graph.hpp"
#include "GraphRep.hpp"
#include "HIPBackend.hpp"

enum class Rep : char { VertexCount, EdgeCount, EdgeList, StructEdgeList, EdgeListCSR, StructEdgeListCSR, CSR };

enum class Dir : char { Forward, Reverse };

struct GraphRep {
    Rep representation;
    Dir direction;
};

inline std::ostream &operator<<(std::ostream &os, const Rep &r)
{
    static const char *names[] = {"VertexCount", "EdgeCount", "EdgeList", "StructEdgeList",
                                  "EdgeListCSR", "StructEdgeListCSR", "CSR"};
    return os << names[static_cast<int>(r)];
}

inline std::ostream &operator<<(std::ostream &os, const Dir &d)
{ return os << (d == Dir::Forward ? "Forward" : "Reverse"); }

template<Rep rep, typename Platform, typename V, typename E>
struct LoaderRep;

template<typename Platform, typename V, typename E>
class GraphLoader
{
    template<Rep, typename, typename, typename>
    friend struct LoaderRep;

    template<typename T>
    using pair = std::tuple<T, T>;
    template<typename T>
    using alloc_t = typename Platform::template alloc_t<T>;

    template<typename T>
    static T &pick(pair<T> &p, Dir d)
    { return d == Dir::Forward ? std::get<0>(p) : std::get<1>(p); }

    // Host arrays of one direction of the input graph, narrowed to V/E.
    struct Source {
        const uint64_t *off;
        const uint32_t *col;
    };

    Source source(const Graph<V, E> &g, Dir d) const
    {
        if (d == Dir::Forward) return {g.raw_vertices.data, g.raw_edges.data};
        return {g.raw_rev_vertices.data, g.raw_rev_edges.data};
    }

    void fillOffsets(V *dst, const Source &s) const
    {
        for (uint64_t v = 0; v <= vertices; v++) dst[v] = static_cast<V>(s.off[v]);
    }

    void fillEndpoints(E *in, E *out, edge<E> *pairs, const Source &s) const
    {
        for (uint64_t v = 0; v < vertices; v++) {
            for (uint64_t j = s.off[v]; j < s.off[v + 1]; j++) {
                if (in) in[j] = static_cast<E>(v);
                if (out) out[j] = static_cast<E>(s.col[j]);
                if (pairs) {
                    pairs[j].in = static_cast<E>(v);
                    pairs[j].out = static_cast<E>(s.col[j]);
                }
            }
        }
    }

    template<typename Struct>
    alloc_t<Struct> header(Platform &p)
    {
        auto a = p.template allocConstant<Struct>();
        a->vertex_count = vertices;
        a->edge_count = edges;
        return a;
    }

    void build(const Graph<V, E> &g, GraphRep r)
    {
        Platform &p = Platform::get();
        const Dir d = r.direction;
        const Source s = source(g, d);
        const size_t nE = std::max<uint64_t>(edges, 1);
        switch (r.representation) {
        case Rep::VertexCount:
        case Rep::EdgeCount:
            break;
        case Rep::EdgeList: {
            auto &a = pick(edgeList, d);
            if (a) break;
            a = header<EdgeList<E>>(p);
            a.allocLocal(&a->inEdges, nE);
            a.allocLocal(&a->outEdges, nE);
            fillEndpoints(a->inEdges, a->outEdges, nullptr, s);
            break;
        }
        case Rep::StructEdgeList: {
            auto &a = pick(structEdgeList, d);
            if (a) break;
            a = header<StructEdgeList<E>>(p);
            a.allocLocal(&a->edges, nE);
            fillEndpoints(nullptr, nullptr, a->edges, s);
            break;
        }
        case Rep::EdgeListCSR: {
            auto &a = pick(edgeListCSR, d);
            if (a) break;
            a = header<EdgeListCSR<V, E>>(p);
            a.allocLocal(&a->vertices, vertices + 1);
            a.allocLocal(&a->inEdges, nE);
            a.allocLocal(&a->outEdges, nE);
            fillOffsets(a->vertices, s);
            fillEndpoints(a->inEdges, a->outEdges, nullptr, s);
            break;
        }
        case Rep::StructEdgeListCSR: {
            auto &a = pick(structEdgeListCSR, d);
            if (a) break;
            a = header<StructEdgeListCSR<V, E>>(p);
            a.allocLocal(&a->vertices, vertices + 1);
            a.allocLocal(&a->edges, nE);
            fillOffsets(a->vertices, s);
            fillEndpoints(nullptr, nullptr, a->edges, s);
            break;
        }
        case Rep::CSR: {
            auto &a = pick(csr, d);
            if (a) break;
            a = header<CSR<V, E>>(p);
            a.allocLocal(&a->vertices, vertices + 1);
            a.allocLocal(&a->edges, nE);
            fillOffsets(a->vertices, s);
            for (uint64_t j = 0; j < edges; j++) a->edges[j] = static_cast<E>(s.col[j]);
            break;
        }
        }
    }

    template<typename F>
    void visit(GraphRep r, F &&f)
    {
        const Dir d = r.direction;
        switch (r.representation) {
        case Rep::VertexCount:
        case Rep::EdgeCount: break;
        case Rep::EdgeList: f(pick(edgeList, d)); break;
        case Rep::StructEdgeList: f(pick(structEdgeList, d)); break;
        case Rep::EdgeListCSR: f(pick(edgeListCSR, d)); break;
        case Rep::StructEdgeListCSR: f(pick(structEdgeListCSR, d)); break;
        case Rep::CSR: f(pick(csr, d)); break;
        }
    }

  public:
    GraphLoader() {}

    template<Rep rep, Dir dir>
    typename LoaderRep<rep, Platform, V, E>::GraphType &getGraph()
    { return pick(this->*LoaderRep<rep, Platform, V, E>::field, dir); }

    std::pair<size_t, size_t> loadGraph(const Graph<V, E> &graph, GraphRep rep)
    { return loadGraph(graph, std::vector<GraphRep>{rep}); }

    std::pair<size_t, size_t> loadGraph(const Graph<V, E> &graph, const std::vector<GraphRep> &reps)
    {
        vertices = graph.vertex_count;
        edges = graph.edge_count;
        vertexCount = {vertices, vertices};
        edgeCount = {edges, edges};
        for (auto r : reps) build(graph, r);
        return {vertices, edges};
    }

    void transferGraph(GraphRep rep)
    { visit(rep, [](auto &a) { a.copyHostToDev(); }); }

    void freeGraph()
    {
        for (Rep r : {Rep::EdgeList, Rep::StructEdgeList, Rep::EdgeListCSR, Rep::StructEdgeListCSR, Rep::CSR})
            for (Dir d : {Dir::Forward, Dir::Reverse}) visit({r, d}, [](auto &a) { a.free(); });
    }

  private:
    uint64_t vertices = 0, edges = 0;
    pair<size_t> vertexCount;
    pair<size_t> edgeCount;
    pair<alloc_t<EdgeList<E>>> edgeList;
    pair<alloc_t<StructEdgeList<E>>> structEdgeList;
    pair<alloc_t<EdgeListCSR<V, E>>> edgeListCSR;
    pair<alloc_t<StructEdgeListCSR<V, E>>> structEdgeListCSR;
    pair<alloc_t<CSR<V, E>>> csr;
};

// LoaderRep<rep>: the loader member holding a representation and the
// type a kernel launcher is handed for it.
template<typename Platform, typename V, typename E>
struct LoaderRep<Rep::VertexCount, Platform, V, E> {
    using Loader = GraphLoader<Platform, V, E>;
    static constexpr auto Loader::*field = &Loader::vertexCount;
    using GraphType = size_t;
};

template<typename Platform, typename V, typename E>
struct LoaderRep<Rep::EdgeCount, Platform, V, E> {
    using Loader = GraphLoader<Platform, V, E>;
    static constexpr auto Loader::*field = &Loader::edgeCount;
    using GraphType = size_t;
};

template<typename Platform, typename V, typename E>
struct LoaderRep<Rep::EdgeList, Platform, V, E> {
    using Loader = GraphLoader<Platform, V, E>;
    static constexpr auto Loader::*field = &Loader::edgeList;
    using GraphType = typename Platform::template alloc_t<EdgeList<E>> &;
};

template<typename Platform, typename V, typename E>
struct LoaderRep<Rep::StructEdgeList, Platform, V, E> {
    using Loader = GraphLoader<Platform, V, E>;
    static constexpr auto Loader::*field = &Loader::structEdgeList;
    using GraphType = typename Platform::template alloc_t<StructEdgeList<E>> &;
};

template<typename Platform, typename V, typename E>
struct LoaderRep<Rep::EdgeListCSR, Platform, V, E> {
    using Loader = GraphLoader<Platform, V, E>;
    static constexpr auto Loader::*field = &Loader::edgeListCSR;
    using GraphType = typename Platform::template alloc_t<EdgeListCSR<V, E>> &;
};

template<typename Platform, typename V, typename E>
struct LoaderRep<Rep::StructEdgeListCSR, Platform, V, E> {
    using Loader = GraphLoader<Platform, V, E>;
    static constexpr auto Loader::*field = &Loader::structEdgeListCSR;
    using GraphType = typename Platform::template alloc_t<StructEdgeListCSR<V, E>> &;
};

template<typename Platform, typename V, typename E>
struct LoaderRep<Rep::CSR, Platform, V, E> {
    using Loader = GraphLoader<Platform, V, E>;
    static constexpr auto Loader::*field = &Loader::csr;
    using GraphType = typename Platform::template alloc_t<CSR<V, E>> &;
};


#endif
