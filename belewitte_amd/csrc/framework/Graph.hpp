// Host graph of the plugin framework: Graph<V,E> (utils/Graph.hpp of
// merijn/Belewitte) backed by libbwgraph, which reads the same `.graph`
// v0/v1 files.  Exposes the counts, the raw CSR / reverse-CSR arrays and
// degreeStatistics; the file is loaded in bulk instead of being accessed
// element by element through an mmap accessor.
#ifndef BW_FRAMEWORK_GRAPH_HPP
#define BW_FRAMEWORK_GRAPH_HPP

#include <memory>
#include <string>

#include "Timer.hpp"
#include "Util.hpp"
#include "belewitte_graph.h"

enum class Degrees { in, out, abs };

template<typename T>
struct RawArray {
    const T *data = nullptr;
    uint64_t size = 0;
    T operator[](uint64_t i) const
    {
        checkError(i < size, "Index too large! Index: ", i, " Max: ", size);
        return data[i];
    }
};

template<typename V, typename E>
class MutableGraph
{
  public:
    explicit MutableGraph(const std::string &file)
      : fileName(file), handle(load(file), bw_graph_free), undirected(bw_graph_is_undirected(handle.get()) != 0),
        vertex_count(bw_graph_vertex_count(handle.get())), edge_count(bw_graph_edge_count(handle.get()))
    {
        raw_vertices = {bw_graph_offsets(handle.get(), 0), vertex_count + 1};
        raw_edges = {bw_graph_cols(handle.get(), 0), edge_count};
        raw_rev_vertices = {bw_graph_offsets(handle.get(), 1), vertex_count + 1};
        raw_rev_edges = {bw_graph_cols(handle.get(), 1), edge_count};
    }

    StatisticalSummary<double> degreeStatistics(Degrees d) const
    {
        double s[7];
        const int kind = d == Degrees::in ? 0 : d == Degrees::out ? 1 : 2;
        if (bw_graph_degree_stats(handle.get(), kind, s) != BW_OK)
            reportError("degreeStatistics: ", bw_graph_last_error());
        std::vector<double> dummy{0.0};
        StatisticalSummary<double> r(dummy);
        r.min = s[0];
        r.lowerQuantile = s[1];
        r.mean = s[2];
        r.median = s[3];
        r.upperQuantile = s[4];
        r.max = s[5];
        r.stdDev = s[6];
        return r;
    }

    const bw_graph *native() const { return handle.get(); }

    const std::string fileName;

  private:
    static bw_graph *load(const std::string &file)
    {
        bw_graph *g = nullptr;
        if (bw_graph_load(&g, file.c_str()) != BW_OK) reportError(bw_graph_last_error());
        return g;
    }

    std::shared_ptr<bw_graph> handle;

  public:
    const bool undirected;
    const uint64_t vertex_count, edge_count;
    RawArray<uint64_t> raw_vertices;
    RawArray<uint32_t> raw_edges;
    RawArray<uint64_t> raw_rev_vertices;
    RawArray<uint32_t> raw_rev_edges;
};

template<typename V, typename E>
using Graph = const MutableGraph<V, E>;

#endif
