// Graph representations handed to kernels (GraphRep.hpp of
// merijn/Belewitte, same field order and types).  With the HIP backend a
// kernel launcher receives the host copy of one of these structs whose array
// members hold device pointers; the launcher passes the arrays to the C ABI
// as plain kernel arguments, so no kernel dereferences a device-side
// descriptor.
#ifndef BW_FRAMEWORK_GRAPHREP_HPP
#define BW_FRAMEWORK_GRAPHREP_HPP

#include <cstdint>

template<typename EDGE>
struct edge {
    EDGE in, out;
    edge(EDGE i, EDGE o) : in(i), out(o) {}
};

template<typename EDGE>
struct EdgeList {
    uint64_t vertex_count, edge_count;
    EDGE *inEdges;
    EDGE *outEdges;
};

template<typename EDGE>
struct StructEdgeList {
    uint64_t vertex_count, edge_count;
    edge<EDGE> *edges;
};

template<typename VERTEX, typename EDGE>
struct EdgeListCSR {
    uint64_t vertex_count, edge_count;
    VERTEX *vertices;
    EDGE *inEdges;
    EDGE *outEdges;
};

template<typename VERTEX, typename EDGE>
struct StructEdgeListCSR {
    uint64_t vertex_count, edge_count;
    VERTEX *vertices;
    edge<EDGE> *edges;
};

template<typename VERTEX, typename EDGE>
struct CSR {
    uint64_t vertex_count, edge_count;
    VERTEX *vertices;
    EDGE *edges;
};

#endif
