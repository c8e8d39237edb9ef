/*
 * belewitte_graph.h -- C ABI of the host-side graph substrate.
 *
 * Builds, stores, loads and saves the graphs that the BFS and PageRank
 * kernels consume: seeded generators, self-loop removal and deduplication,
 * forward and reverse CSR with uint32 vertex ids and uint64 offsets, and the
 * reference's `.graph` v0/v1 file format.  Library: libbwgraph.so (plain C++17
 * and OpenMP, no GPU dependency).
 *
 * Reference interfaces replaced (paths relative to merijn/Belewitte):
 *   - utils/Graph.hpp:784-1063   MutableGraph / Graph<V,E>: mmap reader of the
 *     `.graph` format, raw CSR and reverse-CSR accessors, file-size check.
 *   - utils/Graph.hpp:952-974 + :1112-1119  writeEdges / output: the writer.
 *   - utils/Graph.hpp:1187-1207  degreeStatistics (graph properties of the
 *     per-level switch, ImplementationTemplate.hpp:622-637).
 *   - evolve/gen-graph.cpp:72-123  chain / star / mesh generators.
 *
 * Conventions.  Every function returns 0 on success and a non-zero status on
 * failure; bw_graph_last_error() then returns a message (thread-local).
 * Vertex ids are uint32.  Offsets are uint64 internally; a uint32 copy can be
 * exported when the edge count fits.  A directed graph stores the forward CSR
 * (out-edges, rows sorted by source then destination) and the reverse CSR
 * (in-edges, rows sorted by destination then source).  An undirected graph
 * stores one symmetric CSR that serves as both directions, exactly as the
 * `.graph` format does.
 */
#ifndef BELEWITTE_GRAPH_H
#define BELEWITTE_GRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bw_graph bw_graph;

/* Status codes. */
#define BW_OK 0
#define BW_ERR_ARG 1
#define BW_ERR_IO 2
#define BW_ERR_FORMAT 3
#define BW_ERR_NOMEM 4

const char *bw_graph_last_error(void);

/* R-MAT (Graph500 recursion, no noise, no vertex permutation).  Generates
 * edge_factor * 2^scale edges; edge i descends `scale` levels, choosing the
 * quadrant (0,0),(0,1),(1,0),(1,1) with probabilities a, b, c, 1-a-b-c from a
 * counter-based hash of (seed, i, level).  Self-loops are removed and
 * duplicates merged.  undirected != 0 stores both directions. */
int bw_graph_rmat(bw_graph **out, uint32_t scale, uint32_t edge_factor,
                  double a, double b, double c, uint64_t seed, int undirected);

/* R-MAT restricted to `vertex_count` vertices (not a power of two): levels
 * are drawn at scale ceil(log2(vertex_count)) and an edge with an endpoint
 * >= vertex_count is redrawn from the next counter.  edge_count edges. */
int bw_graph_rmat_n(bw_graph **out, uint64_t vertex_count, uint64_t edge_count,
                    double a, double b, double c, uint64_t seed, int undirected);

/* Erdos-Renyi G(n, m) with i.i.d. uniform endpoints: `pairs` endpoint pairs
 * drawn with replacement, self-loops dropped, duplicates merged. */
int bw_graph_erdos_renyi(bw_graph **out, uint64_t vertex_count, uint64_t pairs,
                         uint64_t seed, int undirected);

/* width x height 4-neighbour grid (undirected), vertex id = y*width + x.
 * Exactly round(rewire * M) of the M undirected edges (chosen by hash rank)
 * get their second endpoint replaced by a uniform vertex, redrawn while it
 * would form a self-loop or an edge already present. */
int bw_graph_grid2d(bw_graph **out, uint32_t width, uint32_t height,
                    double rewire, uint64_t seed);

/* Closed-form fixtures of evolve/gen-graph.cpp:72-123.
 * chain: i -> i+1 for i < n (n+1 vertices).  star: 0 -> i+1 for i < n.
 * mesh: H x W lattice, right and down edges. */
int bw_graph_chain(bw_graph **out, uint64_t n, int undirected);
int bw_graph_star(bw_graph **out, uint64_t n, int undirected);
int bw_graph_mesh(bw_graph **out, uint64_t height, uint64_t width,
                  int undirected);

/* Graph from an explicit edge list (src[i] -> dst[i]).  keep_self_loops = 0
 * drops them; duplicates are always merged. */
int bw_graph_from_edges(bw_graph **out, uint64_t vertex_count, uint64_t n,
                        const uint32_t *src, const uint32_t *dst,
                        int undirected, int keep_self_loops);

/* `.graph` file format, v0 (int32 header [undirected, V, E]) and v1 (10
 * uint32 header words; utils/Graph.hpp:784-807).  Loading copies the data
 * into memory; a wrong file size is BW_ERR_FORMAT ("Invalid file size!"). */
int bw_graph_load(bw_graph **out, const char *path);
int bw_graph_save(const bw_graph *g, const char *path);

void bw_graph_free(bw_graph *g);

uint64_t bw_graph_vertex_count(const bw_graph *g);
uint64_t bw_graph_edge_count(const bw_graph *g);
int bw_graph_is_undirected(const bw_graph *g);

/* Borrowed pointers, valid until bw_graph_free.  reverse != 0 selects the
 * in-edge CSR (the same arrays for an undirected graph). */
const uint64_t *bw_graph_offsets(const bw_graph *g, int reverse);
const uint32_t *bw_graph_cols(const bw_graph *g, int reverse);

/* Copies the offsets narrowed to uint32 into dst[V+1]; BW_ERR_ARG when the
 * edge count does not fit in 32 bits. */
int bw_graph_offsets32(const bw_graph *g, int reverse, uint32_t *dst);

/* Source vertex of maximum out-degree, ties to the lowest id. */
uint32_t bw_graph_max_out_degree_vertex(const bw_graph *g);

/* Degree statistics as in utils/Graph.hpp degreeStatistics: kind 0 = in,
 * 1 = out, 2 = abs (in + out for directed graphs, out for undirected).
 * out[7] = {min, lower quantile, mean, median, upper quantile, max, stddev}
 * with the quantile and stddev definitions of
 * utils/StatisticalSummary.hpp:23-71. */
int bw_graph_degree_stats(const bw_graph *g, int kind, double out[7]);

/* Relabels vertices by decreasing out-degree (ties by id), the experiment
 * of reorder-graph.cpp:40-128.  perm_out[old] = new id (optional).  A pure
 * permutation of each stored CSR: duplicate edges and self-loops keep
 * their multiplicity, so the edge count never changes; rows come out
 * sorted. */
int bw_graph_reorder_by_degree(bw_graph **out, const bw_graph *g,
                               uint32_t *perm_out);

/* 64-bit counter-based hash used by every generator (splitmix64 finaliser
 * over seed + stream + counter); exposed so tests can recompute draws. */
uint64_t bw_hash64(uint64_t seed, uint64_t stream, uint64_t counter);

#ifdef __cplusplus
}
#endif
#endif
