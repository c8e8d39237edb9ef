/*
 * oracle.h -- CPU restatement of merijn/Belewitte's BFS and PageRank kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.  The product path (libbelewitte_hip)
 * does not link or call it.
 *
 * Every function cites the reference file:line it restates.  The reference
 * has no CPU implementation of these kernels; each restatement executes the
 * kernel's grid-stride loop sequentially, in thread order, with the same
 * integer and FP32 arithmetic.  What pins each function (golden vectors,
 * closed forms, invariants, brute force) is listed in tests/test_oracle.py.
 */
#ifndef BELEWITTE_ORACLE_H
#define BELEWITTE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_UNVISITED 2147483647 /* INT_MAX, bfs/bfs.cpp:71-72 */

/* BFS variant ids (bfs/bfs.cpp:121-207 registration order). */
enum oracle_bfs_variant {
    ORACLE_BFS_EDGE_LIST = 0,
    ORACLE_BFS_REV_EDGE_LIST = 1,
    ORACLE_BFS_STRUCT_EDGE_LIST = 2,
    ORACLE_BFS_REV_STRUCT_EDGE_LIST = 3,
    ORACLE_BFS_VERTEX_PUSH = 4,
    ORACLE_BFS_VERTEX_PULL = 5,
    ORACLE_BFS_VERTEX_PUSH_WARP = 6,
    ORACLE_BFS_VERTEX_PULL_WARP = 7
};

/* PageRank variant ids (pagerank/pagerank.cpp:205-351). */
enum oracle_pr_variant {
    ORACLE_PR_EDGE_LIST = 0,
    ORACLE_PR_REV_EDGE_LIST = 1,
    ORACLE_PR_STRUCT_EDGE_LIST = 2,
    ORACLE_PR_REV_STRUCT_EDGE_LIST = 3,
    ORACLE_PR_VERTEX_PUSH = 4,
    ORACLE_PR_VERTEX_PULL = 5,
    ORACLE_PR_VERTEX_PUSH_WARP = 6,
    ORACLE_PR_VERTEX_PULL_WARP = 7,
    ORACLE_PR_VERTEX_PULL_NODIV = 8,
    ORACLE_PR_VERTEX_PULL_WARP_NODIV = 9
};

/* PageRank modes. */
#define ORACLE_PR_DANGLING_DROP 0    /* reference semantics */
#define ORACLE_PR_DANGLING_UNIFORM 1 /* dangling mass spread uniformly */

/* Graph handed to the oracle: forward CSR (out-edges) and reverse CSR
 * (in-edges); for an undirected graph both may point at the same arrays.
 * The edge-list forms are derived inside the oracle exactly as
 * GraphLoader.hpp:123-145 builds them (edge j of row i is (in=i, out=col[j])). */
typedef struct oracle_graph {
    uint64_t V, E;
    const uint64_t *off;
    const uint32_t *col;
    const uint64_t *roff;
    const uint32_t *rcol;
} oracle_graph;

/* ------------------------------------------------------------------ BFS */

/* Textbook FIFO-queue BFS over out-edges: the exact depth oracle.
 * depth[v] = hop distance from root, ORACLE_UNVISITED if unreachable.
 * Returns the number of level kernels the reference loop launches for the
 * same input (max depth + 1), 0 if root >= V (bfs/bfs.cpp:53). */
int64_t oracle_bfs_queue(const oracle_graph *g, uint32_t root, int32_t *depth);

/* One level of the given variant on `levels`, returning the frontier count
 * the kernel's Reduction would accumulate.  warp/chunk only matter for the
 * _warp variants (bfs/vertex_push_warp.cu:20-52, bfs/vertex_pull_warp.cu:19-71). */
uint32_t oracle_bfs_level(const oracle_graph *g, int variant, int32_t *levels,
                          int32_t depth, uint32_t warp, uint32_t chunk);

/* The level loop of bfs/bfs.cpp:67-107 with one variant per level taken
 * from schedule[level % schedule_len].  Writes levels[V] and the per-level
 * frontier counts (counts[k] for launch k, up to max_levels).  Returns the
 * number of launches, 0 if root >= V, -1 if max_levels was exceeded. */
int64_t oracle_bfs_run(const oracle_graph *g, uint32_t root, const int *schedule,
                       int schedule_len, uint32_t warp, uint32_t chunk,
                       int32_t *levels, uint32_t *counts, int64_t max_levels);

/* Baseline timing port: level-synchronous vertex-push (bfs/vertex_push.cu)
 * over `threads` OpenMP threads with a CAS atomic-min.  Same outputs as
 * oracle_bfs_run with ORACLE_BFS_VERTEX_PUSH.  *seconds (optional) receives
 * the monotonic-clock time of the initialisation and level loop. */
int64_t oracle_bfs_vertex_push_omp(const oracle_graph *g, uint32_t root,
                                   int32_t *levels, uint32_t *counts,
                                   int64_t max_levels, int threads, double *seconds);

/* ------------------------------------------------------------- PageRank */

/* Out-degrees as the variant's degree kernel computes them
 * (pagerank/pagerank.cu:40-123); every form yields outdeg(v). */
void oracle_pr_out_degrees(const oracle_graph *g, uint32_t *deg);

/* FP32 restatement of one variant run (pagerank/pagerank.cpp:86-127).
 * mode ORACLE_PR_DANGLING_DROP reproduces the reference exactly, including
 * the nodiv first-iteration quirk and the FP32/FP64 mix of the consolidate
 * kernels.  mode ORACLE_PR_DANGLING_UNIFORM is the corrected PageRank:
 * new = (float)((1-d)/N + d*(sum + D/N)) in FP64 with D the dangling mass,
 * nodiv contributions start pre-divided.  Atomic sums are accumulated in
 * sequential edge / lane order.  ranks[V] receives the final ranks; resid,
 * if non-null, receives the per-iteration L1 residual sum |new - old| (FP64).
 * warp is the virtual-warp width of the _warp variants. */
void oracle_pagerank_f32(const oracle_graph *g, int variant, int iterations,
                         int mode, uint32_t warp, float *ranks, double *resid);

/* FP64 oracle of the same recurrence (no FP32 rounding anywhere).
 * nodiv_quirk != 0 applies the reference's undivided first iteration. */
void oracle_pagerank_f64(const oracle_graph *g, int iterations, int mode,
                         int nodiv_quirk, double *ranks);

/* The binned iteration's arithmetic, restated to check it bit for bit:
 * vertex-pull-nodiv in DROP mode (pagerank/vertex_pull_nodiv.cu:3-29,
 * consolidateRankNoDiv pagerank/pagerank.cu:142-165) with each vertex's
 * in-edge sum taken in 64-bit fixed point, sum of floor(c_u * 2^S) over
 * its in-edges (every stored edge, repeats included), rounded once to FP32
 * and scaled by 2^-S; new rank fmaf(0.85f, sum, (1 - 0.85f) / V), the
 * contribution rank / outdeg (0 without out-edges).  S = 62 - ceil(log2 M),
 * M = 1, or 1 + E/V with the undivided first iteration (nodiv_quirk != 0).
 * This is what the device computes when each partition is a single chunk.
 * ranks[V] receives the final ranks. */
void oracle_pagerank_binned_fixed(const oracle_graph *g, int iterations, int nodiv_quirk, float *ranks);

/* Baseline timing port: config-mode vertex-pull-nodiv over `threads`
 * OpenMP threads; ranks identical to oracle_pagerank_f32 with
 * ORACLE_PR_VERTEX_PULL_NODIV and ORACLE_PR_DANGLING_UNIFORM. */
void oracle_pagerank_pull_omp(const oracle_graph *g, int iterations,
                              float *ranks, int threads, double *seconds);

/* Bucketed order output of pagerank/pagerank.cpp:142-168: masks the low
 * 16 mantissa bits, stable-sorts (rank, id), then re-sorts ids inside
 * buckets of 1000 (the last element stays outside the last bucket).
 * order[V] receives the ids in output order. */
void oracle_pagerank_bucket_order(const float *ranks, uint64_t V, uint64_t *order);

#ifdef __cplusplus
}
#endif
#endif
