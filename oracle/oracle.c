/*
 * oracle.c -- CPU restatement of merijn/Belewitte's BFS and PageRank kernels.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain C99 + OpenMP.
 *
 * Paths are relative to the reference repository.  Each kernel restatement
 * runs the CUDA kernel's grid-stride loop sequentially in thread order; for
 * the BFS kernels the result does not depend on that order (every variant
 * is a level-synchronous relaxation whose writes are only read as
 * `== depth`), for the PageRank atomics it fixes one of the orders the GPU
 * may pick.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#include <time.h>
#endif

/* Damping constant of pagerank/pagerank.hpp:9 (a float). */
static const float DAMPENING = 0.85f;

static double now_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------ BFS */

int64_t oracle_bfs_queue(const oracle_graph *g, uint32_t root, int32_t *depth)
{
    uint64_t V = g->V;
    if (root >= V) return 0;
    for (uint64_t v = 0; v < V; v++) depth[v] = ORACLE_UNVISITED;
    uint32_t *queue = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
    uint64_t head = 0, tail = 0;
    int32_t maxd = 0;
    depth[root] = 0;
    queue[tail++] = root;
    while (head < tail) {
        uint32_t u = queue[head++];
        int32_t du = depth[u];
        if (du > maxd) maxd = du;
        for (uint64_t j = g->off[u]; j < g->off[u + 1]; j++) {
            uint32_t w = g->col[j];
            if (depth[w] == ORACLE_UNVISITED) {
                depth[w] = du + 1;
                queue[tail++] = w;
            }
        }
    }
    free(queue);
    return (int64_t)maxd + 1;
}

/* atomicMin(&levels[i], v) > v, with the int/unsigned mix of the kernels:
 * returns 1 when the old value was larger (a strict decrease). */
static inline int atomic_min_decreased(int32_t *levels, uint32_t i, int32_t v)
{
    int32_t old = levels[i];
    if (v < old) levels[i] = v;
    return old > v;
}

/* bfs/edgelist.cu:3-22 (rev = 0) and bfs/rev_edgelist.cu:3-21 (rev = 1).
 * The EdgeList of GraphLoader.hpp:134-143 is (in = row, out = col) in CSR
 * order; the reverse EdgeList is built the same way from the reverse CSR,
 * so its in[] is the destination and out[] the source. */
static uint32_t level_edge_list(const oracle_graph *g, int rev, int32_t *levels,
                                int32_t depth)
{
    const uint64_t *off = rev ? g->roff : g->off;
    const uint32_t *col = rev ? g->rcol : g->col;
    int32_t nd = depth + 1;
    uint32_t count = 0;
    for (uint64_t row = 0; row < g->V; row++) {
        for (uint64_t j = off[row]; j < off[row + 1]; j++) {
            uint32_t in = (uint32_t)row, out = col[j];
            /* forward: test levels[in], relax levels[out];
             * reverse:  test levels[out], relax levels[in]. */
            uint32_t test = rev ? out : in, relax = rev ? in : out;
            if (levels[test] == depth && atomic_min_decreased(levels, relax, nd)) count++;
        }
    }
    return count;
}

/* bfs/vertex_push.cu:3-26 */
static uint32_t level_vertex_push(const oracle_graph *g, int32_t *levels, int32_t depth)
{
    uint32_t nd = (uint32_t)depth + 1; /* `unsigned newDepth` (:10) */
    uint32_t count = 0;
    for (uint64_t v = 0; v < g->V; v++) {
        if (levels[v] != depth) continue;
        for (uint64_t j = g->off[v]; j < g->off[v + 1]; j++) {
            int32_t old = levels[g->col[j]];
            if ((int32_t)nd < old) levels[g->col[j]] = (int32_t)nd;
            if ((uint32_t)old > nd) count++; /* int compared as unsigned (:19) */
        }
    }
    return count;
}

/* bfs/vertex_pull.cu:3-31 (reverse CSR: in-neighbours, first hit breaks) */
static uint32_t level_vertex_pull(const oracle_graph *g, int32_t *levels, int32_t depth)
{
    int32_t nd = depth + 1;
    uint32_t count = 0;
    for (uint64_t v = 0; v < g->V; v++) {
        if (levels[v] <= nd) continue;
        for (uint64_t j = g->roff[v]; j < g->roff[v + 1]; j++) {
            if (levels[g->rcol[j]] == depth) {
                levels[v] = nd;
                count++;
                break;
            }
        }
    }
    return count;
}

/* bfs/vertex_push_warp.cu:20-52: chunk levels are snapshotted (memcpy_SIMD
 * into shared memory) before the chunk's vertices are expanded; each of the
 * `warp` lanes relaxes every warp-th neighbour. */
static uint32_t level_vertex_push_warp(const oracle_graph *g, int32_t *levels,
                                       int32_t depth, uint32_t warp, uint32_t chunk)
{
    int32_t nd = depth + 1;
    uint32_t count = 0;
    int32_t *snap = (int32_t *)malloc(sizeof(int32_t) * chunk);
    for (uint64_t base = 0; base < g->V; base += chunk) {
        uint64_t end = g->V - base < chunk ? g->V - base : chunk;
        memcpy(snap, levels + base, sizeof(int32_t) * end);
        for (uint64_t v = 0; v < end; v++) {
            if (snap[v] != depth) continue;
            uint64_t s = g->off[base + v], n = g->off[base + v + 1] - s;
            for (uint32_t lane = 0; lane < warp; lane++)
                for (uint64_t i = lane; i < n; i += warp)
                    if (atomic_min_decreased(levels, g->col[s + i], nd)) count++;
        }
    }
    free(snap);
    return count;
}

/* bfs/vertex_pull_warp.cu:19-71: snapshot, any-lane test against the global
 * levels, count `== depth+1` in the chunk, write the chunk back. */
static uint32_t level_vertex_pull_warp(const oracle_graph *g, int32_t *levels,
                                       int32_t depth, uint32_t warp, uint32_t chunk)
{
    int32_t nd = depth + 1;
    uint32_t count = 0;
    (void)warp; /* lanes only split the any-test; the result is their OR */
    int32_t *snap = (int32_t *)malloc(sizeof(int32_t) * chunk);
    for (uint64_t base = 0; base < g->V; base += chunk) {
        uint64_t end = g->V - base < chunk ? g->V - base : chunk;
        memcpy(snap, levels + base, sizeof(int32_t) * end);
        for (uint64_t v = 0; v < end; v++) {
            if (snap[v] <= depth) continue;
            int hit = 0;
            for (uint64_t j = g->roff[base + v]; j < g->roff[base + v + 1]; j++)
                if (levels[g->rcol[j]] == depth) hit = 1;
            if (hit) snap[v] = nd;
        }
        for (uint64_t v = 0; v < end; v++)
            if (snap[v] == nd) count++;
        memcpy(levels + base, snap, sizeof(int32_t) * end);
    }
    free(snap);
    return count;
}

uint32_t oracle_bfs_level(const oracle_graph *g, int variant, int32_t *levels,
                          int32_t depth, uint32_t warp, uint32_t chunk)
{
    switch (variant) {
    case ORACLE_BFS_EDGE_LIST:
    case ORACLE_BFS_STRUCT_EDGE_LIST: /* AoS layout, same loop (bfs/struct_edgelist.cu) */
        return level_edge_list(g, 0, levels, depth);
    case ORACLE_BFS_REV_EDGE_LIST:
    case ORACLE_BFS_REV_STRUCT_EDGE_LIST:
        return level_edge_list(g, 1, levels, depth);
    case ORACLE_BFS_VERTEX_PUSH:
        return level_vertex_push(g, levels, depth);
    case ORACLE_BFS_VERTEX_PULL:
        return level_vertex_pull(g, levels, depth);
    case ORACLE_BFS_VERTEX_PUSH_WARP:
        return level_vertex_push_warp(g, levels, depth, warp, chunk);
    case ORACLE_BFS_VERTEX_PULL_WARP:
        return level_vertex_pull_warp(g, levels, depth, warp, chunk);
    default:
        return 0;
    }
}

/* bfs/bfs.cpp:67-107 */
int64_t oracle_bfs_run(const oracle_graph *g, uint32_t root, const int *schedule,
                       int schedule_len, uint32_t warp, uint32_t chunk,
                       int32_t *levels, uint32_t *counts, int64_t max_levels)
{
    if (root >= g->V) return 0;
    for (uint64_t v = 0; v < g->V; v++) levels[v] = ORACLE_UNVISITED;
    levels[root] = 0;
    int64_t curr = 0;
    uint32_t frontier;
    do {
        if (curr >= max_levels) return -1;
        int variant = schedule[curr % schedule_len];
        frontier = oracle_bfs_level(g, variant, levels, (int32_t)curr, warp, chunk);
        counts[curr] = frontier;
        curr++;
    } while (frontier);
    return curr;
}

int64_t oracle_bfs_vertex_push_omp(const oracle_graph *g, uint32_t root,
                                   int32_t *levels, uint32_t *counts,
                                   int64_t max_levels, int threads, double *seconds)
{
    if (seconds) *seconds = 0.0;
    if (root >= g->V) return 0;
    int64_t V = (int64_t)g->V;
    const int saved_threads = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    /* The "computation" scope of bfs/bfs.cpp:78-108: initialisation and
     * the level loop. */
    const double t0 = now_seconds();
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; v++) levels[v] = ORACLE_UNVISITED;
    levels[root] = 0;
    int64_t curr = 0;
    uint32_t frontier;
    do {
        if (curr >= max_levels) {
            omp_set_num_threads(saved_threads);
            return -1;
        }
        int32_t depth = (int32_t)curr, nd = depth + 1;
        uint32_t count = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : count)
        for (int64_t v = 0; v < V; v++) {
            if (__atomic_load_n(&levels[v], __ATOMIC_RELAXED) != depth) continue;
            for (uint64_t j = g->off[v]; j < g->off[v + 1]; j++) {
                int32_t *p = &levels[g->col[j]];
                int32_t old = __atomic_load_n(p, __ATOMIC_RELAXED);
                while (old > nd &&
                       !__atomic_compare_exchange_n(p, &old, nd, 1, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED)) {
                }
                if (old > nd) count++;
            }
        }
        frontier = count;
        counts[curr] = frontier;
        curr++;
    } while (frontier);
    if (seconds) *seconds = now_seconds() - t0;
    omp_set_num_threads(saved_threads);
    return curr;
}

/* ------------------------------------------------------------- PageRank */

void oracle_pr_out_degrees(const oracle_graph *g, uint32_t *deg)
{
    /* reverseCSRComputeDegrees (pagerank/pagerank.cu:106-123) counts the
     * appearances of u in the reverse CSR; every other form counts the same
     * multiset (edge sources). */
    memset(deg, 0, sizeof(uint32_t) * g->V);
    for (uint64_t j = 0; j < g->E; j++) deg[g->rcol[j]]++;
}

static int is_nodiv(int variant)
{
    return variant == ORACLE_PR_VERTEX_PULL_NODIV ||
           variant == ORACLE_PR_VERTEX_PULL_WARP_NODIV;
}

/* One iteration kernel: accumulates into np[] (zeroed by the caller). */
static void pr_iteration(const oracle_graph *g, int variant, uint32_t warp,
                         const uint32_t *deg, const float *pr, float *np)
{
    uint64_t V = g->V;
    switch (variant) {
    case ORACLE_PR_EDGE_LIST:
    case ORACLE_PR_STRUCT_EDGE_LIST:
        /* pagerank/edgelist.cu:14-22, struct_edgelist.cu:14-23 */
        for (uint64_t row = 0; row < V; row++)
            for (uint64_t j = g->off[row]; j < g->off[row + 1]; j++) {
                uint32_t d = deg[row];
                float r = 0.0f;
                if (d != 0) r = pr[row] / (float)d;
                np[g->col[j]] += r;
            }
        break;
    case ORACLE_PR_REV_EDGE_LIST:
    case ORACLE_PR_REV_STRUCT_EDGE_LIST:
        /* pagerank/rev_edgelist.cu:14-22: origin = reverse out[] (source).
         * Each row's sum is sequential, so rows run in parallel without
         * changing a bit. */
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t r = 0; r < (int64_t)V; r++) {
            const uint64_t row = (uint64_t)r;
            for (uint64_t j = g->roff[row]; j < g->roff[row + 1]; j++) {
                uint32_t src = g->rcol[j];
                uint32_t d = deg[src];
                float c = 0.0f;
                if (d != 0) c = pr[src] / (float)d;
                np[row] += c;
            }
        }
        break;
    case ORACLE_PR_VERTEX_PUSH:
        /* pagerank/vertex_push.cu:16-30 (degree from the CSR itself) */
        for (uint64_t v = 0; v < V; v++) {
            uint64_t s = g->off[v], e = g->off[v + 1];
            uint32_t d = (uint32_t)(e - s);
            if (d == 0) continue;
            float r = pr[v] / (float)d;
            for (uint64_t j = s; j < e; j++) np[g->col[j]] += r;
        }
        break;
    case ORACLE_PR_VERTEX_PUSH_WARP:
        /* pagerank/vertex_push_warp.cu:37-47: lane-strided atomics */
        for (uint64_t v = 0; v < V; v++) {
            uint64_t s = g->off[v], n = g->off[v + 1] - s;
            if (n == 0) continue;
            float r = pr[v] / (float)(uint32_t)n;
            for (uint32_t lane = 0; lane < warp; lane++)
                for (uint64_t i = lane; i < n; i += warp) np[g->col[s + i]] += r;
        }
        break;
    case ORACLE_PR_VERTEX_PULL:
        /* pagerank/vertex_pull.cu:16-30: sequential in-edge sum (rows in
         * parallel, each row's order unchanged) */
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t v = 0; v < (int64_t)V; v++) {
            float acc = 0.0f;
            for (uint64_t j = g->roff[v]; j < g->roff[v + 1]; j++) {
                uint32_t u = g->rcol[j];
                acc += pr[u] / (float)deg[u];
            }
            np[v] = acc;
        }
        break;
    case ORACLE_PR_VERTEX_PULL_NODIV:
        /* pagerank/vertex_pull_nodiv.cu:16-28 */
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t v = 0; v < (int64_t)V; v++) {
            float acc = 0.0f;
            for (uint64_t j = g->roff[v]; j < g->roff[v + 1]; j++) acc += pr[g->rcol[j]];
            np[v] = acc;
        }
        break;
    case ORACLE_PR_VERTEX_PULL_WARP:
    case ORACLE_PR_VERTEX_PULL_WARP_NODIV:
        /* pagerank/vertex_pull_warp.cu:36-47 and _nodiv.cu:35-45: per-lane
         * partial sums, each atomically added to np[v] in lane order. */
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t v = 0; v < (int64_t)V; v++) {
            uint64_t s = g->roff[v], n = g->roff[v + 1] - s;
            for (uint32_t lane = 0; lane < warp; lane++) {
                float part = 0.0f;
                for (uint64_t i = lane; i < n; i += warp) {
                    uint32_t u = g->rcol[s + i];
                    if (variant == ORACLE_PR_VERTEX_PULL_WARP)
                        part += pr[u] / (float)(int32_t)deg[u];
                    else
                        part += pr[u];
                }
                np[v] += part;
            }
        }
        break;
    default:
        break;
    }
}

void oracle_pagerank_f32(const oracle_graph *g, int variant, int iterations,
                         int mode, uint32_t warp, float *pr, double *resid)
{
    uint64_t V = g->V;
    uint32_t *deg = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
    float *np = (float *)calloc(V ? V : 1, sizeof(float));
    oracle_pr_out_degrees(g, deg);
    int nodiv = is_nodiv(variant);

    /* pagerank/pagerank.cpp:88: 1.0f / vertex_count in float */
    float init = 1.0f / (float)V;
    for (uint64_t v = 0; v < V; v++) {
        pr[v] = init;
        if (mode == ORACLE_PR_DANGLING_UNIFORM && nodiv && deg[v] != 0)
            pr[v] = init / (float)deg[v];
    }

    for (int it = 1; it <= iterations; it++) {
        int notLast = iterations > it;
        double dangling = 0.0;
        if (mode == ORACLE_PR_DANGLING_UNIFORM)
            for (uint64_t v = 0; v < V; v++)
                if (deg[v] == 0) dangling += (double)pr[v];

        pr_iteration(g, variant, warp, deg, pr, np);

        double l1 = 0.0;
        for (uint64_t v = 0; v < V; v++) {
            float nr;
            float old = pr[v];
            double oldRank = (double)old;
            if (nodiv && mode == ORACLE_PR_DANGLING_UNIFORM && deg[v] != 0)
                oldRank = (double)old * (double)deg[v];
            if (mode == ORACLE_PR_DANGLING_DROP) {
                if (nodiv) {
                    /* pagerank/pagerank.cu:154, float base, contracted FMA */
                    float base = (1.0f - DAMPENING) / (float)V;
                    nr = fmaf(DAMPENING, np[v], base);
                } else {
                    /* pagerank/pagerank.cu:132, double base */
                    nr = (float)(((1.0 - (double)DAMPENING) / (double)V) +
                                 (double)(DAMPENING * np[v]));
                }
            } else {
                double d = (double)DAMPENING;
                nr = (float)((1.0 - d) / (double)V +
                             d * ((double)np[v] + dangling / (double)V));
            }
            l1 += fabs((double)nr - oldRank);
            /* consolidateRankNoDiv stores contributions (:159) */
            if (nodiv && deg[v] != 0 && notLast) pr[v] = nr / (float)deg[v];
            else pr[v] = nr;
            np[v] = 0.0f;
        }
        if (resid) resid[it - 1] = l1;
    }
    free(deg);
    free(np);
}

void oracle_pagerank_f64(const oracle_graph *g, int iterations, int mode,
                         int nodiv_quirk, double *x)
{
    uint64_t V = g->V;
    uint32_t *deg = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
    double *nx = (double *)malloc(sizeof(double) * (V ? V : 1));
    oracle_pr_out_degrees(g, deg);
    double d = (double)DAMPENING;
    for (uint64_t v = 0; v < V; v++) x[v] = 1.0 / (double)V;
    for (int it = 1; it <= iterations; it++) {
        double dangling = 0.0;
        if (mode == ORACLE_PR_DANGLING_UNIFORM)
            for (uint64_t v = 0; v < V; v++)
                if (deg[v] == 0) dangling += x[v];
        /* Each vertex sums its in-edges sequentially, so the result does
         * not depend on the thread count. */
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t v = 0; v < (int64_t)V; v++) {
            double acc = 0.0;
            for (uint64_t j = g->roff[v]; j < g->roff[v + 1]; j++) {
                uint32_t u = g->rcol[j];
                acc += (nodiv_quirk && it == 1) ? x[u] : x[u] / (double)deg[u];
            }
            nx[v] = (1.0 - d) / (double)V + d * (acc + dangling / (double)V);
        }
        memcpy(x, nx, sizeof(double) * V);
    }
    free(deg);
    free(nx);
}

void oracle_pagerank_binned_fixed(const oracle_graph *g, int iterations, int nodiv_quirk, float *ranks)
{
    const uint64_t V = g->V;
    if (V == 0) return;
    uint32_t *deg = (uint32_t *)malloc(sizeof(uint32_t) * V);
    float *c = (float *)malloc(sizeof(float) * V);
    float *nc = (float *)malloc(sizeof(float) * V);
    oracle_pr_out_degrees(g, deg);
    int lg = 0;
    if (nodiv_quirk) {
        const double M = 1.0 + (double)g->E / (double)V;
        while (ldexp(1.0, lg) < M) lg++;
    }
    const int S = 62 - lg;
    const float init = 1.0f / (float)V;
    const float base = (1.0f - DAMPENING) / (float)V;
    /* Initial contributions: 1/V, divided by the out-degree unless the
     * first iteration is undivided (pagerank/pagerank.cpp:88). */
    for (uint64_t v = 0; v < V; v++) c[v] = (!nodiv_quirk && deg[v]) ? init / (float)deg[v] : init;
    for (int it = 1; it <= iterations; it++) {
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t v = 0; v < (int64_t)V; v++) {
            uint64_t acc = 0;
            for (uint64_t j = g->roff[v]; j < g->roff[v + 1]; j++)
                acc += (uint64_t)ldexp((double)c[g->rcol[j]], S);  /* exact product, truncated */
            const float sum = ldexpf((float)acc, -S);                /* one rounding to FP32 */
            const float nr = fmaf(DAMPENING, sum, base);
            ranks[v] = nr;
            nc[v] = deg[v] ? nr / (float)deg[v] : 0.0f;
        }
        float *t = c;
        c = nc;
        nc = t;
    }
    if (iterations < 1)
        for (uint64_t v = 0; v < V; v++) ranks[v] = init;
    free(deg);
    free(c);
    free(nc);
}

void oracle_pagerank_pull_omp(const oracle_graph *g, int iterations, float *pr,
                              int threads, double *seconds)
{
    int64_t V = (int64_t)g->V;
    /* The caller's thread count is restored on return, so a one-thread
     * baseline does not leave later parallel regions of the process on
     * one thread. */
    const int saved_threads = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    uint32_t *deg = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
    float *np = (float *)malloc(sizeof(float) * (V ? V : 1));
    /* The "computation" scope of pagerank/pagerank.cpp:95-127: degrees and
     * the iterations (buffers are allocated before it). */
    const double t0 = now_seconds();
    oracle_pr_out_degrees(g, deg);
    float init = 1.0f / (float)V;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; v++) pr[v] = deg[v] ? init / (float)deg[v] : init;
    const double d = (double)DAMPENING;
    for (int it = 1; it <= iterations; it++) {
        int notLast = iterations > it;
        double dangling = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : dangling)
        for (int64_t v = 0; v < V; v++)
            if (deg[v] == 0) dangling += (double)pr[v];
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t v = 0; v < V; v++) {
            float acc = 0.0f;
            for (uint64_t j = g->roff[v]; j < g->roff[v + 1]; j++) acc += pr[g->rcol[j]];
            np[v] = acc;
        }
#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < V; v++) {
            float nr = (float)((1.0 - d) / (double)V +
                               d * ((double)np[v] + dangling / (double)V));
            pr[v] = (deg[v] != 0 && notLast) ? nr / (float)deg[v] : nr;
        }
    }
    if (seconds) *seconds = now_seconds() - t0;
    free(deg);
    free(np);
    omp_set_num_threads(saved_threads);
}

typedef struct {
    float rank;
    uint64_t id;
} rank_id;

static int cmp_rank_id(const void *pa, const void *pb)
{
    const rank_id *a = (const rank_id *)pa, *b = (const rank_id *)pb;
    if (a->rank < b->rank) return -1;
    if (b->rank < a->rank) return 1;
    return a->id < b->id ? -1 : (a->id > b->id ? 1 : 0);
}

static int cmp_id(const void *pa, const void *pb)
{
    const rank_id *a = (const rank_id *)pa, *b = (const rank_id *)pb;
    return a->id < b->id ? -1 : (a->id > b->id ? 1 : 0);
}

void oracle_pagerank_bucket_order(const float *ranks, uint64_t V, uint64_t *order)
{
    /* pagerank/pagerank.cpp:12-35 mask_N_bits(16) + roundPrecision */
    const uint32_t mask = ~(uint32_t)0xFFFF;
    rank_id *b = (rank_id *)malloc(sizeof(rank_id) * (V ? V : 1));
    for (uint64_t i = 0; i < V; i++) {
        uint32_t bits;
        memcpy(&bits, &ranks[i], 4);
        bits &= mask;
        memcpy(&b[i].rank, &bits, 4);
        b[i].id = i;
    }
    qsort(b, V, sizeof(rank_id), cmp_rank_id);
    /* :154-163: buckets of 1000, end clamped to size-1 */
    for (uint64_t i = 0; i < V; i += 1000) {
        uint64_t end = i + 1000 < V - 1 ? i + 1000 : V - 1;
        if (end > i) qsort(b + i, end - i, sizeof(rank_id), cmp_id);
    }
    for (uint64_t i = 0; i < V; i++) order[i] = b[i].id;
    free(b);
}
