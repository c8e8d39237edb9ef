/*
 * belewitte_hip.h -- C ABI of the MI355X (gfx950) graph-kernel library,
 * libbelewitte_hip.so.
 *
 * Plain pointers, sizes and status codes only.  Every function returns
 * BW_HIP_OK (0) on success or a non-zero status; bw_hip_last_error() then
 * describes the failure (thread-local).  Pointers documented as "device"
 * must come from bw_hip_dev_alloc (or any hipMalloc'ed memory of the current
 * device); "host" pointers are ordinary or pinned host memory.  A `stream`
 * is a hipStream_t passed as void*; NULL selects the null stream.
 *
 * Reference interfaces replaced (paths relative to merijn/Belewitte):
 *   - CUDA.hpp:30-288, CUDA.cpp:28-241 (CUDABackend: device query, pinned
 *     host + device allocations, copies, launch configuration) and
 *     utils/cuda_utils.hpp:1-18 (CUDA_CHK, memcpy_SIMD) -> bw_hip_* below.
 *   - Backend.cpp:52-70 (computeDivision) -> bw_hip_compute_division.
 *   - the bfs/ kernel files and bfs/bfs.cu (level kernels, frontier
 *     counter) -> bw_bfs_*.
 *   - the pagerank/ kernel files and pagerank/pagerank.cu (iteration,
 *     degree and consolidate kernels, diff counter) -> bw_pr_*.
 * Each declaration names the reference item it replaces.
 */
#ifndef BELEWITTE_HIP_H
#define BELEWITTE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */

#define BW_HIP_OK 0
#define BW_HIP_ERR_ARG 1      /* rejected on the host before any launch   */
#define BW_HIP_ERR_RUNTIME 2  /* a HIP runtime call failed                */
#define BW_HIP_ERR_NODEVICE 3 /* no GPU visible                           */
#define BW_HIP_ERR_NOMEM 4    /* device or pinned allocation failed       */
#define BW_HIP_ERR_LIMIT 5    /* a level / iteration bound was exceeded   */

/* Message of the last failure on this thread (CUDA.cpp:7-26 cudaAssert). */
const char *bw_hip_last_error(void);

/* Version string of the kernel library (KERNEL_COMMIT, Algorithm.hpp:8). */
const char *bw_hip_version(void);

/* ----------------------------------------------------------- backend */

/* Device properties (CUDA.cpp:55-156 queryDevice, CUDA.cpp:158-193). */
typedef struct bw_hip_props {
    char name[256];
    char arch[64];                 /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t compute_units;
    int32_t wave_size;             /* 64 on CDNA                              */
    int32_t max_threads_per_block;
    int32_t max_block_dims[3];
    int32_t max_grid_dims[3];
    uint64_t shared_mem_per_block; /* LDS bytes a block may allocate          */
    uint64_t total_global_mem;
    uint64_t l2_cache_bytes;
    int32_t clock_khz;
    int32_t memory_clock_khz;
    int32_t memory_bus_width_bits;
    int32_t pad_;
} bw_hip_props;

/* CUDABackend ctor, CUDA.cpp:197-217: *count = 0 and BW_HIP_OK when no
 * device is visible. */
int bw_hip_device_count(int32_t *count);
/* CUDA.cpp:158-193 setDevice */
int bw_hip_set_device(int32_t device);
int bw_hip_get_device_props(int32_t device, bw_hip_props *props);

/* cuda_alloc_t allocHostPtr / allocDevPtr, CUDA.hpp:34-61 */
int bw_hip_dev_alloc(void **ptr, size_t bytes);
int bw_hip_dev_free(void *ptr);
int bw_hip_host_alloc(void **ptr, size_t bytes); /* pinned */
int bw_hip_host_free(void *ptr);

/* copyHostToDevImpl / copyDevToHostImpl, CUDA.hpp:68-90.  Synchronous:
 * the copy is complete on return (as cudaMemcpy). */
int bw_hip_memcpy_h2d(void *dst_device, const void *src_host, size_t bytes);
int bw_hip_memcpy_d2h(void *dst_host, const void *src_device, size_t bytes);
int bw_hip_memcpy_d2d(void *dst_device, const void *src_device, size_t bytes);
/* Stream-ordered copy of any direction (hipMemcpyDefault). */
int bw_hip_memcpy_async(void *dst, const void *src, size_t bytes, void *stream);

/* Fill with a kernel (not hipMemset), so it is safe inside graph capture:
 * dst[i] = value for i < count, 32-bit elements. */
int bw_hip_fill_u32(void *dst_device, uint32_t value, uint64_t count, void *stream);

int bw_hip_device_sync(void);
int bw_hip_stream_create(void **stream); /* default priority */
int bw_hip_stream_destroy(void *stream);
int bw_hip_stream_sync(void *stream);
int bw_hip_event_create(void **event);
int bw_hip_event_destroy(void *event);
int bw_hip_event_record(void *event, void *stream);
int bw_hip_event_sync(void *event);
int bw_hip_event_elapsed_ms(float *ms, void *start, void *end);

/* Launch geometry of Backend.cpp:52-70 with the block size rounded to a
 * multiple of 64 (wave64) instead of 32. */
int bw_hip_compute_division(uint64_t count, uint64_t *threads_per_block,
                            uint64_t *blocks);

/* ------------------------------------------------------- launch config */

/* Launch configuration of one kernel (setWorkSizes, CUDA.cpp:195-241, plus
 * the -w/-c warp options of ImplementationTemplate.hpp:431-436).
 * block/grid = 0 lets the library choose; block must be a multiple of 64
 * and at most 1024.  warp is the virtual warp width of the _warp variants
 * (power of two, 1..64: wave64 replaces the reference's 32 limit); chunk is
 * a multiple of warp. */
typedef struct bw_launch {
    uint32_t block;
    uint32_t grid;
    uint32_t warp;
    uint32_t chunk;
    void *stream;
} bw_launch;

/* Frontier-count reductions of bfs/bfs.hpp:32-116 (name suffixes "",
 * "-bulk", "-warpreduce", "-blockreduce"). */
#define BW_RED_NORMAL 0 /* one atomic per discovered vertex          */
#define BW_RED_BULK 1   /* per-thread count, one atomic per thread    */
#define BW_RED_WAVE 2   /* wave64 reduce, one atomic per wave         */
#define BW_RED_BLOCK 3  /* wave reduce + LDS block reduce, one atomic */

/* --------------------------------------------------------------- BFS */

/* Level kernels.  Each call launches one level: for every vertex at
 * `depth` the kernel discovers its unvisited neighbours, sets them to
 * depth+1 and adds the number of newly discovered vertices to *frontier
 * (device uint32, zeroed by the caller; resetFrontier/getFrontier of
 * bfs/bfs.cu:7-18 become bw_hip_fill_u32 / bw_hip_memcpy_d2h).  levels is
 * the device int32 depth array, INT_MAX meaning unvisited (bfs/bfs.cpp:71).
 * CSR offsets are uint32 or uint64 (offset_bytes 4 or 8); ids are uint32.
 * EdgeList arrays are the SoA (in, out) of GraphRep.hpp:17-25, StructEdgeList
 * the AoS (in, out) pairs of GraphRep.hpp:27-38, both in the order
 * GraphLoader.hpp:134-143 builds them; the reverse forms are built from the
 * reverse CSR (in = destination, out = source). */

/* edgeListBfs<R>, bfs/edgelist.cu:3-22 */
int bw_bfs_edge_list(uint64_t edge_count, const uint32_t *in_edges,
                     const uint32_t *out_edges, int32_t *levels, int32_t depth,
                     uint32_t *frontier, int32_t reduction, const bw_launch *cfg);
/* revEdgeListBfs<R>, bfs/rev_edgelist.cu:3-21 */
int bw_bfs_rev_edge_list(uint64_t edge_count, const uint32_t *in_edges,
                         const uint32_t *out_edges, int32_t *levels,
                         int32_t depth, uint32_t *frontier, int32_t reduction,
                         const bw_launch *cfg);
/* structEdgeListBfs<R>, bfs/struct_edgelist.cu:3-22 */
int bw_bfs_struct_edge_list(uint64_t edge_count, const uint32_t *edges,
                            int32_t *levels, int32_t depth, uint32_t *frontier,
                            int32_t reduction, const bw_launch *cfg);
/* revStructEdgeListBfs<R>, bfs/rev_struct_edgelist.cu:3-21 */
int bw_bfs_rev_struct_edge_list(uint64_t edge_count, const uint32_t *edges,
                                int32_t *levels, int32_t depth,
                                uint32_t *frontier, int32_t reduction,
                                const bw_launch *cfg);
/* vertexPushBfs<R>, bfs/vertex_push.cu:3-26 (forward CSR) */
int bw_bfs_vertex_push(uint64_t vertex_count, const void *offsets,
                       int32_t offset_bytes, const uint32_t *cols,
                       int32_t *levels, int32_t depth, uint32_t *frontier,
                       int32_t reduction, const bw_launch *cfg);
/* vertexPullBfs<R>, bfs/vertex_pull.cu:3-31 (reverse CSR) */
int bw_bfs_vertex_pull(uint64_t vertex_count, const void *rev_offsets,
                       int32_t offset_bytes, const uint32_t *rev_cols,
                       int32_t *levels, int32_t depth, uint32_t *frontier,
                       int32_t reduction, const bw_launch *cfg);
/* vertexPushWarpBfs<R>, bfs/vertex_push_warp.cu:4-60 (forward CSR) */
int bw_bfs_vertex_push_warp(uint64_t vertex_count, const void *offsets,
                            int32_t offset_bytes, const uint32_t *cols,
                            int32_t *levels, int32_t depth, uint32_t *frontier,
                            int32_t reduction, const bw_launch *cfg);
/* vertexPullWarpBfs<R>, bfs/vertex_pull_warp.cu:4-71 (reverse CSR) */
int bw_bfs_vertex_pull_warp(uint64_t vertex_count, const void *rev_offsets,
                            int32_t offset_bytes, const uint32_t *rev_cols,
                            int32_t *levels, int32_t depth, uint32_t *frontier,
                            int32_t reduction, const bw_launch *cfg);

/* Device-resident representations of one graph, as GraphLoader.hpp:296-303
 * holds them.  Unused members may be NULL; a run fails with BW_HIP_ERR_ARG
 * when a scheduled variant needs a missing one. */
typedef struct bw_graph_dev {
    uint64_t vertex_count;
    uint64_t edge_count;
    const void *offsets;      /* forward CSR offsets [V+1]              */
    const uint32_t *cols;     /* forward CSR columns [E]                */
    const void *rev_offsets;  /* reverse CSR offsets [V+1]              */
    const uint32_t *rev_cols; /* reverse CSR columns [E]                */
    const uint32_t *el_in;    /* forward EdgeList in[]  (source)        */
    const uint32_t *el_out;   /* forward EdgeList out[] (destination)   */
    const uint32_t *rel_in;   /* reverse EdgeList in[]  (destination)   */
    const uint32_t *rel_out;  /* reverse EdgeList out[] (source)        */
    const uint32_t *sel;      /* forward StructEdgeList (in,out) pairs  */
    const uint32_t *rsel;     /* reverse StructEdgeList (in,out) pairs  */
    int32_t offset_bytes;     /* 4 or 8, for both CSR directions        */
    int32_t pad_;
} bw_graph_dev;

/* BFS variant ids (the 8 kernel families of bfs/bfs.cpp:121-207, then the
 * queue-based forms added for MI355X). */
#define BW_BFS_EDGE_LIST 0
#define BW_BFS_REV_EDGE_LIST 1
#define BW_BFS_STRUCT_EDGE_LIST 2
#define BW_BFS_REV_STRUCT_EDGE_LIST 3
#define BW_BFS_VERTEX_PUSH 4
#define BW_BFS_VERTEX_PULL 5
#define BW_BFS_VERTEX_PUSH_WARP 6
#define BW_BFS_VERTEX_PULL_WARP 7
#define BW_BFS_NUM_REFERENCE_VARIANTS 8

/* One launch of variant `variant` (dispatch over the kernels above). */
int bw_bfs_level(const bw_graph_dev *g, int32_t variant, int32_t reduction,
                 int32_t *levels, int32_t depth, uint32_t *frontier,
                 const bw_launch *cfg);

/* Relabels a graph on the device by decreasing out-degree, ties by id,
 * the same result as bw_graph_reorder_by_degree computes on the host
 * (the relabelling of reorder-graph.cpp:40-128): new id i is the i-th
 * vertex of that order; the rows of each stored direction are permuted and
 * their column lists mapped and sorted.  Inputs: the CSR (offsets, cols)
 * and, for a directed graph, the reverse CSR; an undirected graph passes
 * NULL for rev_offsets, rev_cols, new_rev_offsets and new_rev_cols.
 * Outputs are caller-allocated device arrays of the input sizes; perm
 * (device uint32[V], optional) receives the new id of each old vertex.
 * uint32 offsets, V < 2^31.  Synchronises the stream. */
int bw_graph_relabel_by_degree(uint64_t vertex_count, uint64_t edge_count,
                               const uint32_t *offsets, const uint32_t *cols,
                               const uint32_t *rev_offsets, const uint32_t *rev_cols,
                               uint32_t *new_offsets, uint32_t *new_cols,
                               uint32_t *new_rev_offsets, uint32_t *new_rev_cols,
                               uint32_t *perm, void *stream);

/* Expands the EdgeList / StructEdgeList of GraphLoader.hpp:123-145 from a
 * device CSR on the device: in[j] = row of j, out[j] = cols[j]; pairs gets
 * the interleaved (in, out) form.  Any output may be NULL. */
int bw_build_edge_lists(uint64_t vertex_count, uint64_t edge_count,
                        const void *offsets, int32_t offset_bytes,
                        const uint32_t *cols, uint32_t *in_out,
                        uint32_t *out_out, uint32_t *pairs_out, void *stream);

/* The level loop of bfs/bfs.cpp:67-107: levels = INT_MAX, levels[root] = 0,
 * then per level reset the counter, launch schedule[level % len], read the
 * counter back (a host round trip per level, as the reference does), stop
 * when no vertex was discovered.  counts_host[k] receives the count of launch
 * k and level_ms_host[k] (optional) its device time; *num_levels the number
 * of launches.  root >= V leaves levels untouched and reports 0 levels
 * (bfs/bfs.cpp:53).  BW_HIP_ERR_LIMIT if max_levels launches did not finish. */
int bw_bfs_run(const bw_graph_dev *g, const int32_t *schedule,
               int32_t schedule_len, int32_t reduction, uint32_t root,
               int32_t *levels, uint32_t *counts_host, float *level_ms_host,
               int64_t max_levels, int64_t *num_levels, const bw_launch *cfg);

/* No counterpart in the reference: bw_bfs_run for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same level loop, schedule,
 * results and errors, but it calls no hipMalloc/hipFree and creates no
 * event: the frontier counters, one uint32 per level, live in a static
 * device array of BW_BFS_STATIC_MAX_LEVELS entries (256 KiB, held by each
 * GPU the library runs on until the process ends), cleared by a fill
 * kernel at the start of each call.  max_levels above that bound is
 * rejected with BW_HIP_ERR_ARG before any launch; there is no per-level
 * timing.  Every call on one GPU shares its array, so calls that overlap
 * on one GPU (other host threads, other streams) race. */
#define BW_BFS_STATIC_MAX_LEVELS 65536
int bw_bfs_run_static(const bw_graph_dev *g, const int32_t *schedule,
                      int32_t schedule_len, int32_t reduction, uint32_t root,
                      int32_t *levels, uint32_t *counts_host,
                      int64_t max_levels, int64_t *num_levels,
                      const bw_launch *cfg);

/* ---------------------------------------------------------- PageRank */

/* Degree kernels, pagerank/pagerank.cu:19-123.  deg is device uint32[V]. */
/* zeroInitDegreesKernel :19-26 */
int bw_pr_zero_degrees(uint64_t vertex_count, uint32_t *deg, const bw_launch *cfg);
/* edgeListComputeDegrees :40-51 (origins = EdgeList in[]) and
 * reverseEdgeListComputeDegrees :79-91 (origins = reverse EdgeList out[]) */
int bw_pr_degrees_edge_list(uint64_t edge_count, const uint32_t *origins,
                            uint32_t *deg, const bw_launch *cfg);
/* structEdgeListComputeDegrees :53-64 (reverse = 0, counts .in) and
 * reverseStructEdgeListComputeDegrees :93-104 (reverse = 1, counts .out) */
int bw_pr_degrees_struct_edge_list(uint64_t edge_count, const uint32_t *edges,
                                   int32_t reverse, uint32_t *deg,
                                   const bw_launch *cfg);
/* CSRComputeDegrees :66-77 */
int bw_pr_degrees_csr(uint64_t vertex_count, const void *offsets,
                      int32_t offset_bytes, uint32_t *deg, const bw_launch *cfg);
/* reverseCSRComputeDegrees :106-123 */
int bw_pr_degrees_rev_csr(uint64_t vertex_count, const void *rev_offsets,
                          int32_t offset_bytes, const uint32_t *rev_cols,
                          uint32_t *deg, const bw_launch *cfg);

/* Iteration kernels: accumulate incoming contributions into new_pr
 * (zeroed by the previous consolidate).  pr/new_pr are device float[V]. */
/* edgeListPageRank, pagerank/edgelist.cu:3-23 */
int bw_pr_edge_list(uint64_t edge_count, const uint32_t *in_edges,
                    const uint32_t *out_edges, const uint32_t *deg,
                    const float *pr, float *new_pr, const bw_launch *cfg);
/* revEdgeListPageRank, pagerank/rev_edgelist.cu:3-23 */
int bw_pr_rev_edge_list(uint64_t edge_count, const uint32_t *in_edges,
                        const uint32_t *out_edges, const uint32_t *deg,
                        const float *pr, float *new_pr, const bw_launch *cfg);
/* structEdgeListPageRank, pagerank/struct_edgelist.cu:3-24 */
int bw_pr_struct_edge_list(uint64_t edge_count, const uint32_t *edges,
                           const uint32_t *deg, const float *pr, float *new_pr,
                           const bw_launch *cfg);
/* revStructEdgeListPageRank, pagerank/rev_struct_edgelist.cu:3-24 */
int bw_pr_rev_struct_edge_list(uint64_t edge_count, const uint32_t *edges,
                               const uint32_t *deg, const float *pr,
                               float *new_pr, const bw_launch *cfg);
/* vertexPushPageRank, pagerank/vertex_push.cu:3-31 (degree from the CSR) */
int bw_pr_vertex_push(uint64_t vertex_count, const void *offsets,
                      int32_t offset_bytes, const uint32_t *cols,
                      const float *pr, float *new_pr, const bw_launch *cfg);
/* vertexPullPageRank, pagerank/vertex_pull.cu:3-31 */
int bw_pr_vertex_pull(uint64_t vertex_count, const void *rev_offsets,
                      int32_t offset_bytes, const uint32_t *rev_cols,
                      const uint32_t *deg, const float *pr, float *new_pr,
                      const bw_launch *cfg);
/* vertexPullNoDivPageRank, pagerank/vertex_pull_nodiv.cu:3-29 */
int bw_pr_vertex_pull_nodiv(uint64_t vertex_count, const void *rev_offsets,
                            int32_t offset_bytes, const uint32_t *rev_cols,
                            const float *pr, float *new_pr,
                            const bw_launch *cfg);
/* vertexPushWarpPageRank, pagerank/vertex_push_warp.cu:4-49 */
int bw_pr_vertex_push_warp(uint64_t vertex_count, const void *offsets,
                           int32_t offset_bytes, const uint32_t *cols,
                           const float *pr, float *new_pr, const bw_launch *cfg);
/* vertexPullWarpPageRank, pagerank/vertex_pull_warp.cu:4-49 */
int bw_pr_vertex_pull_warp(uint64_t vertex_count, const void *rev_offsets,
                           int32_t offset_bytes, const uint32_t *rev_cols,
                           const uint32_t *deg, const float *pr, float *new_pr,
                           const bw_launch *cfg);
/* vertexPullWarpNoDivPageRank, pagerank/vertex_pull_warp_nodiv.cu:4-47 */
int bw_pr_vertex_pull_warp_nodiv(uint64_t vertex_count, const void *rev_offsets,
                                 int32_t offset_bytes, const uint32_t *rev_cols,
                                 const float *pr, float *new_pr,
                                 const bw_launch *cfg);

/* Pull plan of the four pull variants (vertexPullPageRank,
 * vertexPullNoDivPageRank, vertexPullWarpPageRank,
 * vertexPullWarpNoDivPageRank; pagerank/vertex_pull*.cu): the rows of the
 * reverse CSR with more than 2048 in-edges, listed once per graph on the
 * device, are summed one wave each, so a skewed graph's hubs run side by
 * side; every other row is summed by the workgroup owning it.  Results are
 * those of bw_pr_vertex_pull* (the non-warp variants bit for bit: each row
 * is still a sequential FP32 sum in CSR order).  The reverse offsets must
 * stay live while the plan is used. */
typedef struct bw_pr_pull_plan bw_pr_pull_plan;
int bw_pr_pull_plan_create(bw_pr_pull_plan **plan, uint64_t vertex_count,
                           const void *rev_offsets, int32_t offset_bytes,
                           void *stream);
int bw_pr_pull_plan_destroy(bw_pr_pull_plan *plan);
int bw_pr_pull_plan_info(const bw_pr_pull_plan *plan, uint32_t *long_rows);
/* One iteration of `variant` (BW_PR_VERTEX_PULL, _NODIV, _WARP,
 * _WARP_NODIV; cfg->warp is the virtual warp of the _warp variants). */
int bw_pr_pull_iteration(const bw_pr_pull_plan *plan, int32_t variant,
                         const uint32_t *rev_cols, const uint32_t *deg,
                         const float *pr, float *new_pr, const bw_launch *cfg);

/* Propagation blocking for the push forms (edgeListPageRank,
 * structEdgeListPageRank, vertexPushPageRank, vertexPushWarpPageRank of
 * pagerank/edgelist.cu:3-23, struct_edgelist.cu:3-24, vertex_push.cu:3-31,
 * vertex_push_warp.cu:4-49).  Those add each edge's contribution to
 * new_pr[destination] with a global atomic; here a layout built once per
 * graph (bw_pr_scatter_create: one pass over the destinations and a scan,
 * 6 bytes of bins per edge) lets one iteration bin every (destination,
 * contribution) by partition of 8192 rows and sum each partition in LDS
 * (FP64), writing each row once per chunk of its bin.  Sums are added in a
 * different order than the reference's atomics, which have no fixed order
 * either.  Source kinds: the forward EdgeList (a = in[], b = out[]), the
 * forward StructEdgeList (a = the (in, out) pairs, b unused) or the forward
 * CSR (a = offsets of offset_bytes, b = columns; the contribution is
 * pr[v] / outdeg(v) from the CSR itself, as vertex_push.cu does; each
 * edge's source row is expanded once when the layout is built, 4 more
 * bytes per edge).  The graph arrays must stay live while the layout is
 * used.  Graphs beyond
 * 2^25 vertices or 2^32 - 1 edges are refused with BW_HIP_ERR_ARG. */
#define BW_PR_SCATTER_EDGE_LIST 0
#define BW_PR_SCATTER_STRUCT 1
#define BW_PR_SCATTER_CSR 2
typedef struct bw_pr_scatter bw_pr_scatter;
int bw_pr_scatter_create(bw_pr_scatter **scatter, int32_t kind,
                         uint64_t vertex_count, uint64_t edge_count,
                         const void *a, const void *b, int32_t offset_bytes,
                         void *stream);
int bw_pr_scatter_destroy(bw_pr_scatter *scatter);
/* One iteration kernel: new_pr[v] += the contributions of v's in-edges
 * (new_pr zeroed by the previous consolidate, as the reference assumes).
 * deg (out-degrees) is needed by the edge-list kinds only. */
int bw_pr_scatter_iteration(const bw_pr_scatter *scatter, const uint32_t *deg,
                            const float *pr, float *new_pr, void *stream);

/* PageRank modes. */
#define BW_PR_DANGLING_DROP 0    /* reference: dangling mass is lost        */
#define BW_PR_DANGLING_UNIFORM 1 /* dangling mass redistributed uniformly   */

/* consolidateRank (nodiv = 0, pagerank/pagerank.cu:125-140) and
 * consolidateRankNoDiv (nodiv = 1, :142-165).  DROP mode keeps the
 * reference arithmetic bit for bit (double base term for nodiv = 0, float
 * base term with a fused multiply-add for nodiv = 1).  UNIFORM mode computes
 * new = (float)((1-d)/V + d*(new_pr + dangling/V)) in FP64, reading the FP64
 * dangling mass from *dangling (device).  The L1 residual |new - old| is
 * accumulated in FP64 into *resid (device, optional): the reference's float
 * `diff` (:4-17, :28-38).  Zeroes new_pr. */
int bw_pr_consolidate(uint64_t vertex_count, const uint32_t *deg, float *pr,
                      float *new_pr, int32_t not_last, int32_t nodiv,
                      int32_t mode, const double *dangling, double *resid,
                      const bw_launch *cfg);

/* pr[v] /= deg[v] where deg[v] != 0: the divided starting contributions of
 * the nodiv variants in UNIFORM mode (no reference counterpart: the
 * reference only has the undivided DROP-mode start, pagerank.cpp:88). */
int bw_pr_predivide(uint64_t vertex_count, const uint32_t *deg, float *pr,
                    const bw_launch *cfg);

/* FP64 dangling mass: *out += sum of pr[v] over deg[v] == 0 (device). */
int bw_pr_dangling_mass(uint64_t vertex_count, const uint32_t *deg,
                        const float *pr, double *out, const bw_launch *cfg);

/* PageRank variant ids (pagerank/pagerank.cpp:205-351). */
#define BW_PR_EDGE_LIST 0
#define BW_PR_REV_EDGE_LIST 1
#define BW_PR_STRUCT_EDGE_LIST 2
#define BW_PR_REV_STRUCT_EDGE_LIST 3
#define BW_PR_VERTEX_PUSH 4
#define BW_PR_VERTEX_PULL 5
#define BW_PR_VERTEX_PUSH_WARP 6
#define BW_PR_VERTEX_PULL_WARP 7
#define BW_PR_VERTEX_PULL_NODIV 8
#define BW_PR_VERTEX_PULL_WARP_NODIV 9
#define BW_PR_NUM_REFERENCE_VARIANTS 10

/* The run of pagerank/pagerank.cpp:86-127 for one variant: ranks = 1/V,
 * degrees once, `iterations` x (iteration kernel + consolidate), with no
 * host round trip inside the loop.  ranks_out (device float[V]) receives
 * the final ranks; resid_host (optional, double[iterations]) the L1
 * residual per iteration.  In UNIFORM mode the nodiv contributions start
 * pre-divided (the reference's undivided first iteration is a DROP-mode
 * quirk only). */
int bw_pr_run(const bw_graph_dev *g, int32_t variant, int32_t iterations,
              int32_t mode, float *ranks_out, double *resid_host,
              const bw_launch *cfg);

/* ------------------------------------------- fused pull PageRank (MI355X) */

/* "vertex-pull-nodiv-fused": the recurrence of vertexPullNoDivPageRank +
 * consolidateRankNoDiv (pagerank/vertex_pull_nodiv.cu:3-29,
 * pagerank/pagerank.cu:142-165) in one launch per iteration, mapped
 * CSR-Adaptive style (LDS-staged row blocks, long rows split across
 * workgroups), Kahan-compensated per-vertex sums, FP64 dangling mass and L1
 * residual.  A plan (work items + out-degrees) is built once per graph from
 * the device reverse CSR (uint32 offsets). */
typedef struct bw_pr_plan bw_pr_plan;

int bw_pr_plan_create(bw_pr_plan **plan, uint64_t vertex_count,
                      uint64_t edge_count, const void *rev_offsets,
                      int32_t offset_bytes, const uint32_t *rev_cols);
int bw_pr_plan_destroy(bw_pr_plan *plan);
int bw_pr_plan_info(const bw_pr_plan *plan, uint32_t *row_tiles,
                    uint32_t *long_rows, uint32_t *long_pieces);
/* Device out-degrees owned by the plan. */
int bw_pr_plan_degrees(const bw_pr_plan *plan, const uint32_t **deg);
/* Launch shape of the row-tile kernel (DROP-mode results are identical for
 * every setting): hot_vertices (0, 8192 or 16384) sources whose
 * contributions each workgroup keeps in LDS -- the most gathered ones when
 * vertices are numbered by decreasing out-degree; rows_grid workgroups
 * looping over the work items (0 = one workgroup per step);
 * tiles_per_workgroup (1, 2 or 4) row tiles a workgroup processes side by
 * side, sharing one hot table (4 only with hot_vertices 0 or 8192). */
int bw_pr_plan_tune(bw_pr_plan *plan, uint32_t hot_vertices,
                    uint32_t rows_grid, uint32_t tiles_per_workgroup);
/* tiles_per_window 1, 2 or 4: the row tiles are processed in windows of
 * up to that many consecutive tiles whose gathers run in source order
 * (built on the first call for a width: the window edges sorted by source
 * id with their CSR positions, 6 extra bytes per edge of device memory;
 * reads the graph's reverse columns given to bw_pr_plan_create, which must
 * still be live).  Every row is still summed in CSR order, so the ranks do
 * not change; rows_grid applies to the window kernel too.  0 returns to
 * the tile kernel.  gathers_per_round (8 or 16): gathers each thread keeps
 * in flight at once (8 needs fewer registers, so more waves fit a CU). */
int bw_pr_plan_sort_windows(bw_pr_plan *plan, int32_t tiles_per_window,
                            int32_t gathers_per_round);
/* rows_log2 13 or 14: switches the plan to the binned iteration
 * (propagation blocking), built on the device (stable radix sorts and scans
 * over the reverse CSR, no copy of the graph to the host) on the first call
 * for a partition size and chunk size from the graph's reverse CSR given to
 * bw_pr_plan_create (which must still be live).  Rows are cut into
 * partitions of 2^rows_log2; each distinct (source, partition) pair holds
 * one contribution per iteration, written once by the source and read once
 * by the partition, whose rows accumulate in LDS in FP64; a partition's
 * edges are split evenly into chunks of at most chunk_tasks (1 to 4096) x
 * 2048 edges, one workgroup and one partial row vector each, a chunk's
 * evenly over the workgroup's 16 waves; lane_edges (8 or 16) edge codes per lane and
 * round.  Same recurrence and modes as the pull kernels
 * (pagerank/vertex_pull_nodiv.cu:3-29, pagerank/pagerank.cu:142-165); the
 * LDS additions run in a run-dependent order, but each partial is an FP64
 * sum rounded once to FP32.  Needs V / 2^rows_log2 <= 2048.  rows_log2 0
 * returns to the pull kernels.  On error the plan keeps the mode and shape
 * it had before the call. */
int bw_pr_plan_bin(bw_pr_plan *plan, int32_t rows_log2, int32_t chunk_tasks,
                   int32_t lane_edges);
/* The same plan built on the host from a copy of the reverse CSR (the
 * sequential builder the device build restates); always rebuilds.  No
 * counterpart in the reference: for checking the device build and for
 * callers that want the host's memory instead of the device's scratch. */
int bw_pr_plan_bin_host(bw_pr_plan *plan, int32_t rows_log2,
                        int32_t chunk_tasks, int32_t lane_edges);
/* Copies one array of the current binned plan to host memory: its size in
 * bytes goes to *bytes; host_dst may be NULL to ask for the size only. */
#define BW_PR_BIN_CODES 0      /* uint16 per edge, partition-major, padded */
#define BW_PR_BIN_TASKS 1      /* uint4 (edge begin, edge end, pair base, partition) */
#define BW_PR_BIN_CHUNKS 2     /* uint4 (task begin, task end, partition, slot) */
#define BW_PR_BIN_CHUNK_BASE 3 /* uint32[K + 1] */
#define BW_PR_BIN_BLOCKS 4     /* uint4 (first vertex, end vertex, run begin, run count) */
#define BW_PR_BIN_RUN_FLAT 5   /* uint32 per run entry */
#define BW_PR_BIN_RUN_BIN 6    /* uint32 per run entry */
#define BW_PR_BIN_PSRC 7       /* uint16 per pair, scatter order */
int bw_pr_plan_bin_array(const bw_pr_plan *plan, int32_t which, void *host_dst,
                         uint64_t capacity_bytes, uint64_t *bytes);
/* Pairs, chunks and scatter blocks of the binned plan (0 when off). */
int bw_pr_plan_bin_info(const bw_pr_plan *plan, uint32_t *pairs,
                        uint32_t *chunks, uint32_t *blocks);

/* Number of doubles of the `acc` workspace of bw_pr_fused_run. */
uint64_t bw_pr_fused_acc_len(int32_t iterations);

/* `iterations` fused launches on cfg->stream.  ranks (device float[V])
 * receives the final ranks; scratch (device float[V]) and acc (device
 * double[bw_pr_fused_acc_len]) are caller-owned workspaces, so the run
 * allocates nothing.  nodiv_quirk != 0 in DROP mode reproduces the
 * reference's undivided first iteration.  resid_host (optional) receives
 * the per-iteration L1 residuals, kernel_ms (optional) the device time of
 * the iteration launches (HIP events on cfg->stream). */
int bw_pr_fused_run(const bw_graph_dev *g, const bw_pr_plan *plan,
                    int32_t iterations, int32_t mode, int32_t nodiv_quirk,
                    float *ranks, float *scratch, double *acc,
                    double *resid_host, float *kernel_ms,
                    const bw_launch *cfg);

/* ------------------------------------------- device-driven BFS (MI355X) */

/* A BFS engine computing the levels of bfs/bfs.cpp:43-118 (depth INT_MAX =
 * unreached, per-level count = newly discovered vertices) with the level
 * loop on the device.  Each level runs one of the native variants below,
 * chosen per level by a policy:
 *   FIXED     always `fixed`;
 *   SCHEDULE  schedule[depth % len] (bw_bfs_engine_set_schedule);
 *   LEARNED   an 8 -> 32 -> 32 -> K MLP evaluated on the device on the 8
 *             frontier features of the level (bw_bfs_engine_set_model);
 *             the role of SwitchImplementation's model lookup,
 *             ImplementationTemplate.hpp:545-908;
 *   HOST      the variant set by bw_bfs_engine_force.
 * Needs both CSR directions with uint32 offsets (an undirected graph may
 * pass the same arrays twice) and V < 2^31. */
typedef struct bw_bfs_engine bw_bfs_engine;

#define BW_BFS_NATIVE_TD 0      /* top-down, queue + edge-balanced LDS scan */
#define BW_BFS_NATIVE_BU 1      /* bottom-up, thread per unvisited vertex   */
#define BW_BFS_NATIVE_BU_WAVE 2 /* bottom-up, wave64 probes for big rows    */
#define BW_BFS_NATIVE_VARIANTS 3
/* Engine variants BW_BFS_ENGINE_REFERENCE + k run the reference's kernel
 * family k (BW_BFS_EDGE_LIST .. BW_BFS_VERTEX_PULL_WARP of the bfs/ kernels): the
 * same work per thread (edge-parallel scans of the stored edge lists,
 * thread-per-vertex push and pull, 32-lane virtual warps over 32-vertex
 * chunks), with the frontier read from the engine's level bitmap and
 * discoveries going to its queue, so any variant can run any level of a
 * device-driven search.  The edge-list families need the matching
 * representations in the bw_graph_dev given to bw_bfs_engine_create. */
#define BW_BFS_ENGINE_REFERENCE 3
#define BW_BFS_ENGINE_VARIANTS 11

#define BW_POLICY_FIXED 0
#define BW_POLICY_SCHEDULE 1
#define BW_POLICY_LEARNED 2
#define BW_POLICY_HOST 3
/* The direction-optimising heuristic of Beamer, Asanovic and Patterson
 * (SC 2012) with alpha = 14, beta = 24.  After a top-down level (or at
 * level 0): bottom-up (BU_WAVE) when the frontier's out-edges m_f exceed
 * m_u / alpha, m_u = E minus the out-degrees of every vertex discovered so
 * far (frontier included); else top-down.  After a bottom-up level:
 * top-down when the frontier is smaller than the previous one and holds
 * fewer than V / beta vertices; else bottom-up. */
#define BW_POLICY_BEAMER 4

/* Features of a level, reduced on the device: [0..2] frontier vertices,
 * frontier out-edges and unvisited vertices as log2 fractions of the graph
 * (log2((1 + x) / V), log2((1 + x) / E), log2((1 + x) / V)); [3] depth;
 * [4..5] log2(1 + x) of the mean and standard deviation of the frontier
 * out-degree; [6] previous variant (-1 at level 0); [7] log2(1 + previous
 * level time in ns); [8] log2 V; [9] out-edges of the vertices not yet
 * discovered, log2((1 + m_u) / E); [10] frontier growth, log2(1 + n_f) -
 * log2(1 + previous frontier size) (that size is 0 at level 0). */
#define BW_BFS_FEATURES 11

/* levels: device int32[V] receiving the depths, or NULL for an
 * engine-owned array (bw_bfs_engine_levels).  max_levels bounds the depth
 * (no fixed 1000-level limit). */
int bw_bfs_engine_create(bw_bfs_engine **engine, const bw_graph_dev *g,
                         int32_t *levels, uint32_t max_levels);
int bw_bfs_engine_destroy(bw_bfs_engine *engine);
int bw_bfs_engine_levels(const bw_bfs_engine *engine, int32_t **levels);

/* Work-division thresholds (the autotuner's knobs): out-degree above which
 * a discovered vertex is a hub whose edges the next top-down level slices
 * over all workgroups (>= 256), and in-degree above which BU_WAVE probes
 * with a whole wave instead of one thread. */
#define BW_BFS_DEFAULT_HUB_DEGREE 4096
#define BW_BFS_DEFAULT_WAVE_DEGREE 8
int bw_bfs_engine_set_tuning(bw_bfs_engine *engine, uint32_t hub_degree,
                             uint32_t wave_degree);

/* Queue sizes (the autotuner's other knobs): discoveries each workgroup
 * stages in LDS before one global append (lds_queue, a multiple of 64 in
 * [256, BW_BFS_MAX_LDS_QUEUE]; the level kernel's dynamic LDS, so a larger
 * queue means fewer workgroups per CU), and the hub queue's capacity per
 * level (hub_cap, a multiple of 256 in [256, BW_BFS_MAX_HUB_CAP]; hubs past
 * it go to the normal queue).  Depths never depend on either. */
#define BW_BFS_DEFAULT_LDS_QUEUE 2048
#define BW_BFS_DEFAULT_HUB_CAP 4096
#define BW_BFS_MAX_LDS_QUEUE 8192
#define BW_BFS_MAX_HUB_CAP 16384
int bw_bfs_engine_set_queues(bw_bfs_engine *engine, uint32_t lds_queue,
                             uint32_t hub_cap);

/* Most roots of one multi-root search. */
#define BW_BFS_MAX_ROOTS 16384

/* Float count of a packed model with `outputs` outputs: mean[11] std[11]
 * W1[32][11] b1[32] W2[32][32] b2[32] W3[K][32] b3[K] (row-major, ReLU
 * hidden layers, outputs = predicted log2 level time per variant; the
 * argmin is taken over the variants the engine can run, i.e. those whose
 * representation its graph holds).  outputs is BW_BFS_NATIVE_VARIANTS or
 * BW_BFS_ENGINE_VARIANTS. */
int32_t bw_bfs_engine_mlp_len(int32_t outputs);
int bw_bfs_engine_set_model(bw_bfs_engine *engine, const float *weights_host,
                            int32_t outputs);
int bw_bfs_engine_set_schedule(bw_bfs_engine *engine,
                               const int32_t *schedule_host, int32_t len);

/* Resets levels and state for a search from `root` and picks the level-0
 * variant. */
int bw_bfs_engine_begin(bw_bfs_engine *engine, uint32_t root, int32_t policy,
                        int32_t fixed, int32_t schedule_len, void *stream);
/* The same from several roots (host array, all at depth 0; a repeated root
 * counts once).  On a disjoint union of graphs with one root each, every
 * component gets exactly its own BFS depths. */
int bw_bfs_engine_begin_multi(bw_bfs_engine *engine, const uint32_t *roots,
                              uint32_t nroots, int32_t policy, int32_t fixed,
                              int32_t schedule_len, void *stream);
/* Enqueues one level (a no-op on the device once the search is done). */
int bw_bfs_engine_level(bw_bfs_engine *engine, int32_t policy, int32_t fixed,
                        int32_t schedule_len, void *stream);
/* Sets the variant of the pending level. */
int bw_bfs_engine_force(bw_bfs_engine *engine, int32_t variant, void *stream);
/* Synchronises `stream` and reports the state and the pending level's
 * features (feat: float[BW_BFS_FEATURES], optional). */
int bw_bfs_engine_status(bw_bfs_engine *engine, uint32_t *done,
                         uint32_t *depth, uint32_t *num_levels,
                         int32_t *choice, float *feat, void *stream);
/* restore == 0 saves the search state, != 0 restores it (for timing every
 * variant on the same level). */
int bw_bfs_engine_snapshot(bw_bfs_engine *engine, int32_t restore,
                           void *stream);
/* Runs the pending level with `variant`; ms = device time of its level
 * kernel (HIP events), count = vertices it discovered. */
int bw_bfs_engine_step_forced(bw_bfs_engine *engine, int32_t variant,
                              float *ms, uint32_t *count, uint32_t *done,
                              void *stream);
/* A whole search: levels enqueued in batches of levels_per_sync (replayed
 * as one captured hipGraph if use_graph; a NULL stream is then replaced by
 * an engine-owned one), the host reading only the done flag between
 * batches.  Optional host outputs, max_levels entries each: per-level
 * counts, device times (ns) and variants.  total_ms = device time from the
 * reset to the end. */
int bw_bfs_engine_run(bw_bfs_engine *engine, uint32_t root, int32_t policy,
                      int32_t fixed, int32_t schedule_len,
                      int32_t levels_per_sync, int32_t use_graph,
                      uint32_t *counts_host, uint64_t *times_ns_host,
                      int32_t *choices_host, uint32_t *num_levels,
                      float *total_ms, void *stream);
/* bw_bfs_engine_run from several roots (see bw_bfs_engine_begin_multi). */
int bw_bfs_engine_run_multi(bw_bfs_engine *engine, const uint32_t *roots,
                            uint32_t nroots, int32_t policy, int32_t fixed,
                            int32_t schedule_len, int32_t levels_per_sync,
                            int32_t use_graph, uint32_t *counts_host,
                            uint64_t *times_ns_host, int32_t *choices_host,
                            uint32_t *num_levels, float *total_ms,
                            void *stream);

/* ------------------------------------------- multi-source BFS (MI355X) */

/* Up to 64 BFS searches on one graph sharing every pass over it (64-bit
 * visit/frontier masks per vertex, bottom-up levels with early exit, the
 * level loop on the device).  Search i computes the depths of
 * bfs/bfs.cpp:43-118 from sources[i].  Needs the forward and reverse CSR
 * with uint32 offsets (offsets / cols for top-down levels, rev_offsets /
 * rev_cols for bottom-up ones). */
typedef struct bw_msbfs bw_msbfs;

#define BW_MSBFS_DEFAULT_WAVE_DEGREE 32
#define BW_MSBFS_DEFAULT_TOPDOWN_DIVISOR 4
#define BW_MSBFS_DEFAULT_HUB_DEGREE 128

int bw_msbfs_create(bw_msbfs **ms, const bw_graph_dev *g, uint32_t max_levels);
int bw_msbfs_destroy(bw_msbfs *ms);
/* wave_degree: in-degree above which a whole wave probes a vertex's
 * in-neighbours (bottom-up levels).  hub_degree (>= 16): in-degree above
 * which a vertex's in-edges are split into chunks of 512, one wave each
 * (bottom-up levels; changing it rebuilds the chunk list).
 * topdown_divisor (0..1024): a level runs top-down when the out-edges of
 * its frontier times the divisor are at most E, bottom-up otherwise; 0
 * keeps every level bottom-up.  Depths do not depend on any of them. */
int bw_msbfs_set_tuning(bw_msbfs *ms, uint32_t wave_degree,
                        uint32_t topdown_divisor, uint32_t hub_degree);
/* depths: device int32[n][V], row i = depths from sources[i] (host array,
 * 1 <= n <= 64).  Levels are launched in batches of levels_per_sync
 * (replayed as a captured hipGraph if use_graph).  Optional host outputs:
 * level_bits (uint64[max_levels], newly reached (search, vertex) pairs per
 * level), total_ms (device time of the whole run). */
int bw_msbfs_run(bw_msbfs *ms, const uint32_t *sources, uint32_t n,
                 int32_t *depths, int32_t levels_per_sync, int32_t use_graph,
                 uint64_t *level_bits, uint32_t *num_levels, float *total_ms,
                 void *stream);

/* ------------------------------------------ the variant predictor (MLP) */

/* FP32 MLP 8 -> 32 -> 32 -> K in the packed layout of
 * bw_bfs_engine_mlp_len, trained and evaluated on the GPU (the role of the
 * model lookup of SwitchImplementation, ImplementationTemplate.hpp:545-908;
 * the reference trains its models offline, benchmark-analysis/model-src).
 * All pointers are device memory.  X is float[n][8] (raw features),
 * Y float[n][K] (targets: log2 level time per variant). */
#define BW_MLP_MAX_BATCH 128

int32_t bw_mlp_params(int32_t outputs);
/* out[n][K] = network(X). */
int bw_mlp_forward(const float *weights, int32_t outputs, const float *X,
                   uint64_t n, float *out, void *stream);
/* Gradient of L = mean over samples and outputs of (out - y)^2 for
 * 1 <= n <= BW_MLP_MAX_BATCH samples: grad[bw_mlp_params(K)] (zero for the
 * mean/std entries), loss[1]. */
int bw_mlp_grad(const float *weights, int32_t outputs, const float *X,
                const float *Y, uint32_t n, float *grad, float *loss,
                void *stream);
/* Minibatch Adam (beta 0.9 / 0.999) for `epochs` passes over the n
 * samples, in place on `weights` (mean/std untouched), in one launch;
 * epoch_loss (optional, float[epochs]) receives each epoch's mean loss. */
int bw_mlp_train(float *weights, int32_t outputs, const float *X,
                 const float *Y, uint32_t n, uint32_t epochs, uint32_t batch,
                 float lr, uint32_t seed, float *epoch_loss, void *stream);

#ifdef __cplusplus
}
#endif
#endif
