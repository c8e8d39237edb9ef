// PageRank kernel library (libpagerankkernel.so): the algorithm of
// pagerank/pagerank.cpp of merijn/Belewitte on the HIP backend.
//
// Registers the 10 reference variants and "switch" under `registerHIP`,
// each a tuple (iteration kernel, consolidate, zeroInitDegrees,
// computeDegrees), with the reference's 100 iterations and output formats.
// Two flags extend it: --iterations N and --dangling drop|uniform (drop is
// the reference's behaviour).  It also registers "vertex-pull-nodiv-fused",
// the MI355X-native fused pull iteration (FusedPageRank below).

#include <algorithm>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "Algorithm.hpp"
#include "belewitte_graph.h"
#include "HIPBackend.hpp"
#include "ImplementationTemplate.hpp"
#include "Timer.hpp"

namespace {

// FP64 accumulators of the run: [0] L1 residual of the reference's `diff`
// (pagerank/pagerank.cu:4-17), [1] dangling mass.
double *accumulators()
{
    static double *acc = [] {
        void *p = nullptr;
        BW_CHK(bw_hip_dev_alloc(&p, 2 * sizeof(double)));
        return static_cast<double *>(p);
    }();
    return acc;
}

int danglingMode = BW_PR_DANGLING_DROP;
const unsigned *currentDegrees = nullptr;
bool predividePending = false;

void resetDiff()
{
    // Two doubles, zeroed with the 32-bit fill kernel.
    BW_CHK(bw_hip_fill_u32(accumulators(), 0u, 4, nullptr));
}

double getDiff()
{
    double v[2];
    BW_CHK(bw_hip_memcpy_d2h(v, accumulators(), sizeof v));
    return v[0];
}

bw_launch config(size_t warp = 0, size_t chunk = 0)
{
    bw_launch l = HIPBackend::get().launchConfig();
    l.warp = static_cast<uint32_t>(warp);
    l.chunk = static_cast<uint32_t>(chunk);
    return l;
}

// Degree kernels (pagerank/pagerank.cu:19-123).
void zeroInitDegreesKernel(size_t count, unsigned *degrees)
{
    auto c = config();
    BW_CHK(bw_pr_zero_degrees(count, degrees, &c));
}

void edgeListComputeDegrees(EdgeList<unsigned> *g, unsigned *degrees)
{
    auto c = config();
    BW_CHK(bw_pr_degrees_edge_list(g->edge_count, g->inEdges, degrees, &c));
}

void reverseEdgeListComputeDegrees(EdgeList<unsigned> *g, unsigned *degrees)
{
    auto c = config();
    BW_CHK(bw_pr_degrees_edge_list(g->edge_count, g->outEdges, degrees, &c));
}

void structEdgeListComputeDegrees(StructEdgeList<unsigned> *g, unsigned *degrees)
{
    auto c = config();
    BW_CHK(bw_pr_degrees_struct_edge_list(g->edge_count, reinterpret_cast<const uint32_t *>(g->edges), 0, degrees, &c));
}

void reverseStructEdgeListComputeDegrees(StructEdgeList<unsigned> *g, unsigned *degrees)
{
    auto c = config();
    BW_CHK(bw_pr_degrees_struct_edge_list(g->edge_count, reinterpret_cast<const uint32_t *>(g->edges), 1, degrees, &c));
}

void reverseCSRComputeDegrees(CSR<unsigned, unsigned> *g, unsigned *degrees)
{
    auto c = config();
    BW_CHK(bw_pr_degrees_rev_csr(g->vertex_count, g->vertices, 4, g->edges, degrees, &c));
}

// Binned layouts of the push forms (bw_pr_scatter_*), one per source kind,
// built on a run's first iteration and released when the next run starts;
// a graph the layout refuses keeps the direct atomic kernels.
struct RunScatter {
    bw_pr_scatter *layout = nullptr;
    const void *key = nullptr;
    bool refused = false;
};
RunScatter runScatter[3];

// The pull plan of the run (long rows one wave each), built on the first
// pull iteration of a run for the reverse CSR it is given.
bw_pr_pull_plan *runPull = nullptr;
const void *runPullKey = nullptr;

void resetScatter()
{
    for (auto &r : runScatter) {
        if (r.layout) BW_CHK(bw_pr_scatter_destroy(r.layout));
        r = RunScatter{};
    }
    if (runPull) BW_CHK(bw_pr_pull_plan_destroy(runPull));
    runPull = nullptr;
    runPullKey = nullptr;
}

bw_pr_pull_plan *pullPlanFor(CSR<unsigned, unsigned> *g)
{
    if (runPullKey != g->vertices) {
        if (runPull) BW_CHK(bw_pr_pull_plan_destroy(runPull));
        runPull = nullptr;
        BW_CHK(bw_pr_pull_plan_create(&runPull, g->vertex_count, g->vertices, 4, nullptr));
        runPullKey = g->vertices;
    }
    return runPull;
}

bw_pr_scatter *scatterFor(int32_t kind, uint64_t V, uint64_t E, const void *a, const void *b)
{
    RunScatter &r = runScatter[kind];
    if (r.key != a) {
        if (r.layout) BW_CHK(bw_pr_scatter_destroy(r.layout));
        r = RunScatter{};
        r.key = a;
        const int st = bw_pr_scatter_create(&r.layout, kind, V, E, a, b, 4, nullptr);
        if (st == BW_HIP_ERR_ARG) r.refused = true;
        else BW_CHK(st);
    }
    return r.refused ? nullptr : r.layout;
}

// Iteration kernels.
void edgeListPageRank(EdgeList<unsigned> *g, unsigned *deg, float *pr, float *np)
{
    if (auto *sc = g->edge_count ? scatterFor(BW_PR_SCATTER_EDGE_LIST, g->vertex_count, g->edge_count, g->inEdges,
                                              g->outEdges)
                                 : nullptr) {
        BW_CHK(bw_pr_scatter_iteration(sc, deg, pr, np, nullptr));
        return;
    }
    auto c = config();
    BW_CHK(bw_pr_edge_list(g->edge_count, g->inEdges, g->outEdges, deg, pr, np, &c));
}

void revEdgeListPageRank(EdgeList<unsigned> *g, unsigned *deg, float *pr, float *np)
{
    auto c = config();
    BW_CHK(bw_pr_rev_edge_list(g->edge_count, g->inEdges, g->outEdges, deg, pr, np, &c));
}

void structEdgeListPageRank(StructEdgeList<unsigned> *g, unsigned *deg, float *pr, float *np)
{
    if (auto *sc = g->edge_count ? scatterFor(BW_PR_SCATTER_STRUCT, g->vertex_count, g->edge_count, g->edges, nullptr)
                                 : nullptr) {
        BW_CHK(bw_pr_scatter_iteration(sc, deg, pr, np, nullptr));
        return;
    }
    auto c = config();
    BW_CHK(bw_pr_struct_edge_list(g->edge_count, reinterpret_cast<const uint32_t *>(g->edges), deg, pr, np, &c));
}

void revStructEdgeListPageRank(StructEdgeList<unsigned> *g, unsigned *deg, float *pr, float *np)
{
    auto c = config();
    BW_CHK(bw_pr_rev_struct_edge_list(g->edge_count, reinterpret_cast<const uint32_t *>(g->edges), deg, pr, np, &c));
}

void vertexPushPageRank(CSR<unsigned, unsigned> *g, unsigned *, float *pr, float *np)
{
    if (auto *sc = g->edge_count ? scatterFor(BW_PR_SCATTER_CSR, g->vertex_count, g->edge_count, g->vertices, g->edges)
                                 : nullptr) {
        BW_CHK(bw_pr_scatter_iteration(sc, nullptr, pr, np, nullptr));
        return;
    }
    auto c = config();
    BW_CHK(bw_pr_vertex_push(g->vertex_count, g->vertices, 4, g->edges, pr, np, &c));
}

void vertexPullPageRank(CSR<unsigned, unsigned> *g, unsigned *deg, float *pr, float *np)
{
    auto c = config();
    BW_CHK(bw_pr_pull_iteration(pullPlanFor(g), BW_PR_VERTEX_PULL, g->edges, deg, pr, np, &c));
}

void predivideOnce(uint64_t V, const unsigned *deg, float *pr)
{
    if (!predividePending) return;
    predividePending = false;
    BW_CHK(bw_pr_predivide(V, deg, pr, nullptr));
}

void vertexPullNoDivPageRank(CSR<unsigned, unsigned> *g, unsigned *deg, float *pr, float *np)
{
    predivideOnce(g->vertex_count, deg, pr);
    auto c = config();
    BW_CHK(bw_pr_pull_iteration(pullPlanFor(g), BW_PR_VERTEX_PULL_NODIV, g->edges, deg, pr, np, &c));
}

void vertexPushWarpPageRank(size_t w, size_t ch, CSR<unsigned, unsigned> *g, unsigned *, float *pr, float *np)
{
    // The warp geometry only shapes the atomic kernel; the binned layout
    // does not depend on it.
    if (auto *sc = g->edge_count ? scatterFor(BW_PR_SCATTER_CSR, g->vertex_count, g->edge_count, g->vertices, g->edges)
                                 : nullptr) {
        BW_CHK(bw_pr_scatter_iteration(sc, nullptr, pr, np, nullptr));
        return;
    }
    auto c = config(w, ch);
    BW_CHK(bw_pr_vertex_push_warp(g->vertex_count, g->vertices, 4, g->edges, pr, np, &c));
}

void vertexPullWarpPageRank(size_t w, size_t ch, CSR<unsigned, unsigned> *g, unsigned *deg, float *pr, float *np)
{
    auto c = config(w, ch);
    BW_CHK(bw_pr_pull_iteration(pullPlanFor(g), BW_PR_VERTEX_PULL_WARP, g->edges, deg, pr, np, &c));
}

void vertexPullWarpNoDivPageRank(size_t w, size_t ch, CSR<unsigned, unsigned> *g, unsigned *deg, float *pr,
                                 float *np)
{
    predivideOnce(g->vertex_count, deg, pr);
    auto c = config(w, ch);
    BW_CHK(bw_pr_pull_iteration(pullPlanFor(g), BW_PR_VERTEX_PULL_WARP_NODIV, g->edges, deg, pr, np, &c));
}

// Consolidate kernels (pagerank/pagerank.cu:125-165).
void consolidateRank(size_t V, unsigned *deg, float *pr, float *np, bool notLast)
{
    auto c = config();
    const unsigned *d = danglingMode == BW_PR_DANGLING_UNIFORM ? currentDegrees : nullptr;
    BW_CHK(bw_pr_consolidate(V, d ? d : deg, pr, np, notLast, 0, danglingMode, accumulators() + 1, accumulators(), &c));
}

void consolidateRankNoDiv(size_t V, unsigned *deg, float *pr, float *np, bool notLast)
{
    auto c = config();
    BW_CHK(bw_pr_consolidate(V, deg, pr, np, notLast, 1, danglingMode, accumulators() + 1, accumulators(), &c));
}

uint32_t mask_N_bits(unsigned n) { return ~((n >= 32) ? 0xFFFFFFFFu : ((1u << n) - 1u)); }

float roundPrecision(float v, uint32_t mask)
{
    uint32_t bits;
    std::memcpy(&bits, &v, sizeof bits);
    bits &= mask;
    std::memcpy(&v, &bits, sizeof v);
    return v;
}

// The `.output` of a PageRank run (pagerank/pagerank.cpp:134-170): ranks
// with 7 significant digits under --validate, otherwise the vertex ids in
// bucketed rank order.
template<typename Ranks>
void writeRanks(std::ofstream &outputFile, const Ranks &ranks, size_t n, bool validate)
{
    const unsigned maskBits = 16, bucketSize = 1000;
    auto oldLocale = outputFile.imbue(std::locale::classic());
    if (validate) {
        outputFile << std::setprecision(std::numeric_limits<float>::digits10 + 1);
        for (size_t i = 0; i < n; i++) outputFile << i << "\t" << ranks[i] << "\n";
    } else {
        const uint32_t mask = mask_N_bits(maskBits);
        std::vector<std::pair<float, size_t>> buckets;
        buckets.reserve(n);
        for (size_t i = 0; i < n; i++) buckets.emplace_back(roundPrecision(ranks[i], mask), i);
        std::stable_sort(buckets.begin(), buckets.end());
        for (size_t i = 0; i < buckets.size(); i += bucketSize) {
            // The last element stays outside the last sort
            // (pagerank/pagerank.cpp:155-156).
            const size_t end = std::min(i + bucketSize, buckets.size() - 1);
            if (end > i)
                std::stable_sort(buckets.begin() + static_cast<long>(i), buckets.begin() + static_cast<long>(end),
                                 [](const auto &a, const auto &b) { return a.second < b.second; });
        }
        for (const auto &p : buckets) outputFile << p.second << "\n";
    }
    outputFile.imbue(oldLocale);
}

}  // namespace

template<typename Platform, typename Vertex, typename Edge, bool switching>
struct PageRank : public ImplementationTemplate<Platform, Vertex, Edge, switching> {
    using Impl = ImplementationTemplate<Platform, Vertex, Edge, switching>;
    using Impl::backend;
    using Impl::isSwitching;
    using Impl::options;
    using Impl::run_count;
    using Impl::setKernelConfig;
    using Impl::validate;
    using Impl::vertex_count;

    template<typename... Args>
    using Kernel = typename Impl::template GraphKernel<Args...>;

    int max_iterations = 100;  // pagerank/pagerank.cpp:55
    std::string dangling = "drop";

    Kernel<unsigned *> zeroInitDegrees;
    Kernel<unsigned *> computeDegrees;
    Kernel<unsigned *, float *, float *> kernel;
    Kernel<unsigned *, float *, float *, bool> consolidate;

    PageRank(Kernel<unsigned *, float *, float *> k, Kernel<unsigned *, float *, float *, bool> c,
             Kernel<unsigned *> zeroInit, Kernel<unsigned *> compute)
      : zeroInitDegrees(zeroInit), computeDegrees(compute), kernel(k), consolidate(c)
    {
        options.add("iterations", "NUM", max_iterations, "Number of PageRank iterations.");
        options.add("dangling", "MODE", dangling,
                    "Dangling mass: drop (the reference's behaviour) or uniform (redistributed).");
    }

    void runImplementation(std::ofstream &outputFile) override
    {
        if (dangling != "drop" && dangling != "uniform") reportError("--dangling must be drop or uniform");
        danglingMode = dangling == "uniform" ? BW_PR_DANGLING_UNIFORM : BW_PR_DANGLING_DROP;

        Timer initResults("initResults", run_count);
        Timer pagerankTime("computation", run_count);
        Timer pagerankStepTime("0:computation", run_count);
        Timer resultTransfer("resultTransfer", run_count);

        auto pageranks = backend.template alloc<float>(vertex_count);
        auto new_pageranks = backend.template alloc<float>(vertex_count);
        // UNIFORM mode needs out-degrees for every variant (the push
        // variants have no degree kernels in the reference).
        auto uniformDegrees = backend.template alloc<unsigned>(
            danglingMode == BW_PR_DANGLING_UNIFORM ? vertex_count : 1);

        for (size_t i = 0; i < run_count; i++) {
            initResults.start();
            std::fill(pageranks.begin(), pageranks.end(), 1.0f / static_cast<float>(vertex_count));
            pageranks.copyHostToDev();
            std::fill(new_pageranks.begin(), new_pageranks.end(), 0.0f);
            new_pageranks.copyHostToDev();
            initResults.stop();

            pagerankTime.start();
            pagerankStepTime.start();
            if constexpr (isSwitching) this->predictInitial();

            const uint64_t degreeBufferSize = (zeroInitDegrees || computeDegrees) ? vertex_count : 0;
            auto degrees = backend.template alloc<unsigned>(std::max<uint64_t>(degreeBufferSize, 1));
            if (zeroInitDegrees) setKernelConfig(zeroInitDegrees);
            zeroInitDegrees(this->loader, degrees);
            if (computeDegrees) setKernelConfig(computeDegrees);
            computeDegrees(this->loader, degrees);

            currentDegrees = nullptr;
            if (danglingMode == BW_PR_DANGLING_UNIFORM) {
                auto &fwd = this->loader.template getGraph<Rep::CSR, Dir::Forward>();
                if (degreeBufferSize) {
                    currentDegrees = degrees.devicePtr();
                } else if (fwd) {
                    BW_CHK(bw_pr_zero_degrees(vertex_count, uniformDegrees.devicePtr(), nullptr));
                    BW_CHK(bw_pr_degrees_csr(vertex_count, fwd->vertices, 4, uniformDegrees.devicePtr(), nullptr));
                    currentDegrees = uniformDegrees.devicePtr();
                } else {
                    reportError("--dangling uniform needs out-degrees this variant does not compute");
                }
            }
            // UNIFORM mode starts the nodiv variants from divided
            // contributions; their launchers apply it on the first call.
            predividePending = danglingMode == BW_PR_DANGLING_UNIFORM;
            // Each run builds its push layouts afresh (inside the
            // computation timer, as the degree kernels are).
            resetScatter();

            int j = 0;
            do {
                j++;
                resetDiff();
                if (danglingMode == BW_PR_DANGLING_UNIFORM)
                    BW_CHK(bw_pr_dangling_mass(vertex_count, currentDegrees, pageranks.devicePtr(), accumulators() + 1,
                                               nullptr));
                setKernelConfig(kernel);
                kernel(this->loader, degrees, pageranks, new_pageranks);
                setKernelConfig(consolidate);
                consolidate(this->loader, degrees, pageranks, new_pageranks, max_iterations > j);
                (void)getDiff();
            } while (j < max_iterations);
            pagerankStepTime.stop();
            pagerankTime.stop();

            resultTransfer.start();
            pageranks.copyDevToHost();
            resultTransfer.stop();
        }

        resetScatter();
        writeRanks(outputFile, pageranks, pageranks.size, validate);
    }
};

// "vertex-pull-nodiv-fused": the MI355X-native fused pull iteration of
// libbelewitte_hip (bw_pr_fused_run), which has no counterpart in the
// reference.  It computes vertex-pull-nodiv (pagerank/vertex_pull_nodiv.cu):
// DROP mode keeps the reference's undivided first iteration, UNIFORM mode
// redistributes the dangling mass.  Same lifecycle, --iterations and
// --dangling flags and output formats as the implementations above.  By
// default the vertices are relabelled by decreasing out-degree on the
// device after the upload (bw_graph_relabel_by_degree; --no-reorder keeps
// the file's order; the output is in the file's order either way) and each iteration is the binned form (propagation blocking):
// partitions of 2^14 rows (--bins 13/14; --bins 0 selects the pull
// kernels), chunks of 128 tasks of 2048 edges (--bin-chunk N), 8 edges per
// lane.  A graph the binned layout cannot hold (more than 2048 partitions,
// or too many pairs or edge codes) falls back to the pull kernels with a
// note on stderr.  --windows 0/1/2/4 (source-sorted windows of that many
// row tiles) and --rows-grid N (persistent row workgroups, default 3 per
// CU) shape the pull kernels only.
// Timer "computation": device time of the iterations.
struct FusedPageRank : public ImplementationBase {
    int max_iterations = 100;
    std::string dangling = "drop";
    bool keepOrder = false;
    int windows = 2;
    int rowsGrid = 0;
    int binRows = 14;
    int binChunk = 128;
    bw_graph *host = nullptr;
    bw_graph_dev dev{};
    std::vector<uint32_t> perm;  // file id -> device id, when relabelled
    std::vector<void *> bufs;

    FusedPageRank()
    {
        options.add("iterations", "NUM", max_iterations, "Number of PageRank iterations.");
        options.add("dangling", "MODE", dangling,
                    "Dangling mass: drop (the reference's behaviour) or uniform (redistributed).");
        options.add("no-reorder", keepOrder, true, "Keep the file's vertex order on the device.");
        options.add("windows", "NUM", windows, "Row tiles per source-sorted window (0: off, 1, 2, 4).");
        options.add("rows-grid", "NUM", rowsGrid, "Persistent row workgroups (0: 3 per CU).");
        options.add("bins", "LOG2", binRows,
                    "Binned iteration with 2^LOG2 rows per partition (13 or 14; 0: the pull kernels).");
        options.add("bin-chunk", "NUM", binChunk, "Tasks of 2048 edges per chunk of the binned iteration.");
    }

    ~FusedPageRank() override { FusedPageRank::freeGraph(); }

    void *upload(const void *src, size_t bytes)
    {
        void *d = nullptr;
        BW_CHK(bw_hip_dev_alloc(&d, std::max<size_t>(bytes, 4)));
        bufs.push_back(d);
        if (bytes && src) BW_CHK(bw_hip_memcpy_h2d(d, src, bytes));
        return d;
    }

    // Frees one buffer of upload() before freeGraph().
    void release(void *p)
    {
        auto it = std::find(bufs.begin(), bufs.end(), p);
        if (it == bufs.end()) return;
        bufs.erase(it);
        BW_CHK(bw_hip_dev_free(p));
    }

    void loadGraph(const std::string file) override
    {
        freeGraph();
        if (bw_graph_load(&host, file.c_str()) != 0) reportError("Failed to load '", file, "': ", bw_graph_last_error());
        const uint64_t V = bw_graph_vertex_count(host), E = bw_graph_edge_count(host);
        std::vector<uint32_t> off(V + 1);
        if (bw_graph_offsets32(host, 1, off.data()) != 0) reportError(bw_graph_last_error());
        dev.vertex_count = V;
        dev.edge_count = E;
        dev.offset_bytes = 4;
        void *roff = upload(off.data(), 4 * (V + 1));
        void *rcols = upload(bw_graph_cols(host, 1), 4 * E);
        if (keepOrder || V == 0) {  // (no vertex: the plan refuses it later)
            perm.clear();
            dev.rev_offsets = roff;
            dev.rev_cols = static_cast<uint32_t *>(rcols);
            return;
        }
        // The relabelling reads both directions (a symmetric graph's two
        // are equal and relabel alike); only the new reverse CSR is kept.
        if (bw_graph_offsets32(host, 0, off.data()) != 0) reportError(bw_graph_last_error());
        void *foff = upload(off.data(), 4 * (V + 1));
        void *fcols = upload(bw_graph_cols(host, 0), 4 * E);
        void *nfoff = upload(nullptr, 4 * (V + 1)), *nfcols = upload(nullptr, 4 * E);
        void *nroff = upload(nullptr, 4 * (V + 1)), *nrcols = upload(nullptr, 4 * E);
        void *dperm = upload(nullptr, 4 * V);
        BW_CHK(bw_graph_relabel_by_degree(V, E, static_cast<const uint32_t *>(foff), static_cast<const uint32_t *>(fcols),
                                          static_cast<const uint32_t *>(roff), static_cast<const uint32_t *>(rcols),
                                          static_cast<uint32_t *>(nfoff), static_cast<uint32_t *>(nfcols),
                                          static_cast<uint32_t *>(nroff), static_cast<uint32_t *>(nrcols),
                                          static_cast<uint32_t *>(dperm), nullptr));
        perm.assign(V, 0);
        if (V) BW_CHK(bw_hip_memcpy_d2h(perm.data(), dperm, 4 * V));
        for (void *p : {roff, rcols, foff, fcols, nfoff, nfcols, dperm}) release(p);
        dev.rev_offsets = nroff;
        dev.rev_cols = static_cast<uint32_t *>(nrcols);
    }

    void freeGraph() override
    {
        for (void *p : bufs) (void)bw_hip_dev_free(p);
        bufs.clear();
        if (host) bw_graph_free(host);
        host = nullptr;
        dev = bw_graph_dev{};
    }

    void runImplementation(std::ofstream &outputFile) override
    {
        if (dangling != "drop" && dangling != "uniform") reportError("--dangling must be drop or uniform");
        checkError(max_iterations >= 1, "--iterations must be at least 1");
        const int32_t mode = dangling == "uniform" ? BW_PR_DANGLING_UNIFORM : BW_PR_DANGLING_DROP;
        const uint64_t V = dev.vertex_count;
        bw_pr_plan *plan = nullptr;
        BW_CHK(bw_pr_plan_create(&plan, V, dev.edge_count, dev.rev_offsets, 4, dev.rev_cols));
        int32_t device = 0;
        bw_hip_props props;
        BW_CHK(bw_hip_get_device_props(device, &props));
        const uint32_t grid = rowsGrid > 0 ? static_cast<uint32_t>(rowsGrid) : 3u * static_cast<uint32_t>(props.compute_units);
        BW_CHK(bw_pr_plan_tune(plan, 0, grid, 1));
        BW_CHK(bw_pr_plan_sort_windows(plan, windows, 16));
        // The binned layout holds at most 2048 partitions: larger graphs
        // take 2^14-row partitions, and beyond that the pull kernels.
        int rows = binRows;
        auto partitions = [V](int r) { return (V + (uint64_t{1} << r) - 1) >> r; };
        if (rows == 13 && partitions(13) > 2048) rows = 14;
        if (rows != 0 && partitions(rows) > 2048) {
            std::cerr << "vertex-pull-nodiv-fused: " << V << " vertices need more than 2048 partitions of 2^"
                      << rows << " rows; using the pull kernels" << std::endl;
            rows = 0;
        }
        int status = bw_pr_plan_bin(plan, rows, binChunk, 8);
        if (rows != 0 && status == BW_HIP_ERR_ARG) {
            // Refused on the host (too many pairs or edge codes for the
            // layout); the plan is unchanged, so the pull kernels run.
            std::cerr << "vertex-pull-nodiv-fused: the binned layout refused the graph (" << bw_hip_last_error()
                      << "); using the pull kernels" << std::endl;
            status = bw_pr_plan_bin(plan, 0, binChunk, 8);
        }
        BW_CHK(status);
        float *ranks = static_cast<float *>(upload(nullptr, 4 * V));
        float *scratch = static_cast<float *>(upload(nullptr, 4 * V));
        double *acc = static_cast<double *>(upload(nullptr, 8 * bw_pr_fused_acc_len(max_iterations)));
        Timer computation("computation", run_count);
        for (size_t i = 0; i < run_count; i++) {
            float ms = 0;
            BW_CHK(bw_pr_fused_run(&dev, plan, max_iterations, mode, mode == BW_PR_DANGLING_DROP ? 1 : 0, ranks,
                                   scratch, acc, nullptr, &ms, nullptr));
            computation.add(static_cast<double>(ms) * 1e6);
        }
        BW_CHK(bw_pr_plan_destroy(plan));
        std::vector<float> dranks(V), out(V);
        BW_CHK(bw_hip_memcpy_d2h(dranks.data(), ranks, 4 * V));
        for (uint64_t v = 0; v < V; v++) out[v] = perm.empty() ? dranks[v] : dranks[perm[v]];
        writeRanks(outputFile, out, V, validate);
    }
};

extern "C" register_algorithm_t registerHIP;
extern "C" void registerHIP(Algorithm &result)
{
    INITIALISE_ALGORITHM(result);
    KernelBuilder<HIPBackend, unsigned, unsigned> make_kernel;

    auto zeroInit = make_kernel(zeroInitDegreesKernel, work_division::vertex, tag_t(Rep::VertexCount));
    auto revCSRDegrees =
        make_kernel(reverseCSRComputeDegrees, work_division::vertex, tag_t(Rep::CSR), tag_t(Dir::Reverse));
    auto consolidate = make_kernel(consolidateRank, work_division::vertex, tag_t(Rep::VertexCount));
    auto consolidateNoDiv = make_kernel(consolidateRankNoDiv, work_division::vertex, tag_t(Rep::VertexCount));
    auto pushMem = [](size_t chunk) { return chunk * sizeof(float) + (1 + chunk) * sizeof(unsigned); };
    auto pullMem = [](size_t chunk) { return (1 + chunk) * sizeof(unsigned); };
    using Deg = decltype(zeroInit);

    KernelMap prMap{std::pair{
        "edge-list",
        std::tuple{make_kernel(edgeListPageRank, work_division::edge, tag_t(Rep::EdgeList)), consolidate, zeroInit,
                   make_kernel(edgeListComputeDegrees, work_division::edge, tag_t(Rep::EdgeList))}}};
    prMap["rev-edge-list"] = {
        make_kernel(revEdgeListPageRank, work_division::edge, tag_t(Rep::EdgeList), tag_t(Dir::Reverse)), consolidate,
        zeroInit,
        make_kernel(reverseEdgeListComputeDegrees, work_division::edge, tag_t(Rep::EdgeList), tag_t(Dir::Reverse))};
    prMap["struct-edge-list"] = {
        make_kernel(structEdgeListPageRank, work_division::edge, tag_t(Rep::StructEdgeList)), consolidate, zeroInit,
        make_kernel(structEdgeListComputeDegrees, work_division::edge, tag_t(Rep::StructEdgeList))};
    prMap["rev-struct-edge-list"] = {
        make_kernel(revStructEdgeListPageRank, work_division::edge, tag_t(Rep::StructEdgeList), tag_t(Dir::Reverse)),
        consolidate, zeroInit,
        make_kernel(reverseStructEdgeListComputeDegrees, work_division::edge, tag_t(Rep::StructEdgeList),
                    tag_t(Dir::Reverse))};
    prMap["vertex-push"] = std::make_tuple(make_kernel(vertexPushPageRank, work_division::vertex, tag_t(Rep::CSR)),
                                           consolidate, Deg(nullptr), Deg(nullptr));
    prMap["vertex-pull"] = {
        make_kernel(vertexPullPageRank, work_division::vertex, tag_t(Rep::CSR), tag_t(Dir::Reverse)), consolidate,
        zeroInit, revCSRDegrees};
    prMap["vertex-push-warp"] = std::make_tuple(
        make_kernel(vertexPushWarpPageRank, work_division::vertex, pushMem, tag_t(Rep::CSR)), consolidate,
        Deg(nullptr), Deg(nullptr));
    prMap["vertex-pull-warp"] = {
        make_kernel(vertexPullWarpPageRank, work_division::vertex, pullMem, tag_t(Rep::CSR), tag_t(Dir::Reverse)),
        consolidate, zeroInit, revCSRDegrees};
    prMap["vertex-pull-nodiv"] = {
        make_kernel(vertexPullNoDivPageRank, work_division::vertex, tag_t(Rep::CSR), tag_t(Dir::Reverse)),
        consolidateNoDiv, zeroInit, revCSRDegrees};
    prMap["vertex-pull-warp-nodiv"] = {
        make_kernel(vertexPullWarpNoDivPageRank, work_division::vertex, pullMem, tag_t(Rep::CSR), tag_t(Dir::Reverse)),
        consolidateNoDiv, zeroInit, revCSRDegrees};

    for (auto &[name, kernel] : prMap) result.addImplementation(name, make_implementation<PageRank>(kernel));
    result.addImplementation("switch", make_switch_implementation<PageRank>(prMap));
    result.addImplementation("vertex-pull-nodiv-fused", std::make_unique<FusedPageRank>());
}
