// BFS kernel library (libbfskernel.so): the algorithm of bfs/bfs.cpp of
// merijn/Belewitte on the HIP backend.
//
// Registers the 32 implementations (8 kernel families x 4 frontier
// reductions) and "switch" under `registerHIP`.  The kernels are launchers
// with the reference kernels' signatures that call the gfx950 kernels of
// libbelewitte_hip through its C ABI.

#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "Algorithm.hpp"
#include "HIPBackend.hpp"
#include "ImplementationTemplate.hpp"
#include "Timer.hpp"

namespace {

// The frontier counter of bfs/bfs.cu:5-18 as a device buffer.
unsigned *frontierCounter()
{
    static unsigned *counter = [] {
        void *p = nullptr;
        BW_CHK(bw_hip_dev_alloc(&p, sizeof(unsigned)));
        return static_cast<unsigned *>(p);
    }();
    return counter;
}

void resetFrontier() { BW_CHK(bw_hip_fill_u32(frontierCounter(), 0u, 1, nullptr)); }

unsigned getFrontier()
{
    unsigned v = 0;
    BW_CHK(bw_hip_memcpy_d2h(&v, frontierCounter(), sizeof v));
    return v;
}

bw_launch config(size_t warp = 0, size_t chunk = 0)
{
    bw_launch l = HIPBackend::get().launchConfig();
    l.warp = static_cast<uint32_t>(warp);
    l.chunk = static_cast<uint32_t>(chunk);
    return l;
}

enum bfs_variant { normal = BW_RED_NORMAL, bulk = BW_RED_BULK, warpreduce = BW_RED_WAVE, blockreduce = BW_RED_BLOCK };

template<bfs_variant V>
struct Reduction;
template<> struct Reduction<normal> { static constexpr const char *suffix = ""; };
template<> struct Reduction<bulk> { static constexpr const char *suffix = "-bulk"; };
template<> struct Reduction<warpreduce> { static constexpr const char *suffix = "-warpreduce"; };
template<> struct Reduction<blockreduce> { static constexpr const char *suffix = "-blockreduce"; };

// Launchers: same parameter lists as the reference kernels.
template<bfs_variant R>
void edgeListBfs(EdgeList<unsigned> *g, int *levels, int depth)
{
    auto c = config();
    BW_CHK(bw_bfs_edge_list(g->edge_count, g->inEdges, g->outEdges, levels, depth, frontierCounter(), R, &c));
}

template<bfs_variant R>
void revEdgeListBfs(EdgeList<unsigned> *g, int *levels, int depth)
{
    auto c = config();
    BW_CHK(bw_bfs_rev_edge_list(g->edge_count, g->inEdges, g->outEdges, levels, depth, frontierCounter(), R, &c));
}

template<bfs_variant R>
void structEdgeListBfs(StructEdgeList<unsigned> *g, int *levels, int depth)
{
    auto c = config();
    BW_CHK(bw_bfs_struct_edge_list(g->edge_count, reinterpret_cast<const uint32_t *>(g->edges), levels, depth,
                                   frontierCounter(), R, &c));
}

template<bfs_variant R>
void revStructEdgeListBfs(StructEdgeList<unsigned> *g, int *levels, int depth)
{
    auto c = config();
    BW_CHK(bw_bfs_rev_struct_edge_list(g->edge_count, reinterpret_cast<const uint32_t *>(g->edges), levels, depth,
                                       frontierCounter(), R, &c));
}

template<bfs_variant R>
void vertexPushBfs(CSR<unsigned, unsigned> *g, int *levels, int depth)
{
    auto c = config();
    BW_CHK(bw_bfs_vertex_push(g->vertex_count, g->vertices, 4, g->edges, levels, depth, frontierCounter(), R, &c));
}

template<bfs_variant R>
void vertexPullBfs(CSR<unsigned, unsigned> *g, int *levels, int depth)
{
    auto c = config();
    BW_CHK(bw_bfs_vertex_pull(g->vertex_count, g->vertices, 4, g->edges, levels, depth, frontierCounter(), R, &c));
}

template<bfs_variant R>
void vertexPushWarpBfs(size_t warp, size_t chunk, CSR<unsigned, unsigned> *g, int *levels, int depth)
{
    auto c = config(warp, chunk);
    BW_CHK(bw_bfs_vertex_push_warp(g->vertex_count, g->vertices, 4, g->edges, levels, depth, frontierCounter(), R,
                                   &c));
}

template<bfs_variant R>
void vertexPullWarpBfs(size_t warp, size_t chunk, CSR<unsigned, unsigned> *g, int *levels, int depth)
{
    auto c = config(warp, chunk);
    BW_CHK(bw_bfs_vertex_pull_warp(g->vertex_count, g->vertices, 4, g->edges, levels, depth, frontierCounter(), R,
                                   &c));
}

}  // namespace

template<typename Platform, typename Vertex, typename Edge, bool switching>
struct BFS : public ImplementationTemplate<Platform, Vertex, Edge, switching> {
    using Impl = ImplementationTemplate<Platform, Vertex, Edge, switching>;
    using Impl::backend;
    using Impl::isSwitching;
    using Impl::loader;
    using Impl::options;
    using Impl::run_count;
    using Impl::setKernelConfig;
    using Impl::vertex_count;

    template<typename... Args>
    using Kernel = typename Impl::template GraphKernel<Args...>;

    unsigned root;
    prop_ref absFrontier, relFrontier, absVisited, relVisited;
    Kernel<int *, int> kernel;

    explicit BFS(Kernel<int *, int> k)
      : root(0), absFrontier("frontier abs", *this), relFrontier("frontier rel", *this),
        absVisited("visited abs", *this), relVisited("visited rel", *this), kernel(k)
    { options.add('r', "root", "NUM", root, "Starting vertex for BFS."); }

    void setProps(unsigned frontier)
    {
        absVisited = absVisited.get() + absFrontier.get();
        relVisited = absVisited.get() / static_cast<double>(vertex_count);
        absFrontier = frontier;
        relFrontier = absFrontier.get() / static_cast<double>(vertex_count);
    }

    void runImplementation(std::ofstream &outputFile) override
    {
        if (root >= vertex_count) return;

        Timer initResults("initResults", run_count);
        Timer bfs("computation", run_count);
        // One timer per level, created on demand (no fixed level cap).
        std::vector<std::unique_ptr<Timer>> levelTimers;
        Timer resultTransfer("resultTransfer", run_count);

        auto results = backend.template alloc<int>(vertex_count);
        for (size_t i = 0; i < run_count; i++) {
            initResults.start();
            std::fill(results.begin(), results.end(), std::numeric_limits<int>::max());
            results[root] = 0;
            results.copyHostToDev();
            initResults.stop();

            bfs.start();
            unsigned frontier = 1;
            int curr = 0;
            if constexpr (isSwitching) {
                absFrontier = 1;
                relFrontier = 1.0 / static_cast<double>(vertex_count);
                absVisited = 0;
                relVisited = 0;
                this->predictInitial();
            }
            setKernelConfig(kernel);
            do {
                if (static_cast<size_t>(curr) >= levelTimers.size())
                    levelTimers.push_back(std::make_unique<Timer>(std::to_string(curr) + ":bfsLevel", run_count));
                auto &levelTimer = *levelTimers[static_cast<size_t>(curr)];
                resetFrontier();
                levelTimer.start();
                kernel(loader, results, curr++);
                frontier = getFrontier();
                levelTimer.stop();
                if constexpr (isSwitching) {
                    if (frontier) {
                        setProps(frontier);
                        if (this->predict()) setKernelConfig(kernel);
                    }
                }
            } while (frontier);
            bfs.stop();

            resultTransfer.start();
            results.copyDevToHost();
            resultTransfer.stop();
        }
        for (size_t i = 0; i < results.size; i++) outputFile << i << "\t" << results[i] << "\n";
    }
};

template<bfs_variant R>
static auto insertVariant()
{
    KernelBuilder<HIPBackend, unsigned, unsigned> make_kernel;
    const std::string s = Reduction<R>::suffix;
    auto warpMem = [](size_t chunk) { return chunk * sizeof(int) + (chunk + 1) * sizeof(unsigned); };

    KernelMap kernelMap{std::pair{std::string("edge-list") + s,
                                  std::tuple{make_kernel(edgeListBfs<R>, work_division::edge, tag_t(Rep::EdgeList))}}};
    kernelMap["rev-edge-list" + s] = {
        make_kernel(revEdgeListBfs<R>, work_division::edge, tag_t(Rep::EdgeList), tag_t(Dir::Reverse))};
    kernelMap["struct-edge-list" + s] = {
        make_kernel(structEdgeListBfs<R>, work_division::edge, tag_t(Rep::StructEdgeList))};
    kernelMap["rev-struct-edge-list" + s] = {
        make_kernel(revStructEdgeListBfs<R>, work_division::edge, tag_t(Rep::StructEdgeList), tag_t(Dir::Reverse))};
    kernelMap["vertex-push" + s] = {make_kernel(vertexPushBfs<R>, work_division::vertex, tag_t(Rep::CSR))};
    kernelMap["vertex-pull" + s] = {
        make_kernel(vertexPullBfs<R>, work_division::vertex, tag_t(Rep::CSR), tag_t(Dir::Reverse))};
    kernelMap["vertex-push-warp" + s] = {
        make_kernel(vertexPushWarpBfs<R>, work_division::vertex, warpMem, tag_t(Rep::CSR))};
    kernelMap["vertex-pull-warp" + s] = {
        make_kernel(vertexPullWarpBfs<R>, work_division::vertex, warpMem, tag_t(Rep::CSR), tag_t(Dir::Reverse))};
    return kernelMap;
}

extern "C" register_algorithm_t registerHIP;
extern "C" void registerHIP(Algorithm &result)
{
    INITIALISE_ALGORITHM(result);
    auto kernelMap = insertVariant<normal>();
    kernelMap += insertVariant<bulk>();
    kernelMap += insertVariant<warpreduce>();
    kernelMap += insertVariant<blockreduce>();
    for (auto &[name, kernel] : kernelMap) result.addImplementation(name, make_implementation<BFS>(kernel));
    result.addImplementation("switch", make_switch_implementation<BFS>(kernelMap));
}
