// BFS kernel library (libbfskernel.so): the algorithm of bfs/bfs.cpp of
// merijn/Belewitte on the HIP backend.
//
// Registers the 32 implementations (8 kernel families x 4 frontier
// reductions) and "switch" under `registerHIP`.  The kernels are launchers
// with the reference kernels' signatures that call the gfx950 kernels of
// libbelewitte_hip through its C ABI.

#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "Algorithm.hpp"
#include "belewitte_graph.h"
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

namespace {

// Reads the "weights" array of a switch model file (the JSON written by
// belewitte_amd.switch.Model.save); *outputs receives its output count:
// one per native variant or one per engine variant.
std::vector<float> readModel(const std::string &path, int32_t *outputs)
{
    std::ifstream in(path);
    checkError(in.good(), "Cannot open model file '", path, "'");
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    checkError(text.find("\"belewitte-bfs-mlp\"") != std::string::npos, "'", path, "' is not a BFS switch model");
    const size_t key = text.find("\"weights\"");
    const size_t open = key == std::string::npos ? key : text.find('[', key);
    checkError(open != std::string::npos, "'", path, "' has no weights");
    std::vector<float> w;
    const char *p = text.c_str() + open + 1;
    while (*p) {
        while (*p == ' ' || *p == ',' || *p == '\n' || *p == '\r' || *p == '\t') p++;
        if (*p == ']') break;
        char *end = nullptr;
        const float x = std::strtof(p, &end);
        checkError(end != p, "malformed weight in '", path, "'");
        w.push_back(x);
        p = end;
    }
    const int32_t n = static_cast<int32_t>(w.size());
    if (n == bw_bfs_engine_mlp_len(BW_BFS_ENGINE_VARIANTS)) *outputs = BW_BFS_ENGINE_VARIANTS;
    else if (n == bw_bfs_engine_mlp_len(BW_BFS_NATIVE_VARIANTS)) *outputs = BW_BFS_NATIVE_VARIANTS;
    else
        reportError("'", path, "' holds ", w.size(), " weights, expected ",
                    bw_bfs_engine_mlp_len(BW_BFS_ENGINE_VARIANTS), " or ", bw_bfs_engine_mlp_len(BW_BFS_NATIVE_VARIANTS));
    return w;
}

// Default model: data/bfs_switch_model.json two directories above this
// library (belewitte_amd/lib/kernels -> belewitte_amd/data).
std::string defaultModelPath()
{
    Dl_info info;
    if (!dladdr(reinterpret_cast<void *>(&defaultModelPath), &info) || !info.dli_fname) return "";
    std::string so = info.dli_fname;
    const size_t slash = so.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
    return dir + "/../../data/bfs_switch_model.json";
}

// Engine variant ids: the native ones, then the reference's families at
// BW_BFS_ENGINE_REFERENCE + their bfs/ kernel id.
int32_t nativeVariant(const std::string &name)
{
    static const char *const names[BW_BFS_ENGINE_VARIANTS] = {
        "td", "bu", "bu-wave", "edge-list", "rev-edge-list", "struct-edge-list", "rev-struct-edge-list",
        "vertex-push", "vertex-pull", "vertex-push-warp", "vertex-pull-warp"};
    for (int32_t i = 0; i < BW_BFS_ENGINE_VARIANTS; i++)
        if (name == names[i]) return i;
    reportError("Unknown BFS engine variant '", name,
                "' (td, bu, bu-wave, or a reference family: edge-list, rev-edge-list, struct-edge-list, "
                "rev-struct-edge-list, vertex-push, vertex-pull, vertex-push-warp, vertex-pull-warp)");
}

bool isEdgeListFamily(int32_t v)
{ return v >= BW_BFS_ENGINE_REFERENCE && v <= BW_BFS_ENGINE_REFERENCE + BW_BFS_REV_STRUCT_EDGE_LIST; }

}  // namespace

// The MI355X-native searches, which have no counterpart in the reference:
// the device-driven engine of libbelewitte_hip (level loop on the device,
// per-level variant chosen on the device).  Same lifecycle, options (-r)
// and output (`i\t<depth>`) as the BFS implementations above; timers
// "computation" (device time of the search) and "<n>:bfsLevel" (device
// time of each level).
struct NativeBFS : public ImplementationBase {
    int32_t policy;
    std::string variant = "td";
    std::string modelPath;
    std::string schedule;
    unsigned root = 0;
    size_t levelsPerSync = 8;
    bw_graph *host = nullptr;
    bw_graph_dev dev{};
    std::vector<void *> bufs;

    explicit NativeBFS(int32_t p) : policy(p)
    {
        options.add('r', "root", "NUM", root, "Starting vertex for BFS.");
        options.add("variant", "NAME", variant,
                    "Fixed variant (td, bu, bu-wave or a reference family), also the learned fallback.");
        options.add("model", "FILE", modelPath, "Switch model (default: the one shipped with the library).");
        options.add("schedule", "LIST", schedule, "Comma-separated per-level variants, repeated.");
        options.add("levels-per-sync", "NUM", levelsPerSync, "Levels launched between done-flag reads.");
    }

    ~NativeBFS() override { NativeBFS::freeGraph(); }

    void *upload(const void *src, size_t bytes)
    {
        void *d = nullptr;
        BW_CHK(bw_hip_dev_alloc(&d, std::max<size_t>(bytes, 4)));
        bufs.push_back(d);
        if (bytes) BW_CHK(bw_hip_memcpy_h2d(d, src, bytes));
        return d;
    }

    void loadGraph(const std::string file) override
    {
        freeGraph();
        if (bw_graph_load(&host, file.c_str()) != 0) reportError("Failed to load '", file, "': ", bw_graph_last_error());
        const uint64_t V = bw_graph_vertex_count(host), E = bw_graph_edge_count(host);
        std::vector<uint32_t> off(V + 1);
        dev.vertex_count = V;
        dev.edge_count = E;
        dev.offset_bytes = 4;
        for (int rev = 0; rev < 2; rev++) {
            if (bw_graph_offsets32(host, rev, off.data()) != 0) reportError(bw_graph_last_error());
            void *o = upload(off.data(), 4 * (V + 1));
            auto *c = static_cast<uint32_t *>(upload(bw_graph_cols(host, rev), 4 * E));
            if (rev) {
                dev.rev_offsets = o;
                dev.rev_cols = c;
            } else {
                dev.offsets = o;
                dev.cols = c;
            }
        }
    }

    void buildEdgeLists()
    {
        const uint64_t V = dev.vertex_count, E = dev.edge_count;
        auto alloc = [&](size_t bytes) {
            void *d = nullptr;
            BW_CHK(bw_hip_dev_alloc(&d, std::max<size_t>(bytes, 8)));
            bufs.push_back(d);
            return static_cast<uint32_t *>(d);
        };
        uint32_t *in = alloc(4 * E), *out = alloc(4 * E), *pairs = alloc(8 * E);
        uint32_t *rin = alloc(4 * E), *rout = alloc(4 * E), *rpairs = alloc(8 * E);
        BW_CHK(bw_build_edge_lists(V, E, dev.offsets, 4, dev.cols, in, out, pairs, nullptr));
        BW_CHK(bw_build_edge_lists(V, E, dev.rev_offsets, 4, dev.rev_cols, rin, rout, rpairs, nullptr));
        BW_CHK(bw_hip_device_sync());
        dev.el_in = in;
        dev.el_out = out;
        dev.sel = pairs;
        dev.rel_in = rin;
        dev.rel_out = rout;
        dev.rsel = rpairs;
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
        const uint64_t V = dev.vertex_count;
        if (root >= V) return;
        const uint32_t maxLevels = static_cast<uint32_t>(std::min<uint64_t>(V + 1, 1u << 22));
        const int32_t fixed = nativeVariant(variant);
        std::vector<float> w;
        int32_t outputs = 0;
        std::vector<int32_t> sched;
        bool lists = isEdgeListFamily(fixed);
        if (policy == BW_POLICY_LEARNED) {
            w = readModel(modelPath.empty() ? defaultModelPath() : modelPath, &outputs);
            lists = lists || outputs == BW_BFS_ENGINE_VARIANTS;
        } else if (policy == BW_POLICY_SCHEDULE) {
            size_t start = 0;
            while (start <= schedule.size()) {
                const size_t comma = schedule.find(',', start);
                sched.push_back(nativeVariant(schedule.substr(start, comma - start)));
                lists = lists || isEdgeListFamily(sched.back());
                if (comma == std::string::npos) break;
                start = comma + 1;
            }
        }
        // The edge-list families read edge lists, built on the device from
        // the CSR (once per graph, outside the timers).
        if (lists && !dev.el_in) buildEdgeLists();
        bw_bfs_engine *e = nullptr;
        BW_CHK(bw_bfs_engine_create(&e, &dev, nullptr, maxLevels));
        int32_t slen = 1;
        if (policy == BW_POLICY_LEARNED) {
            BW_CHK(bw_bfs_engine_set_model(e, w.data(), outputs));
        } else if (policy == BW_POLICY_SCHEDULE) {
            BW_CHK(bw_bfs_engine_set_schedule(e, sched.data(), static_cast<int32_t>(sched.size())));
            slen = static_cast<int32_t>(sched.size());
        }
        Timer computation("computation", run_count);
        std::vector<std::unique_ptr<Timer>> levelTimers;
        std::vector<uint32_t> counts(maxLevels);
        std::vector<uint64_t> times(maxLevels);
        uint32_t levels = 0;
        for (size_t i = 0; i < run_count; i++) {
            float ms = 0;
            BW_CHK(bw_bfs_engine_run(e, root, policy, fixed, slen, static_cast<int32_t>(levelsPerSync), 1,
                                     counts.data(), times.data(), nullptr, &levels, &ms, nullptr));
            computation.add(static_cast<double>(ms) * 1e6);
            for (uint32_t d = 0; d < levels; d++) {
                if (d >= levelTimers.size())
                    levelTimers.push_back(std::make_unique<Timer>(std::to_string(d) + ":bfsLevel", run_count));
                levelTimers[d]->add(static_cast<double>(times[d]));
            }
        }
        int32_t *devLevels = nullptr;
        BW_CHK(bw_bfs_engine_levels(e, &devLevels));
        std::vector<int32_t> depth(V);
        BW_CHK(bw_hip_memcpy_d2h(depth.data(), devLevels, 4 * V));
        BW_CHK(bw_bfs_engine_destroy(e));
        for (uint64_t v = 0; v < V; v++) outputFile << v << "\t" << depth[v] << "\n";
    }
};

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
    result.addImplementation("native", std::make_unique<NativeBFS>(BW_POLICY_FIXED));
    result.addImplementation("native-schedule", std::make_unique<NativeBFS>(BW_POLICY_SCHEDULE));
    result.addImplementation("native-learned", std::make_unique<NativeBFS>(BW_POLICY_LEARNED));
    result.addImplementation("native-beamer", std::make_unique<NativeBFS>(BW_POLICY_BEAMER));
}
