// Multi-source BFS: up to 64 searches on one graph in one level loop.
//
// Each search s computes exactly the depths of the reference's BFS from its
// source (bfs/bfs.cpp:43-118 of merijn/Belewitte: INT_MAX for unreached).
// The searches share every pass over the graph: vertex v holds a 64-bit
// `visit` mask (bit s = reached by search s) and a `frontier` mask (bit s =
// at the current depth of search s).  A level is one bottom-up pass: every
// vertex not yet reached by all searches ORs the frontier masks of its
// in-neighbours, stopping as soon as that covers every search still missing
// it; the new bits give depth + 1 in those searches.  One thread probes a
// vertex with few in-edges, a whole wave (64 in-edges per step) one above
// the tunable wave threshold.  The level loop runs on the device (a done
// flag read by every kernel), launched in hipGraph batches.

#include "common.hpp"

#include <algorithm>
#include <cstring>

namespace bw {
namespace {

constexpr int kBlock = 256;

struct MsState {
    uint32_t depth;
    uint32_t done;
    uint32_t num_levels;
    uint32_t pad;
};

struct MsArgs {
    uint64_t V;
    const uint32_t *roff;
    const uint32_t *rcol;
    uint64_t *visit;
    uint64_t *front[2];  // level d reads front[d & 1], writes front[(d & 1) ^ 1]
    int32_t *depths;     // [nsrc][V]
    uint64_t all;        // mask of the nsrc searches
    unsigned long long *level_bits;  // per level: newly set bits
    MsState *st;
    uint32_t max_levels;
    uint32_t wave_degree;
};

__device__ __forceinline__ void record(const MsArgs &a, uint32_t v, uint64_t nb, int32_t nd)
{
    while (nb) {
        const int s = __ffsll(static_cast<unsigned long long>(nb)) - 1;
        nb &= nb - 1;
        a.depths[static_cast<uint64_t>(s) * a.V + v] = nd;
    }
}

__global__ __launch_bounds__(kBlock) void msbfs_level(MsArgs a)
{
    __shared__ unsigned long long part[kBlock / kWave];
    const MsState &st = *a.st;
    if (st.done) return;
    const uint32_t d = st.depth;
    const int32_t nd = static_cast<int32_t>(d) + 1;
    // Selects rather than indexes, so the argument array stays in registers.
    const uint64_t *fin = (d & 1) ? a.front[1] : a.front[0];
    uint64_t *fout = (d & 1) ? a.front[0] : a.front[1];
    const int lane = __lane_id();
    unsigned long long found = 0;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlock; base < a.V;
         base += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t v64 = base + threadIdx.x;
        const uint32_t v = static_cast<uint32_t>(v64);
        const bool live = v64 < a.V;
        const uint64_t vis = live ? a.visit[v] : a.all;
        const bool open = vis != a.all;
        uint32_t s = 0, e = 0;
        if (open) {
            s = a.roff[v];
            e = a.roff[v + 1];
        }
        // The lane probes the first wave_degree in-edges itself; a row still
        // missing searches after that is finished by the whole wave.
        const uint32_t lim = e - s > a.wave_degree ? s + a.wave_degree : e;
        uint64_t acc = 0;
        for (uint32_t j = s; j < lim; j++) {
            acc |= fin[a.rcol[j]];
            if ((acc | vis) == a.all) break;
        }
        uint64_t m = __ballot(open && lim < e && (acc | vis) != a.all);
        while (m) {
            const int l = __ffsll(static_cast<unsigned long long>(m)) - 1;
            m &= m - 1;
            const uint32_t vs = __shfl(lim, l, kWave), ve = __shfl(e, l, kWave);
            const uint64_t lvis = __shfl(vis | acc, l, kWave);
            uint64_t wacc = 0;
            for (uint32_t k0 = vs; k0 < ve; k0 += kWave) {
                const uint32_t j = k0 + lane;
                uint64_t x = j < ve ? fin[a.rcol[j]] : 0;
                for (int o = kWave / 2; o > 0; o >>= 1) x |= __shfl_xor(x, o, kWave);
                wacc |= x;
                if ((wacc | lvis) == a.all) break;
            }
            if (lane == l) acc |= wacc;
        }
        const uint64_t nb = acc & ~vis;
        if (live) {
            fout[v] = nb;
            if (nb) {
                a.visit[v] = vis | nb;
                record(a, v, nb, nd);
                found += __popcll(nb);
            }
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) found += __shfl_xor(found, o, kWave);
    if (lane == 0) part[threadIdx.x / kWave] = found;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / kWave; w++) t += part[w];
        if (t) atomicAdd(&a.level_bits[d], t);
    }
}

__global__ void msbfs_epilogue(MsArgs a)
{
    MsState &st = *a.st;
    if (st.done) return;
    const uint32_t d = st.depth;
    st.num_levels = d + 1;
    if (a.level_bits[d] == 0 || d + 1 >= a.max_levels) st.done = 1;
    else st.depth = d + 1;
}

// Depth 0: source i gets bit i (a repeated source vertex gets several).
__global__ void msbfs_init(MsArgs a, const uint32_t *src, uint32_t n)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t v = src[i];
        a.visit[v] |= 1ull << i;
        a.front[0][v] |= 1ull << i;
        a.depths[static_cast<uint64_t>(i) * a.V + v] = 0;
    }
    a.st->depth = 0;
    a.st->done = 0;
    a.st->num_levels = 0;
}

}  // namespace
}  // namespace bw

using namespace bw;

struct bw_msbfs {
    uint64_t V = 0;
    const uint32_t *roff = nullptr, *rcol = nullptr;
    uint64_t *visit = nullptr, *front[2] = {nullptr, nullptr};
    unsigned long long *level_bits = nullptr;
    MsState *st = nullptr;
    uint32_t *src = nullptr;
    uint32_t max_levels = 0;
    uint32_t wave_degree = BW_MSBFS_DEFAULT_WAVE_DEGREE;
    uint32_t grid = 0;
    hipStream_t own_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t *exec_depths = nullptr;
    uint64_t exec_all = 0;
    int32_t exec_lps = 0;
    uint32_t exec_wave = 0;
    void *exec_stream = nullptr;

    ~bw_msbfs()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        void *bufs[] = {visit, front[0], front[1], level_bits, st, src};
        for (void *p : bufs)
            if (p) (void)hipFree(p);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

namespace {

MsArgs make_args(const bw_msbfs *m, int32_t *depths, uint32_t n)
{
    MsArgs a;
    a.V = m->V;
    a.roff = m->roff;
    a.rcol = m->rcol;
    a.visit = m->visit;
    a.front[0] = m->front[0];
    a.front[1] = m->front[1];
    a.depths = depths;
    a.all = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
    a.level_bits = m->level_bits;
    a.st = m->st;
    a.max_levels = m->max_levels;
    a.wave_degree = m->wave_degree;
    return a;
}

int launch_levels(const bw_msbfs *m, const MsArgs &a, int32_t count, hipStream_t s)
{
    for (int32_t i = 0; i < count; i++) {
        hipLaunchKernelGGL(msbfs_level, dim3(m->grid), dim3(kBlock), 0, s, a);
        BW_LAUNCH_CHECK();
        hipLaunchKernelGGL(msbfs_epilogue, dim3(1), dim3(1), 0, s, a);
        BW_LAUNCH_CHECK();
    }
    return BW_HIP_OK;
}

}  // namespace

extern "C" int bw_msbfs_create(bw_msbfs **out, const bw_graph_dev *g, uint32_t max_levels)
{
    if (!out || !g || max_levels == 0) return fail(BW_HIP_ERR_ARG, "bw_msbfs_create: invalid argument");
    *out = nullptr;
    if (g->offset_bytes != 4 || !g->rev_offsets || (g->edge_count && !g->rev_cols))
        return fail(BW_HIP_ERR_ARG, "multi-source BFS needs the reverse CSR with uint32 offsets");
    if (g->vertex_count == 0 || g->vertex_count >= 0x7FFFFFFFull)
        return fail(BW_HIP_ERR_ARG, "multi-source BFS needs 0 < V < 2^31");
    auto m = new bw_msbfs();
    m->V = g->vertex_count;
    m->roff = static_cast<const uint32_t *>(g->rev_offsets);
    m->rcol = g->rev_cols;
    m->max_levels = max_levels;
    m->grid = static_cast<uint32_t>(device_cus()) * 8;
    bool ok = hipMalloc(&m->visit, 8 * m->V) == hipSuccess && hipMalloc(&m->front[0], 8 * m->V) == hipSuccess &&
              hipMalloc(&m->front[1], 8 * m->V) == hipSuccess &&
              hipMalloc(&m->level_bits, 8ull * max_levels) == hipSuccess &&
              hipMalloc(&m->st, sizeof(MsState)) == hipSuccess && hipMalloc(&m->src, 4 * 64) == hipSuccess;
    if (!ok) {
        delete m;
        return fail(BW_HIP_ERR_NOMEM, "device allocation of the multi-source BFS failed");
    }
    *out = m;
    return BW_HIP_OK;
}

extern "C" int bw_msbfs_destroy(bw_msbfs *m)
{
    delete m;
    return BW_HIP_OK;
}

extern "C" int bw_msbfs_set_tuning(bw_msbfs *m, uint32_t wave_degree)
{
    if (!m) return fail(BW_HIP_ERR_ARG, "NULL handle");
    m->wave_degree = wave_degree;
    return BW_HIP_OK;
}

extern "C" int bw_msbfs_run(bw_msbfs *m, const uint32_t *sources, uint32_t n, int32_t *depths,
                            int32_t levels_per_sync, int32_t use_graph, uint64_t *level_bits_host,
                            uint32_t *num_levels, float *total_ms, void *stream)
{
    if (!m || !sources || !depths || !num_levels || levels_per_sync < 1)
        return fail(BW_HIP_ERR_ARG, "bw_msbfs_run: invalid argument");
    if (n < 1 || n > 64) return fail(BW_HIP_ERR_ARG, "multi-source BFS takes 1..64 sources");
    for (uint32_t i = 0; i < n; i++)
        if (sources[i] >= m->V) return fail(BW_HIP_ERR_ARG, "source " + std::to_string(sources[i]) + " out of range");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (use_graph && !s) {
        if (!m->own_stream) BW_HIP_CHECK(hipStreamCreateWithFlags(&m->own_stream, hipStreamNonBlocking));
        s = m->own_stream;
    }
    const MsArgs a = make_args(m, depths, n);
    if (use_graph && (!m->exec || m->exec_depths != depths || m->exec_all != a.all ||
                      m->exec_lps != levels_per_sync || m->exec_wave != m->wave_degree || m->exec_stream != s)) {
        if (m->exec) (void)hipGraphExecDestroy(m->exec);
        m->exec = nullptr;
        hipGraph_t graph = nullptr;
        BW_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int cst = launch_levels(m, a, levels_per_sync, s);
        const hipError_t ce = hipStreamEndCapture(s, &graph);
        if (cst != BW_HIP_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return cst;
        }
        if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture", __FILE__, __LINE__);
        const hipError_t ie = hipGraphInstantiate(&m->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            m->exec = nullptr;
            return hip_fail(ie, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        m->exec_depths = depths;
        m->exec_all = a.all;
        m->exec_lps = levels_per_sync;
        m->exec_wave = m->wave_degree;
        m->exec_stream = s;
    }
    hipEvent_t ev0, ev1;
    BW_HIP_CHECK(hipEventCreate(&ev0));
    BW_HIP_CHECK(hipEventCreate(&ev1));
    int st = BW_HIP_OK;
    auto step = [&](int rc) {
        if (st == BW_HIP_OK && rc != BW_HIP_OK) st = rc;
    };
    if (hipEventRecord(ev0, s) != hipSuccess) st = fail(BW_HIP_ERR_RUNTIME, "event record failed");
    step(bw_hip_fill_u32(depths, static_cast<uint32_t>(kUnvisited), static_cast<uint64_t>(n) * m->V, s));
    step(bw_hip_fill_u32(m->visit, 0u, 2 * m->V, s));
    step(bw_hip_fill_u32(m->front[0], 0u, 2 * m->V, s));
    step(bw_hip_fill_u32(m->level_bits, 0u, 2ull * m->max_levels, s));
    if (st == BW_HIP_OK && hipMemcpyAsync(m->src, sources, 4 * n, hipMemcpyHostToDevice, s) != hipSuccess)
        st = fail(BW_HIP_ERR_RUNTIME, "copying the sources failed");
    if (st == BW_HIP_OK) {
        hipLaunchKernelGGL(msbfs_init, dim3(1), dim3(64), 0, s, a, m->src, n);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) st = hip_fail(le, "msbfs_init", __FILE__, __LINE__);
    }
    uint32_t done = 0, batches = 0;
    const uint32_t max_batches = m->max_levels / static_cast<uint32_t>(levels_per_sync) + 2;
    while (st == BW_HIP_OK && !done) {
        if (batches++ > max_batches) {
            st = fail(BW_HIP_ERR_LIMIT, "multi-source BFS exceeded its level bound");
            break;
        }
        if (use_graph) {
            const hipError_t le = hipGraphLaunch(m->exec, s);
            if (le != hipSuccess) st = hip_fail(le, "hipGraphLaunch", __FILE__, __LINE__);
        } else {
            st = launch_levels(m, a, levels_per_sync, s);
        }
        if (st != BW_HIP_OK) break;
        if (hipMemcpyAsync(&done, &m->st->done, sizeof done, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            st = fail(BW_HIP_ERR_RUNTIME, "reading the multi-source BFS state failed");
    }
    if (st == BW_HIP_OK) {
        if (hipEventRecord(ev1, s) != hipSuccess || hipEventSynchronize(ev1) != hipSuccess)
            st = fail(BW_HIP_ERR_RUNTIME, "event failed");
        else if (total_ms)
            (void)hipEventElapsedTime(total_ms, ev0, ev1);
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (st != BW_HIP_OK) return st;
    MsState hs;
    BW_HIP_CHECK(hipMemcpy(&hs, m->st, sizeof hs, hipMemcpyDeviceToHost));
    *num_levels = hs.num_levels;
    const size_t k = std::min(hs.num_levels, m->max_levels);
    if (k) {
        unsigned long long last = 0;
        BW_HIP_CHECK(hipMemcpy(&last, m->level_bits + k - 1, sizeof last, hipMemcpyDeviceToHost));
        if (last != 0) return fail(BW_HIP_ERR_LIMIT, "multi-source BFS reached max_levels with work left");
    }
    if (level_bits_host && k) BW_HIP_CHECK(hipMemcpy(level_bits_host, m->level_bits, 8 * k, hipMemcpyDeviceToHost));
    return BW_HIP_OK;
}
