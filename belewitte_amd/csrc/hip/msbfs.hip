// Multi-source BFS: up to 64 searches on one graph in one level loop.
//
// Each search s computes exactly the depths of the reference's BFS from its
// source (bfs/bfs.cpp:43-118 of merijn/Belewitte: INT_MAX for unreached).
// The searches share every pass over the graph: vertex v holds a 64-bit
// `visit` mask (bit s = reached by search s) and a `frontier` mask (bit s =
// at the current depth of search s).  A level runs in one of two
// directions, chosen on the device from the frontier's out-edge count:
//  - bottom-up: every vertex not yet reached by all searches ORs the
//    frontier masks of its in-neighbours, stopping as soon as that covers
//    every search still missing it.  One thread probes a vertex with few
//    in-edges, a whole wave (256 in-edges per round trip) one above the
//    tunable wave threshold; a vertex above the tunable hub degree is cut
//    into chunks of in-edges, one wave each, that OR into next[] with
//    atomics (msbfs_bu_hubs);
//  - top-down: the frontier vertices are queued with their out-edge ranges
//    reserved in one counter, then every frontier edge u -> w ORs u's mask
//    into next[w] (edge-parallel, so a hub's edges spread over the grid).
//    OR is order-independent, so the result is the same bit for bit.
// Either way the new bits give depth + 1 in those searches.  The level loop
// runs on the device (a done flag read by every kernel), launched in
// hipGraph batches.

#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace bw {
namespace {

constexpr int kBlock = 256;
constexpr int kUnroll = 4;       // bottom-up wave path: chunks of 64 in-edges per round trip
constexpr uint32_t kSlots = 8;   // top-down expand: consecutive edge slots per thread
// Bottom-up: vertices with more in-edges than the tunable hub degree are
// split into chunks of kHubChunk in-edges, one wave each (msbfs_bu_hubs),
// so that a few waves holding many such vertices do not hold up the level.
constexpr uint32_t kHubChunk = 512;
constexpr int kLaneUnroll = 4;   // bottom-up lane path: in-edges per round trip

struct MsState {
    uint32_t depth;
    uint32_t done;
    uint32_t num_levels;
    uint32_t topdown;                // direction of the current level (1: top-down)
    unsigned long long queue;        // top-down: (queued vertices << 32) | out-edges reserved
    unsigned long long next_edges;   // out-edges of the vertices reached at this level
};

struct MsArgs {
    uint64_t V;
    uint64_t E;                // forward edges
    const uint32_t *roff;      // reverse CSR (bottom-up)
    const uint32_t *rcol;
    const uint32_t *off;       // forward CSR (top-down)
    const uint32_t *col;
    uint64_t *visit;
    uint64_t *front[2];  // level d reads front[d & 1], writes front[(d & 1) ^ 1]
    uint64_t *next;      // ORed frontier masks per vertex from the top-down edges or the hub
                         // chunks (zero between levels)
    uint2 *queue;        // top-down: (vertex, first reserved edge) per frontier vertex
    const uint4 *hubs;   // bottom-up: (vertex, first in-edge, end) per hub chunk
    uint32_t nhubs;
    uint32_t hub_degree; // in-degree above which a vertex is covered by hub chunks
    int32_t *depths;     // [nsrc][V]
    uint64_t all;        // mask of the nsrc searches
    unsigned long long *level_bits;  // per level: newly set bits
    MsState *st;
    uint32_t max_levels;
    uint32_t wave_degree;
    uint32_t td_div;     // top-down when next_edges * td_div <= E (0: never)
};

__device__ __forceinline__ uint64_t wave_or64(uint64_t x)
{
    for (int o = kWave / 2; o > 0; o >>= 1) x |= __shfl_xor(x, o, kWave);
    return x;
}

// Depth nd for the searches in nb (0 for lanes with nothing).  Called by all
// lanes of the wave together: the lanes store one search's row at a time,
// so the stores of one instruction go to adjacent vertices.
__device__ __forceinline__ void record(const MsArgs &a, uint64_t v, uint64_t nb, int32_t nd)
{
    uint64_t any = wave_or64(nb);
    while (any) {
        const int s = __ffsll(static_cast<unsigned long long>(any)) - 1;
        any &= any - 1;
        if ((nb >> s) & 1ull) a.depths[static_cast<uint64_t>(s) * a.V + v] = nd;
    }
}

__device__ __forceinline__ void block_add(unsigned long long x, unsigned long long *dst,
                                          unsigned long long (*part)[kBlock / kWave])
{
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    if (__lane_id() == 0) (*part)[threadIdx.x / kWave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / kWave; w++) t += (*part)[w];
        if (t) atomicAdd(dst, t);
    }
    __syncthreads();
}

// One level over all vertices.  Bottom-up: pulls the frontier masks of the
// in-neighbours.  Top-down: takes the masks ORed into next[] by
// msbfs_td_expand (and clears them).  Then the new bits update the visit
// mask, the next frontier and the depths, and the block adds its new bits
// and the next frontier's out-edges to the level's counters.
__global__ __launch_bounds__(kBlock) void msbfs_level(MsArgs a)
{
    __shared__ unsigned long long part[kBlock / kWave];
    MsState &st = *a.st;
    if (st.done) return;
    const uint32_t d = st.depth;
    const bool td = st.topdown != 0;
    const int32_t nd = static_cast<int32_t>(d) + 1;
    // Selects rather than indexes, so the argument array stays in registers.
    const uint64_t *fin = (d & 1) ? a.front[1] : a.front[0];
    uint64_t *fout = (d & 1) ? a.front[0] : a.front[1];
    const int lane = __lane_id();
    unsigned long long found = 0, nedges = 0;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlock; base < a.V;
         base += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t v64 = base + threadIdx.x;
        const uint32_t v = static_cast<uint32_t>(v64);
        const bool live = v64 < a.V;
        const uint64_t vis = live ? a.visit[v] : a.all;
        uint64_t acc = 0;
        if (td) {
            if (live) {
                acc = a.next[v];
                if (acc) a.next[v] = 0;
            }
        } else {
            bool open = vis != a.all;
            uint32_t s = 0, e = 0;
            if (open) {
                s = a.roff[v];
                e = a.roff[v + 1];
                if (e - s > a.hub_degree) {  // covered by msbfs_bu_hubs
                    acc = a.next[v];
                    if (acc) a.next[v] = 0;
                    open = false;
                    e = s;
                }
            }
            // The lane probes the first wave_degree in-edges itself; a row
            // still missing searches after that is finished by the whole wave.
            const uint32_t lim = e - s > a.wave_degree ? s + a.wave_degree : e;
            for (uint32_t j0 = s; j0 < lim; j0 += kLaneUnroll) {
                uint32_t c[kLaneUnroll];
#pragma unroll
                for (int u = 0; u < kLaneUnroll; u++) c[u] = j0 + u < lim ? a.rcol[j0 + u] : 0xFFFFFFFFu;
#pragma unroll
                for (int u = 0; u < kLaneUnroll; u++) acc |= c[u] != 0xFFFFFFFFu ? fin[c[u]] : 0;
                if ((acc | vis) == a.all) break;
            }
            uint64_t m = __ballot(open && lim < e && (acc | vis) != a.all);
            while (m) {
                const int l = __ffsll(static_cast<unsigned long long>(m)) - 1;
                m &= m - 1;
                const uint32_t vs = __shfl(lim, l, kWave), ve = __shfl(e, l, kWave);
                const uint64_t lvis = __shfl(vis | acc, l, kWave);
                // kUnroll chunks of 64 in-edges per round trip: their loads
                // are in flight together before the coverage test.
                uint64_t wacc = 0;
                for (uint32_t k0 = vs; k0 < ve; k0 += kWave * kUnroll) {
                    uint64_t x = 0;
#pragma unroll
                    for (int u = 0; u < kUnroll; u++) {
                        const uint32_t j = k0 + u * kWave + lane;
                        x |= j < ve ? fin[a.rcol[j]] : 0;
                    }
                    wacc |= wave_or64(x);
                    if ((wacc | lvis) == a.all) break;
                }
                if (lane == l) acc |= wacc;
            }
        }
        const uint64_t nb = acc & ~vis;
        if (live) {
            fout[v] = nb;
            if (nb) {
                a.visit[v] = vis | nb;
                found += __popcll(nb);
                nedges += a.off[v + 1] - a.off[v];
            }
        }
        record(a, v64, nb, nd);
    }
    block_add(found, &a.level_bits[d], &part);
    block_add(nedges, &st.next_edges, &part);
}

// Top-down, step 1: queues the frontier vertices with out-edges, reserving
// a contiguous range of edge slots for each.  Workgroup b takes a
// contiguous chunk of vertices: it counts its entries and edges, reserves
// them with one atomic, then writes them, so within the queue the reserved
// ranges increase with the entry.
__device__ __forceinline__ void td_scan(const MsArgs &a, MsState &st)
{
    __shared__ uint32_t wcnt[kBlock / kWave], wdeg[kBlock / kWave];
    __shared__ unsigned long long base_slot;
    const uint64_t *fin = (st.depth & 1) ? a.front[1] : a.front[0];
    const int lane = __lane_id(), wid = threadIdx.x / kWave;
    const uint64_t chunk = (a.V + gridDim.x - 1) / gridDim.x;
    const uint64_t v0 = blockIdx.x * chunk, v1 = min(a.V, v0 + chunk);
    // Pass 1: this thread's entries and edges.
    uint32_t cnt = 0, deg = 0;
    for (uint64_t v = v0 + threadIdx.x; v < v1; v += kBlock)
        if (fin[v]) {
            const uint32_t d = a.off[v + 1] - a.off[v];
            cnt += d != 0;
            deg += d;
        }
    // Exclusive prefixes over the workgroup (threads in order).
    uint32_t pc = cnt, pd = deg;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t yc = __shfl_up(pc, static_cast<unsigned>(o), kWave);
        const uint32_t yd = __shfl_up(pd, static_cast<unsigned>(o), kWave);
        if (lane >= o) {
            pc += yc;
            pd += yd;
        }
    }
    if (lane == kWave - 1) {
        wcnt[wid] = pc;
        wdeg[wid] = pd;
    }
    __syncthreads();
    uint32_t bc = 0, bd = 0, tc = 0, td = 0;
    for (int w = 0; w < kBlock / kWave; w++) {
        if (w < wid) {
            bc += wcnt[w];
            bd += wdeg[w];
        }
        tc += wcnt[w];
        td += wdeg[w];
    }
    if (threadIdx.x == 0)
        base_slot = tc ? atomicAdd(&st.queue, (static_cast<unsigned long long>(tc) << 32) | td) : 0ull;
    __syncthreads();
    if (tc == 0) return;
    uint32_t idx = static_cast<uint32_t>(base_slot >> 32) + bc + pc - cnt;
    uint32_t pos = static_cast<uint32_t>(base_slot) + bd + pd - deg;
    // Pass 2: the same vertices in the same order, written out.
    for (uint64_t v = v0 + threadIdx.x; v < v1; v += kBlock)
        if (fin[v]) {
            const uint32_t d = a.off[v + 1] - a.off[v];
            if (d) {
                a.queue[idx++] = make_uint2(static_cast<uint32_t>(v), pos);
                pos += d;
            }
        }
}

// Top-down, step 2: each thread takes kSlots consecutive reserved edge
// slots.  The queue entry holding the first is found by binary search, the
// next ones by stepping forward; the edge u -> w in a slot ORs u's frontier
// bits that w has not reached yet into next[w].
__global__ __launch_bounds__(kBlock) void msbfs_td_expand(MsArgs a)
{
    const MsState &st = *a.st;
    if (st.done || !st.topdown) return;
    const uint64_t *fin = (st.depth & 1) ? a.front[1] : a.front[0];
    const unsigned long long q = st.queue;
    const uint32_t n = static_cast<uint32_t>(q >> 32), T = static_cast<uint32_t>(q);
    const uint64_t nchunks = (static_cast<uint64_t>(T) + kSlots - 1) / kSlots;
    for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; c < nchunks;
         c += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint32_t t0 = static_cast<uint32_t>(c * kSlots);
        const uint32_t t1 = static_cast<uint32_t>(min(static_cast<uint64_t>(T), c * kSlots + kSlots));
        uint32_t lo = 0, hi = n;  // queue[lo].y <= t0 < queue[hi].y, with queue[n].y taken as T
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.queue[mid].y <= t0) lo = mid;
            else hi = mid;
        }
        uint2 e = a.queue[lo];
        uint32_t end = lo + 1 < n ? a.queue[lo + 1].y : T;
        uint64_t m = fin[e.x];
        for (uint32_t t = t0; t < t1; t++) {
            while (t >= end) {  // next entry (entries have at least one edge)
                lo++;
                e = a.queue[lo];
                end = lo + 1 < n ? a.queue[lo + 1].y : T;
                m = fin[e.x];
            }
            const uint32_t w = a.col[a.off[e.x] + (t - e.y)];
            const uint64_t nw = m & ~a.visit[w];
            if (nw && (nw & ~a.next[w])) atomicOr(reinterpret_cast<unsigned long long *>(&a.next[w]), nw);
        }
    }
}

// Bottom-up, hubs: one wave per chunk of a hub's in-edges ORs their
// frontier masks (stopping once the hub's missing searches are covered,
// counting what other chunks already found) into next[hub].
__device__ __forceinline__ void bu_hubs(const MsArgs &a, const MsState &st)
{
    const uint64_t *fin = (st.depth & 1) ? a.front[1] : a.front[0];
    const int lane = __lane_id();
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    for (uint32_t c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; c < a.nhubs; c += nwaves) {
        const uint4 h = a.hubs[c];
        const uint64_t vis = a.visit[h.x];
        if (vis == a.all) continue;
        uint64_t acc = 0;
        for (uint32_t k0 = h.y; k0 < h.z; k0 += kWave * kUnroll) {
            uint64_t x = 0;
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const uint32_t j = k0 + u * kWave + lane;
                x |= j < h.z ? fin[a.rcol[j]] : 0;
            }
            acc |= wave_or64(x);
            if ((acc | vis | a.next[h.x]) == a.all) break;
        }
        const uint64_t nb = acc & ~vis;
        if (lane == 0 && (nb & ~a.next[h.x])) atomicOr(reinterpret_cast<unsigned long long *>(&a.next[h.x]), nb);
    }
}

// The first kernel of a level, whichever its direction: the top-down
// queue (td_scan) or the bottom-up hub chunks (bu_hubs).  One launch, where
// two kernels would each return at once on the levels of the other
// direction.
__global__ __launch_bounds__(kBlock) void msbfs_scan_or_hubs(MsArgs a)
{
    MsState &st = *a.st;
    if (st.done) return;
    if (st.topdown) td_scan(a, st);
    else if (a.nhubs) bu_hubs(a, st);
}

// Closes level d; picks the direction of level d + 1 from its frontier's
// out-edges.
__global__ void msbfs_epilogue(MsArgs a)
{
    MsState &st = *a.st;
    if (st.done) return;
    const uint32_t d = st.depth;
    st.num_levels = d + 1;
    if (a.level_bits[d] == 0 || d + 1 >= a.max_levels) {
        st.done = 1;
    } else {
        st.depth = d + 1;
        st.topdown = a.td_div != 0 && st.next_edges * a.td_div <= a.E;
    }
    st.queue = 0;
    st.next_edges = 0;
}

// Depth 0: source i gets bit i (a repeated source vertex gets several).
__global__ void msbfs_init(MsArgs a, const uint32_t *src, uint32_t n)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long edges = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t v = src[i];
        a.visit[v] |= 1ull << i;
        a.front[0][v] |= 1ull << i;
        a.depths[static_cast<uint64_t>(i) * a.V + v] = 0;
        edges += a.off[v + 1] - a.off[v];
    }
    a.st->depth = 0;
    a.st->done = 0;
    a.st->num_levels = 0;
    a.st->topdown = a.td_div != 0 && edges * a.td_div <= a.E;
    a.st->queue = 0;
    a.st->next_edges = 0;
}

}  // namespace
}  // namespace bw

using namespace bw;

struct bw_msbfs {
    uint64_t V = 0, E = 0;
    const uint32_t *roff = nullptr, *rcol = nullptr;
    const uint32_t *off = nullptr, *col = nullptr;
    uint64_t *visit = nullptr, *front[2] = {nullptr, nullptr};
    uint64_t *next = nullptr;
    uint2 *queue = nullptr;
    uint4 *hubs = nullptr;
    uint32_t nhubs = 0;
    uint32_t hub_degree = 0;  // the hub chunks are built for this degree (0: none yet)
    unsigned long long *level_bits = nullptr;
    MsState *st = nullptr;
    uint32_t *src = nullptr;
    uint32_t max_levels = 0;
    uint32_t wave_degree = BW_MSBFS_DEFAULT_WAVE_DEGREE;
    uint32_t td_div = BW_MSBFS_DEFAULT_TOPDOWN_DIVISOR;
    uint32_t grid = 0;
    hipStream_t own_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t *exec_depths = nullptr;
    uint64_t exec_all = 0;
    int32_t exec_lps = 0;
    uint32_t exec_wave = 0;
    uint32_t exec_td = 0;
    const uint4 *exec_hubs = nullptr;
    uint32_t exec_nhubs = 0;
    void *exec_stream = nullptr;

    ~bw_msbfs()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        void *bufs[] = {visit, front[0], front[1], next, queue, hubs, level_bits, st, src};
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
    a.E = m->E;
    a.roff = m->roff;
    a.rcol = m->rcol;
    a.off = m->off;
    a.col = m->col;
    a.visit = m->visit;
    a.front[0] = m->front[0];
    a.front[1] = m->front[1];
    a.next = m->next;
    a.queue = m->queue;
    a.hubs = m->hubs;
    a.nhubs = m->nhubs;
    a.hub_degree = m->hub_degree;
    a.depths = depths;
    a.all = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
    a.level_bits = m->level_bits;
    a.st = m->st;
    a.max_levels = m->max_levels;
    a.wave_degree = m->wave_degree;
    a.td_div = m->td_div;
    return a;
}

// Hub chunks of the bottom-up levels for vertices with more than
// hub_degree in-edges, from the reverse offsets.
int build_hubs(bw_msbfs *m, uint32_t hub_degree)
{
    std::vector<uint32_t> roff(m->V + 1);
    BW_HIP_CHECK(hipMemcpy(roff.data(), m->roff, 4 * (m->V + 1), hipMemcpyDeviceToHost));
    std::vector<uint4> hubs;
    for (uint64_t v = 0; v < m->V; v++)
        if (roff[v + 1] - roff[v] > hub_degree)
            for (uint32_t j = roff[v]; j < roff[v + 1]; j += kHubChunk)
                hubs.push_back(make_uint4(static_cast<uint32_t>(v), j, std::min(j + kHubChunk, roff[v + 1]), 0));
    uint4 *d = nullptr;
    if (!hubs.empty() && (hipMalloc(&d, sizeof(uint4) * hubs.size()) != hipSuccess ||
                          hipMemcpy(d, hubs.data(), sizeof(uint4) * hubs.size(), hipMemcpyHostToDevice) != hipSuccess)) {
        if (d) (void)hipFree(d);
        return fail(BW_HIP_ERR_NOMEM, "device allocation of the hub chunks failed");
    }
    if (m->hubs) (void)hipFree(m->hubs);
    m->hubs = d;
    m->nhubs = static_cast<uint32_t>(hubs.size());
    m->hub_degree = hub_degree;
    return BW_HIP_OK;
}

int launch_levels(const bw_msbfs *m, const MsArgs &a, int32_t count, hipStream_t s)
{
    for (int32_t i = 0; i < count; i++) {
        hipLaunchKernelGGL(msbfs_scan_or_hubs, dim3(m->grid), dim3(kBlock), 0, s, a);
        BW_LAUNCH_CHECK();
        hipLaunchKernelGGL(msbfs_td_expand, dim3(m->grid), dim3(kBlock), 0, s, a);
        BW_LAUNCH_CHECK();
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
    if (g->offset_bytes != 4 || !g->rev_offsets || !g->offsets || (g->edge_count && (!g->rev_cols || !g->cols)))
        return fail(BW_HIP_ERR_ARG, "multi-source BFS needs the forward and reverse CSR with uint32 offsets");
    if (g->edge_count >= 0xFFFFFFFFull) return fail(BW_HIP_ERR_ARG, "multi-source BFS needs E < 2^32 - 1");
    if (g->vertex_count == 0 || g->vertex_count >= 0x7FFFFFFFull)
        return fail(BW_HIP_ERR_ARG, "multi-source BFS needs 0 < V < 2^31");
    auto m = new bw_msbfs();
    m->V = g->vertex_count;
    m->E = g->edge_count;
    m->roff = static_cast<const uint32_t *>(g->rev_offsets);
    m->rcol = g->rev_cols;
    m->off = static_cast<const uint32_t *>(g->offsets);
    m->col = g->cols;
    m->max_levels = max_levels;
    m->grid = static_cast<uint32_t>(device_cus()) * 8;
    bool ok = hipMalloc(&m->visit, 8 * m->V) == hipSuccess && hipMalloc(&m->front[0], 8 * m->V) == hipSuccess &&
              hipMalloc(&m->front[1], 8 * m->V) == hipSuccess && hipMalloc(&m->next, 8 * m->V) == hipSuccess &&
              hipMalloc(&m->queue, sizeof(uint2) * m->V) == hipSuccess &&
              hipMalloc(&m->level_bits, 8ull * max_levels) == hipSuccess &&
              hipMalloc(&m->st, sizeof(MsState)) == hipSuccess && hipMalloc(&m->src, 4 * 64) == hipSuccess;
    if (!ok) {
        delete m;
        return fail(BW_HIP_ERR_NOMEM, "device allocation of the multi-source BFS failed");
    }
    const int hs = build_hubs(m, BW_MSBFS_DEFAULT_HUB_DEGREE);
    if (hs != BW_HIP_OK) {
        delete m;
        return hs;
    }
    *out = m;
    return BW_HIP_OK;
}

extern "C" int bw_msbfs_destroy(bw_msbfs *m)
{
    delete m;
    return BW_HIP_OK;
}

extern "C" int bw_msbfs_set_tuning(bw_msbfs *m, uint32_t wave_degree, uint32_t topdown_divisor,
                                   uint32_t hub_degree)
{
    if (!m) return fail(BW_HIP_ERR_ARG, "NULL handle");
    if (topdown_divisor > 1024) return fail(BW_HIP_ERR_ARG, "topdown_divisor must be 0..1024");
    if (hub_degree < 16) return fail(BW_HIP_ERR_ARG, "hub_degree must be at least 16");
    if (hub_degree != m->hub_degree) BW_TRY(build_hubs(m, hub_degree));
    m->wave_degree = wave_degree;
    m->td_div = topdown_divisor;
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
                      m->exec_lps != levels_per_sync || m->exec_wave != m->wave_degree || m->exec_td != m->td_div ||
                      m->exec_hubs != m->hubs || m->exec_nhubs != m->nhubs || m->exec_stream != s)) {
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
        m->exec_td = m->td_div;
        m->exec_hubs = m->hubs;
        m->exec_nhubs = m->nhubs;
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
    step(bw_hip_fill_u32(m->next, 0u, 2 * m->V, s));
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
