// Device-driven BFS engine for gfx950: queue-based top-down and bottom-up
// level kernels, a per-level epilogue that computes the frontier features
// and picks the next level's variant (fixed, scheduled, host-forced, or by a
// small MLP), and a level loop that needs no host round trip per level.
//
// Every variant computes the level of merijn/Belewitte's BFS
// (bfs/bfs.cpp:93-107, bfs/vertex_push.cu, bfs/vertex_pull.cu): vertices at
// `depth` discover their unvisited neighbours, which get depth + 1; the
// frontier count is the number of newly discovered vertices.  So every
// per-level schedule of variants yields the same depths and counts.  The
// variants differ only in how they find the work:
//   TD       top-down over the frontier queue; a workgroup takes 256 queue
//            vertices, scans their degrees in LDS and expands their edges 256
//            at a time (edge-balanced); vertices whose out-degree exceeds
//            the hub threshold (tunable, 4096 by default) are queued
//            separately and their edges are sliced across all workgroups.
//   BU       bottom-up: one thread per unvisited vertex probes its
//            in-neighbours and stops at the first one in the frontier.
//   BU_WAVE  bottom-up where each lane probes only its vertex's first
//            wave-threshold in-edges (tunable, 32 by default) and the whole
//            wave finishes the rows still unresolved (256 in-edges per step,
//            ballot early exit), so one long row never serialises a lane.
// A search may start from several roots at once (all at depth 0), which
// runs many small graphs stored as one disjoint union in one search.
// Frontier membership for the bottom-up probes is a bitmap (V bits, three
// rotating: read, written, cleared), visited-ness a fourth bitmap, so the
// probes hit a few MiB instead of the 4-byte depth array.  Discovered
// vertices are appended to the next frontier queue through an LDS buffer
// (wave-aggregated LDS atomics, one global atomic per flush).

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

namespace bw {
namespace {

constexpr int kBlock = 256;
// Staged discoveries per workgroup (dynamic LDS of the level kernel) and
// hubs per level (more go to the normal queue): runtime sizes up to these
// maxima (bw_bfs_engine_set_queues).
constexpr uint32_t kLdsQueueMax = BW_BFS_MAX_LDS_QUEUE;
constexpr uint32_t kHubCapMax = BW_BFS_MAX_HUB_CAP;
constexpr uint32_t kParentScan = 32;   // in-edges scanned for a vertex's likely parent
constexpr uint32_t kNoParent = 0xFFFFFFFFu;
constexpr int kFeatures = BW_BFS_FEATURES;
constexpr int kHidden = 32;
constexpr int kMaxOut = 16;
constexpr int kVariants = BW_BFS_ENGINE_VARIANTS;  // native, then the reference families
constexpr int kRef = BW_BFS_ENGINE_REFERENCE;

enum Variant : int32_t { TD = 0, BU = 1, BU_WAVE = 2 };
enum Policy : int32_t { FIXED = BW_POLICY_FIXED, SCHEDULE = BW_POLICY_SCHEDULE, LEARNED = BW_POLICY_LEARNED,
                        HOST = BW_POLICY_HOST, BEAMER = BW_POLICY_BEAMER };

struct BfsState {
    uint32_t depth;     // level being processed: the frontier is at `depth`
    uint32_t done;
    uint32_t qcount;    // normal frontier queue size
    uint32_t hubcount;  // hub frontier queue size
    uint32_t qout;      // appended to the next normal queue during the level
    uint32_t hubout;    // appended to the next hub queue (may exceed the hub capacity)
    int32_t choice;     // variant of this level
    uint32_t num_levels;
    int32_t prev_choice;        // variant of the previous level, -1 at level 0
    float prev_time_ns;         // device time of the previous level
    uint64_t visited;           // discovered so far, root included
    uint64_t frontier_edges;    // out-degree sum of the frontier
    double frontier_sq;         // sum of squared out-degrees of the frontier
    uint64_t t_step;            // wall clock at the start of the level's step kernel
    uint32_t fill_levels;       // the previous level left its depths to this one (dense top-down)
    uint32_t stalled;           // a level's variant was of a kind that was not launched (see run_impl)
    uint64_t explored_edges;    // out-degree sum of every vertex discovered so far
    uint64_t prev_frontier;     // frontier size of the previous level (0 at level 0)
};

struct Graph {
    uint64_t V;
    uint64_t E;
    const uint32_t *off;   // out-edges (top-down)
    const uint32_t *col;
    const uint32_t *roff;  // in-edges (bottom-up)
    const uint32_t *rcol;
    // The edge-list representations the reference's edge-list families
    // read (NULL when not resident): SoA forward (source, destination) and
    // reverse (destination, source), AoS pairs of each.
    const uint32_t *el_in, *el_out, *rel_in, *rel_out, *sel, *rsel;
};

// Whether variant c can run on this graph (the edge-list families need
// their representation).
__host__ __device__ inline bool variant_available(const Graph &g, int32_t c)
{
    if (c < 0 || c >= kVariants) return false;
    switch (c - kRef) {
    case BW_BFS_EDGE_LIST: return g.E == 0 || (g.el_in && g.el_out);
    case BW_BFS_REV_EDGE_LIST: return g.E == 0 || (g.rel_in && g.rel_out);
    case BW_BFS_STRUCT_EDGE_LIST: return g.E == 0 || g.sel;
    case BW_BFS_REV_STRUCT_EDGE_LIST: return g.E == 0 || g.rsel;
    default: return true;
    }
}

struct Buffers {
    int32_t *levels;
    uint32_t *q[2];
    uint32_t *hubq[2];
    uint64_t *hub_prefix;  // hub_cap + 1
    uint32_t *bits[3];     // frontier bitmaps, level d reads bits[d % 3]
    uint32_t *vis;         // visited bitmap
    const uint32_t *parent;  // per vertex: its likely parent (in-neighbour of largest out-degree among
                             // its first kParentScan in-edges), kNoParent without in-edges
    BfsState *st;
    uint64_t *part;        // per workgroup: out-degree sum, (double) sum of squares
    uint32_t *counts;      // per level: discovered vertices
    uint64_t *times;       // per level: device time in ns
    int32_t *choices;      // per level: variant
    uint32_t max_levels;
    uint32_t words;        // bitmap words
    float ns_per_tick;
    uint32_t hub_degree;   // out-degree above which a discovered vertex is a hub
    uint32_t wave_degree;  // in-degree above which BU_WAVE probes with a whole wave
    uint32_t lds_queue;    // staged discoveries per workgroup (<= kLdsQueueMax)
    uint32_t hub_cap;      // hub queue capacity, a multiple of kBlock (<= kHubCapMax)
};

struct PolicyArgs {
    int32_t policy;
    int32_t fixed;
    const int32_t *schedule;
    int32_t schedule_len;
    const float *mlp;  // packed weights, see mlp_len
    int32_t nout;
};

__device__ __forceinline__ uint32_t out_degree(const Graph &g, uint32_t v) { return g.off[v + 1] - g.off[v]; }
__device__ __forceinline__ bool test_bit(const uint32_t *bits, uint32_t v) { return (bits[v >> 5] >> (v & 31)) & 1u; }

// --------------------------------------------------------- queue emission

struct Emitter {
    uint32_t *lbuf;
    uint32_t *lcount;
    uint32_t *lbase;
    uint32_t cap;  // lbuf's size
    uint64_t outsum = 0;
    double sq = 0.0;

    // Called by the lanes that discovered a vertex (any subset of a wave).
    // mark: set w's bit in the next-frontier bitmap here (top-down, where
    // any workgroup may discover any vertex; try_discover has set its
    // visited bit); bottom-up levels set whole words per wave instead (see
    // level_bu).
    __device__ __forceinline__ void emit(const Graph &g, const Buffers &b, uint32_t d, uint32_t w, bool mark)
    {
        const uint32_t nq = (d & 1) ^ 1;
        const uint32_t dw = out_degree(g, w);
        outsum += dw;
        sq += static_cast<double>(dw) * static_cast<double>(dw);
        if (mark) atomicOr(&b.bits[(d + 1) % 3][w >> 5], 1u << (w & 31));
        if (dw > b.hub_degree) {
            // Slots are handed out by a counter that only grows, so every
            // slot below min(count, hub_cap) is written exactly once.
            const uint32_t h = atomicAdd(&b.st->hubout, 1u);
            if (h < b.hub_cap) {
                b.hubq[nq][h] = w;
                return;
            }
        }
        const uint64_t m = __ballot(1);
        const int lane = __lane_id();
        const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
        const uint32_t n = __popcll(m);
        const uint32_t rank = __popcll(m & ((1ull << lane) - 1ull));
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(lcount, n);
        base = __shfl(base, leader, kWave);
        const uint32_t pos = base + rank;
        if (pos < cap) {
            lbuf[pos] = w;
        } else {
            b.q[nq][atomicAdd(&b.st->qout, 1u)] = w;
        }
    }

    // Block-uniform: moves the staged vertices to the global queue when
    // more than `reserve` further slots might not fit.
    __device__ __forceinline__ void flush(const Buffers &b, uint32_t d, uint32_t reserve)
    {
        __syncthreads();
        const uint32_t n = min(*lcount, cap);
        __syncthreads();
        if (n == 0 || n + reserve <= cap) return;
        if (threadIdx.x == 0) *lbase = atomicAdd(&b.st->qout, n);
        __syncthreads();
        const uint32_t g0 = *lbase;
        uint32_t *q = b.q[(d & 1) ^ 1];
        for (uint32_t i = threadIdx.x; i < n; i += kBlock) q[g0 + i] = lbuf[i];
        __syncthreads();
        if (threadIdx.x == 0) *lcount = 0;
        __syncthreads();
    }
};

// Top-down discovery of w: the thread whose fetch-or sets w's visited bit
// owns w and records its depth (unless the level defers its depths).
// Edges to visited vertices cost one read of the visited bitmap (V bits,
// far more cacheable than the depth array).
__device__ __forceinline__ bool try_discover(const Buffers &b, uint32_t w, int32_t nd, bool write_depth)
{
    const uint32_t bit = 1u << (w & 31);
    if (b.vis[w >> 5] & bit) return false;
    if (atomicOr(&b.vis[w >> 5], bit) & bit) return false;
    if (write_depth) b.levels[w] = nd;
    return true;
}

// A dense top-down level (frontier out-edges above V/16) leaves the depths
// of its discoveries to the next level, which writes depth d + 1 for every
// vertex of its frontier bitmap with coalesced stores instead of one
// scattered store per discovery.  Never at the last levels the limit
// allows, so the next level always runs.
__device__ __forceinline__ bool defer_depths(const BfsState &st, const Buffers &b, uint64_t V)
{
    return st.choice == TD && st.frontier_edges * 16 > V && st.depth + 2 < b.max_levels;
}

// ------------------------------------------------------------- variants

__device__ void level_td(const Graph &g, const Buffers &b, uint32_t d, Emitter &em, uint32_t *sstart,
                         uint64_t *sscan, uint64_t *stotal)
{
    const int32_t nd = static_cast<int32_t>(d) + 1;
    const uint32_t cur = d & 1;
    const BfsState &st = *b.st;
    const uint32_t tid = threadIdx.x;
    const bool wdepth = !defer_depths(st, b, g.V);

    // Hub edges, sliced over all workgroups 256 edges at a time.
    const uint32_t nh = st.hubcount;
    if (nh) {
        const uint64_t H = b.hub_prefix[nh];
        for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * kBlock; c < H;
             c += static_cast<uint64_t>(gridDim.x) * kBlock) {
            const uint64_t k = c + tid;
            if (k < H) {
                uint32_t lo = 0, hi = nh;  // largest h with prefix[h] <= k
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) / 2;
                    if (b.hub_prefix[mid] <= k) lo = mid;
                    else hi = mid;
                }
                const uint32_t u = b.hubq[cur][lo];
                const uint32_t w = g.col[g.off[u] + static_cast<uint32_t>(k - b.hub_prefix[lo])];
                if (try_discover(b, w, nd, wdepth)) em.emit(g, b, d, w, true);
            }
            em.flush(b, d, kBlock);
        }
    }

    // Normal frontier: batches of up to 256 queue vertices, edges expanded
    // 256 at a time with the owner found by binary search in the LDS degree
    // scan.  A small frontier is spread over all workgroups (batches down to
    // one vertex), so its edges never pile up on a few workgroups.
    const uint32_t n = st.qcount;
    const uint32_t bs = max(1u, min(static_cast<uint32_t>(kBlock), (n + gridDim.x - 1) / gridDim.x));
    for (uint32_t base = blockIdx.x * bs; base < n; base += gridDim.x * bs) {
        const uint32_t i = base + tid;
        uint32_t deg = 0, start = 0;
        if (tid < bs && i < n) {
            const uint32_t u = b.q[cur][i];
            start = g.off[u];
            deg = g.off[u + 1] - start;
        }
        sstart[tid] = start;
        sscan[tid] = deg;
        __syncthreads();
        // Inclusive Hillis-Steele scan over the 256 degrees.
        for (uint32_t o = 1; o < kBlock; o <<= 1) {
            const uint64_t add = tid >= o ? sscan[tid - o] : 0;
            __syncthreads();
            sscan[tid] += add;
            __syncthreads();
        }
        if (tid == 0) *stotal = sscan[kBlock - 1];
        __syncthreads();
        const uint64_t total = *stotal;
        for (uint64_t c = 0; c < total; c += kBlock) {
            const uint64_t k = c + tid;
            if (k < total) {
                uint32_t lo = 0, hi = kBlock - 1;  // first j with inclusive scan > k
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) / 2;
                    if (sscan[mid] > k) hi = mid;
                    else lo = mid + 1;
                }
                const uint64_t before = lo ? sscan[lo - 1] : 0;
                const uint32_t w = g.col[sstart[lo] + static_cast<uint32_t>(k - before)];
                if (try_discover(b, w, nd, wdepth)) em.emit(g, b, d, w, true);
            }
            em.flush(b, d, kBlock);
        }
        __syncthreads();
    }
}

template<bool USE_PARENT>
__device__ void level_bu(const Graph &g, const Buffers &b, uint32_t d, Emitter &em, bool wave, bool fill)
{
    const int32_t nd = static_cast<int32_t>(d) + 1;
    const uint32_t *front = b.bits[d % 3];
    const uint64_t V = g.V;
    const int lane = __lane_id();
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlock; base < V;
         base += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t v64 = base + threadIdx.x;
        const uint32_t v = static_cast<uint32_t>(v64);
        // The depths a dense top-down level left: its discoveries are this
        // level's frontier.
        if (fill && v64 < V && test_bit(front, v)) b.levels[v] = static_cast<int32_t>(d);
        const bool unvisited = v64 < V && !test_bit(b.vis, v);
        // The likely parent first: one coalesced load and one probe resolve
        // most vertices of a big level; the in-edges are scanned otherwise.
        const uint32_t pv = USE_PARENT && unvisited ? b.parent[v] : kNoParent;
        const bool quick = pv != kNoParent && test_bit(front, pv);
        uint32_t s = 0, e = 0;
        if (unvisited && !quick) {
            s = g.roff[v];
            e = g.roff[v + 1];
        }
        // Each lane first probes its own vertex's in-edges, four independent
        // probes per step; BU_WAVE stops a lane after wave_degree edges and
        // hands the rest of a still-unresolved row to the whole wave.
        const uint32_t lim = wave && e - s > b.wave_degree ? s + b.wave_degree : e;
        bool discovered = quick;
        for (uint32_t j = s; j < lim; j += 4) {
            bool hit = false;
#pragma unroll
            for (uint32_t q = 0; q < 4; q++)
                if (j + q < lim) hit |= test_bit(front, g.rcol[j + q]);
            if (hit) {
                discovered = true;
                break;
            }
        }
        if (wave) {
            uint64_t m = __ballot(!discovered && lim < e);
            while (m) {
                const int l = __ffsll(static_cast<unsigned long long>(m)) - 1;
                m &= m - 1;
                const uint32_t vs = __shfl(lim, l, kWave), ve = __shfl(e, l, kWave);
                bool found = false;
                for (uint32_t k0 = vs; k0 < ve && !found; k0 += 4 * kWave) {
                    bool hit = false;
#pragma unroll
                    for (uint32_t q = 0; q < 4; q++) {
                        const uint32_t j = k0 + q * kWave + lane;
                        if (j < ve) hit |= test_bit(front, g.rcol[j]);
                    }
                    found = __ballot(hit) != 0;
                }
                if (found && lane == l) discovered = true;
            }
        }
        if (discovered) {
            b.levels[v] = nd;
            em.emit(g, b, d, v, false);
        }
        // The wave's 64 vertices are two whole bitmap words that no other
        // wave touches in a bottom-up level: plain stores, no atomics.  The
        // next-frontier words were cleared two levels ago.
        const uint64_t fm = __ballot(discovered);
        if (fm) {
            const uint64_t w0 = (base + (threadIdx.x & ~static_cast<uint32_t>(kWave - 1))) >> 5;
            const uint32_t half = lane == 0 ? static_cast<uint32_t>(fm) : static_cast<uint32_t>(fm >> 32);
            if ((lane == 0 || lane == 32) && half) {
                const uint64_t wi = w0 + (lane >> 5);
                b.bits[(d + 1) % 3][wi] = half;
                b.vis[wi] |= half;
            }
        }
        em.flush(b, d, kBlock);
    }
}

// ------------------------------------------ the reference's families

// The reference's kernel families as engine variants (bfs/edgelist.cu,
// rev_edgelist.cu, struct_edgelist.cu, rev_struct_edgelist.cu,
// vertex_push.cu, vertex_pull.cu, vertex_push_warp.cu, vertex_pull_warp.cu):
// each thread does the work the reference's thread does, with the frontier
// test on the engine's level bitmap and discovery through its visited
// bitmap and queues (so the next level of any variant can follow).

// Edge-parallel scan of every stored edge: an edge whose tested endpoint is
// in the frontier relaxes its other endpoint.  Forward lists test the
// source, reverse lists (destination, source) the source as well.
template<int FAMILY>
__device__ void level_ref_edges(const Graph &g, const Buffers &b, uint32_t d, Emitter &em)
{
    const int32_t nd = static_cast<int32_t>(d) + 1;
    const uint32_t *front = b.bits[d % 3];
    const uint64_t E = g.E;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlock * 4; base < E;
         base += static_cast<uint64_t>(gridDim.x) * kBlock * 4) {
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint64_t e = base + k * kBlock + threadIdx.x;
            if (e >= E) continue;
            uint32_t test, relax;
            if constexpr (FAMILY == BW_BFS_EDGE_LIST) {
                test = g.el_in[e];
                relax = g.el_out[e];
            } else if constexpr (FAMILY == BW_BFS_REV_EDGE_LIST) {
                test = g.rel_out[e];
                relax = g.rel_in[e];
            } else if constexpr (FAMILY == BW_BFS_STRUCT_EDGE_LIST) {
                test = g.sel[2 * e];
                relax = g.sel[2 * e + 1];
            } else {
                test = g.rsel[2 * e + 1];
                relax = g.rsel[2 * e];
            }
            if (test_bit(front, test) && try_discover(b, relax, nd, true)) em.emit(g, b, d, relax, true);
        }
        em.flush(b, d, 4 * kBlock);
    }
}

// vertex_push.cu: a thread per vertex expands its out-edges when it is in
// the frontier.
__device__ void level_ref_push(const Graph &g, const Buffers &b, uint32_t d, Emitter &em)
{
    const int32_t nd = static_cast<int32_t>(d) + 1;
    const uint32_t *front = b.bits[d % 3];
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlock; base < g.V;
         base += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t v = base + threadIdx.x;
        if (v < g.V && test_bit(front, static_cast<uint32_t>(v))) {
            for (uint32_t j = g.off[v]; j < g.off[v + 1]; j++) {
                const uint32_t w = g.col[j];
                if (try_discover(b, w, nd, true)) em.emit(g, b, d, w, true);
            }
        }
        em.flush(b, d, kBlock);
    }
}

// vertex_push_warp.cu / vertex_pull_warp.cu geometry: a virtual warp of 32
// lanes (half a wave) takes a chunk of 32 consecutive vertices and visits
// them one after another.
constexpr uint32_t kRefWarp = 32;

__device__ void level_ref_push_warp(const Graph &g, const Buffers &b, uint32_t d, Emitter &em)
{
    const int32_t nd = static_cast<int32_t>(d) + 1;
    const uint32_t *front = b.bits[d % 3];
    const uint32_t lane = threadIdx.x % kRefWarp;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlock; base < g.V;
         base += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t cb = base + (threadIdx.x & ~(kRefWarp - 1));
        for (uint32_t i = 0; i < kRefWarp && cb + i < g.V; i++) {
            const uint32_t v = static_cast<uint32_t>(cb + i);
            if (!test_bit(front, v)) continue;
            for (uint32_t j = g.off[v] + lane; j < g.off[v + 1]; j += kRefWarp) {
                const uint32_t w = g.col[j];
                if (try_discover(b, w, nd, true)) em.emit(g, b, d, w, true);
            }
        }
        em.flush(b, d, kBlock);
    }
}

// vertex_pull_warp.cu: a virtual warp probes each unvisited vertex of its
// chunk, its lanes striding over the in-edges and stopping at the first
// stride with a hit (a ballot over the half wave).  A wave's two virtual
// warps cover 64 consecutive vertices, two whole bitmap words, which it
// writes with plain stores.
__device__ void level_ref_pull_warp(const Graph &g, const Buffers &b, uint32_t d, Emitter &em)
{
    const int32_t nd = static_cast<int32_t>(d) + 1;
    const uint32_t *front = b.bits[d % 3];
    const uint32_t lane = threadIdx.x % kRefWarp;
    const uint64_t half = (threadIdx.x & kRefWarp) ? ~0ull << 32 : 0xFFFFFFFFull;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlock; base < g.V;
         base += static_cast<uint64_t>(gridDim.x) * kBlock) {
        // The wave's 64 vertices, its half's 32 of them.
        const uint64_t c0 = base + (threadIdx.x & ~(kRefWarp - 1));
        uint32_t found_mask = 0;
        for (uint32_t i = 0; i < kRefWarp; i++) {
            const uint64_t v = c0 + i;
            const bool open = v < g.V && !test_bit(b.vis, static_cast<uint32_t>(v));
            bool found = false;
            if (open) {
                const uint32_t s = g.roff[v], e = g.roff[v + 1];
                for (uint32_t j0 = s; j0 < e && !found; j0 += kRefWarp) {
                    const uint32_t j = j0 + lane;
                    found = (__ballot(j < e && test_bit(front, g.rcol[j])) & half) != 0;
                }
            }
            if (found) {
                found_mask |= 1u << i;
                if (lane == 0) {
                    b.levels[v] = nd;
                    em.emit(g, b, d, static_cast<uint32_t>(v), false);
                }
            }
        }
        if (lane == 0 && found_mask) {
            const uint64_t wi = c0 >> 5;
            b.bits[(d + 1) % 3][wi] = found_mask;
            b.vis[wi] |= found_mask;
        }
        em.flush(b, d, kBlock);
    }
}

// ---------------------------------------------------------- level kernel

// The level kernel in two instantiations: REF = false runs the native
// variants, REF = true the reference's families; each returns at once when
// the level's variant is the other kind.  Kept apart so that the native
// kernel's code is compiled on its own (with the reference families in the
// same kernel its levels ran about 1.5x slower on MI355X).
template<bool REF>
__global__ __launch_bounds__(kBlock) void bfs_engine_step(Graph g, Buffers b)
{
    extern __shared__ uint32_t lbuf[];  // b.lds_queue entries
    __shared__ uint32_t lcount, lbase;
    __shared__ uint64_t stotal;
    __shared__ uint32_t sstart[kBlock];
    __shared__ uint64_t sscan[kBlock];
    __shared__ uint64_t rout[kBlock / kWave];
    __shared__ double rsq[kBlock / kWave];

    BfsState &st = *b.st;
    if (st.done) return;
    const uint32_t d = st.depth;
    const int32_t variant = st.choice;
    if ((variant >= kRef) != REF) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) st.t_step = wall_clock64();

    // Clear the bitmap that the next level will write.
    uint32_t *clear = b.bits[(d + 2) % 3];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < b.words; i += gridDim.x * kBlock) clear[i] = 0;

    if (threadIdx.x == 0) lcount = 0;
    __syncthreads();
    Emitter em{lbuf, &lcount, &lbase, b.lds_queue};
    const bool fill = st.fill_levels != 0;
    // Bottom-up forms write the deferred depths in their vertex loop; the
    // others first write them from this level's frontier bitmap.
    auto fill_depths = [&] {
        if (!fill) return;
        const uint32_t *front = b.bits[d % 3];
        for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; v < g.V;
             v += static_cast<uint64_t>(gridDim.x) * kBlock)
            if (test_bit(front, static_cast<uint32_t>(v))) b.levels[v] = static_cast<int32_t>(d);
    };
    if constexpr (!REF) {
        if (variant == TD) {
            fill_depths();
            level_td(g, b, d, em, sstart, sscan, &stotal);
        } else {
            level_bu<true>(g, b, d, em, variant == BU_WAVE, fill);
        }
    } else {
        if (variant == kRef + BW_BFS_VERTEX_PULL) {
            level_bu<false>(g, b, d, em, false, fill);
        } else {
            fill_depths();
            switch (variant - kRef) {
            case BW_BFS_EDGE_LIST: level_ref_edges<BW_BFS_EDGE_LIST>(g, b, d, em); break;
            case BW_BFS_REV_EDGE_LIST: level_ref_edges<BW_BFS_REV_EDGE_LIST>(g, b, d, em); break;
            case BW_BFS_STRUCT_EDGE_LIST: level_ref_edges<BW_BFS_STRUCT_EDGE_LIST>(g, b, d, em); break;
            case BW_BFS_REV_STRUCT_EDGE_LIST: level_ref_edges<BW_BFS_REV_STRUCT_EDGE_LIST>(g, b, d, em); break;
            case BW_BFS_VERTEX_PUSH: level_ref_push(g, b, d, em); break;
            case BW_BFS_VERTEX_PUSH_WARP: level_ref_push_warp(g, b, d, em); break;
            default: level_ref_pull_warp(g, b, d, em); break;
            }
        }
    }
    em.flush(b, d, em.cap);  // reserve everything: moves all that is staged

    // Workgroup partials of the discovered vertices' out-degrees.
    uint64_t os = em.outsum;
    double sq = em.sq;
    for (int o = kWave / 2; o > 0; o >>= 1) {
        os += __shfl_xor(os, o, kWave);
        sq += __shfl_xor(sq, o, kWave);
    }
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) {
        rout[wid] = os;
        rsq[wid] = sq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a = 0;
        double q = 0.0;
        for (int w = 0; w < kBlock / kWave; w++) {
            a += rout[w];
            q += rsq[w];
        }
        b.part[2 * blockIdx.x] = a;
        b.part[2 * blockIdx.x + 1] = static_cast<uint64_t>(__double_as_longlong(q));
    }
}

// ------------------------------------------------------------- the policy

// Packed MLP: mean[F] std[F] W1[32][F] b1[32] W2[32][32] b2[32] W3[K][32] b3[K],
// F = kFeatures.
__host__ __device__ constexpr int mlp_len(int K)
{ return 2 * kFeatures + kHidden * kFeatures + kHidden + kHidden * kHidden + kHidden + K * kHidden + K; }

// The features of the pending level: frontier size, frontier out-edges and
// unvisited vertices as log2 fractions of the graph (log2((1 + x) / V),
// log2((1 + x) / E)), which carry over between graph scales; depth;
// log2(1 + x) of the mean and standard deviation of the frontier
// out-degree; previous variant; log2(1 + previous level time in ns); the
// scale log2 V (the variants' relative speed shifts with how much of the
// bitmaps the caches hold); the undiscovered vertices' out-edges as a
// log2 fraction of E (m_u of Beamer's rule); the frontier's growth over the
// previous level in log2.
__host__ __device__ void features(const BfsState &st, uint64_t V, uint64_t E, float *f)
{
    const double n = static_cast<double>(st.qcount) + static_cast<double>(st.hubcount);
    const double edges = static_cast<double>(st.frontier_edges);
    const double mean = n > 0 ? edges / n : 0.0;
    const double var = n > 0 ? fmax(st.frontier_sq / n - mean * mean, 0.0) : 0.0;
    const double lv = log2(static_cast<double>(V)), le = log2(static_cast<double>(E > 0 ? E : 1));
    f[0] = static_cast<float>(log2(1.0 + n) - lv);
    f[1] = static_cast<float>(log2(1.0 + edges) - le);
    f[2] = static_cast<float>(log2(1.0 + static_cast<double>(V - st.visited)) - lv);
    f[3] = static_cast<float>(st.depth);
    f[4] = static_cast<float>(log2(1.0 + mean));
    f[5] = static_cast<float>(log2(1.0 + sqrt(var)));
    f[6] = static_cast<float>(st.prev_choice);
    f[7] = static_cast<float>(log2(1.0 + static_cast<double>(st.prev_time_ns)));
    f[8] = static_cast<float>(lv);
    const double mu = E > st.explored_edges ? static_cast<double>(E - st.explored_edges) : 0.0;
    f[9] = static_cast<float>(log2(1.0 + mu) - le);
    f[10] = static_cast<float>(log2(1.0 + n) - log2(1.0 + static_cast<double>(st.prev_frontier)));
}

// The model in wave 0 of a kernel, weights in registers: lane j holds row
// j % kHidden of W1 and W2 and lane k row min(k, K - 1) of W3 (lanes past
// the rows hold copies).  Loaded at the kernel's start so that the loads'
// latency overlaps the kernel's other work.
struct LaneModel {
    float w1[kFeatures], b1;
    float w2[kHidden], b2;
    float w3[kHidden], b3;
};

__device__ __forceinline__ void load_lane_model(const PolicyArgs &p, LaneModel &m)
{
    if (p.policy != LEARNED) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int j = lane & (kHidden - 1), k = min(lane, p.nout - 1);
    const float *W1 = p.mlp + 2 * kFeatures, *b1 = W1 + kHidden * kFeatures;
    const float *W2 = b1 + kHidden, *b2 = W2 + kHidden * kHidden;
    const float *W3 = b2 + kHidden, *b3 = W3 + p.nout * kHidden;
#pragma unroll
    for (int i = 0; i < kFeatures; i++) m.w1[i] = W1[j * kFeatures + i];
    m.b1 = b1[j];
#pragma unroll
    for (int i = 0; i < kHidden; i++) m.w2[i] = W2[j * kHidden + i];
    m.b2 = b2[j];
#pragma unroll
    for (int i = 0; i < kHidden; i++) m.w3[i] = W3[k * kHidden + i];
    m.b3 = b3[k];
}

// Lane i's value of x, read by the whole wave (v_readlane; i is uniform and
// every lane is active).
__device__ __forceinline__ float lane_value(float x, int i)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), i));
}

__device__ __forceinline__ double lane_value(double x, int i)
{
    const uint64_t u = static_cast<uint64_t>(__double_as_longlong(x));
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u & 0xFFFFFFFFu), i));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), i));
    return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

// Wave 0: MLP forward on the features fv (the same in every lane; one
// output per variant, the predicted log2 time) and the argmin -- the first
// minimum -- over the variants this graph can run (-1 if none).  The sums
// run in the order of the batch inference kernel, so both pick alike.
__device__ int32_t mlp_choose_wave(const Graph &g, const PolicyArgs &p, const LaneModel &m, const float *fv)
{
    const float *mean = p.mlp, *sd = p.mlp + kFeatures;
    float a = m.b1;
#pragma unroll
    for (int i = 0; i < kFeatures; i++) a += m.w1[i] * ((fv[i] - mean[i]) / sd[i]);
    const float h1 = fmaxf(a, 0.0f);
    a = m.b2;
#pragma unroll
    for (int i = 0; i < kHidden; i++) a += m.w2[i] * lane_value(h1, i);
    const float h2 = fmaxf(a, 0.0f);
    a = m.b3;
#pragma unroll
    for (int i = 0; i < kHidden; i++) a += m.w3[i] * lane_value(h2, i);
    int32_t best = -1;
    float bv = 0.0f;
    for (int k = 0; k < p.nout; k++) {
        const float v = lane_value(a, k);
        if (variant_available(g, k) && (best < 0 || v < bv)) {
            best = k;
            bv = v;
        }
    }
    return best;
}

// features() spread over wave 0: lane 0 reads the state and forms the ten
// logarithms' arguments, lane i takes the i-th logarithm, and every lane
// assembles the same features from them (one double-precision log2 of
// latency instead of ten in a row).
__device__ void features_wave(const BfsState &st, uint64_t V, uint64_t E, float *f)
{
    const int lane = threadIdx.x & (kWave - 1);
    double arg[10] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
    double depth = 0.0, prev = 0.0;
    if (lane == 0) {
        const double n = static_cast<double>(st.qcount) + static_cast<double>(st.hubcount);
        const double edges = static_cast<double>(st.frontier_edges);
        const double mean = n > 0 ? edges / n : 0.0;
        const double var = n > 0 ? fmax(st.frontier_sq / n - mean * mean, 0.0) : 0.0;
        arg[0] = static_cast<double>(V);
        arg[1] = static_cast<double>(E > 0 ? E : 1);
        arg[2] = 1.0 + n;
        arg[3] = 1.0 + edges;
        arg[4] = 1.0 + static_cast<double>(V - st.visited);
        arg[5] = 1.0 + mean;
        arg[6] = 1.0 + sqrt(var);
        arg[7] = 1.0 + static_cast<double>(st.prev_time_ns);
        arg[8] = 1.0 + (E > st.explored_edges ? static_cast<double>(E - st.explored_edges) : 0.0);
        arg[9] = 1.0 + static_cast<double>(st.prev_frontier);
        depth = static_cast<double>(st.depth);
        prev = static_cast<double>(st.prev_choice);
    }
    double mine = 1.0;
#pragma unroll
    for (int i = 0; i < 10; i++) {
        const double a = lane_value(arg[i], 0);
        if (lane == i) mine = a;
    }
    const double l = log2(mine);
    double lg[10];
#pragma unroll
    for (int i = 0; i < 10; i++) lg[i] = lane_value(l, i);
    const double lv = lg[0], le = lg[1];
    f[0] = static_cast<float>(lg[2] - lv);
    f[1] = static_cast<float>(lg[3] - le);
    f[2] = static_cast<float>(lg[4] - lv);
    f[3] = static_cast<float>(lane_value(depth, 0));
    f[4] = static_cast<float>(lg[5]);
    f[5] = static_cast<float>(lg[6]);
    f[6] = static_cast<float>(lane_value(prev, 0));
    f[7] = static_cast<float>(lg[7]);
    f[8] = static_cast<float>(lv);
    f[9] = static_cast<float>(lg[8] - le);
    f[10] = static_cast<float>(lg[2] - lg[9]);
}

// Wave 0, after lane 0 has brought the state to the pending level; the
// result is valid in lane 0.
__device__ int32_t choose(const PolicyArgs &p, const BfsState &st, const Graph &g, const LaneModel &m)
{
    const int lane = threadIdx.x & (kWave - 1);
    if (p.policy == LEARNED) {
        float fv[kFeatures];
        features_wave(st, g.V, g.E, fv);
        const int32_t c = mlp_choose_wave(g, p, m, fv);
        return c >= 0 ? c : p.fixed;
    }
    if (lane != 0) return p.fixed;
    switch (p.policy) {
    case SCHEDULE: return p.schedule[st.depth % static_cast<uint32_t>(p.schedule_len)];
    case HOST: return st.choice;
    case BEAMER: {
        // After a top-down level: bottom-up once the frontier's out-edges
        // m_f exceed the unexplored vertices' out-edges m_u / 14.  After a
        // bottom-up level: top-down again once the frontier both shrinks
        // and holds fewer than V / 24 vertices.
        const double mf = static_cast<double>(st.frontier_edges);
        const double mu = static_cast<double>(g.E) - static_cast<double>(st.explored_edges);
        const uint64_t nf = static_cast<uint64_t>(st.qcount) + st.hubcount;
        const bool bottom_up = st.prev_choice == BU_WAVE || st.prev_choice == BU;
        if (!bottom_up) return mf > mu / 14.0 ? BU_WAVE : TD;
        const bool stay = nf >= st.prev_frontier || static_cast<double>(nf) > static_cast<double>(g.V) / 24.0;
        return stay ? BU_WAVE : TD;
    }
    default: return p.fixed;
    }
}

// Wave 0: picks the pending level's variant unless the search is done.
__device__ __forceinline__ void choose_next(const PolicyArgs &p, BfsState &st, const Graph &g, const LaneModel &m)
{
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t done = 0;
    if (lane == 0) done = st.done;
    if (__shfl(done, 0, kWave)) return;
    const int32_t c = choose(p, st, g, m);
    if (lane == 0) st.choice = c;
}

// The per-level epilogue: one workgroup.  Records the finished level,
// advances the state to the next frontier and picks its variant.
__global__ __launch_bounds__(kBlock) void bfs_engine_epilogue(Graph g, Buffers b, PolicyArgs p, uint32_t nparts)
{
    __shared__ uint64_t rout[kBlock / kWave];
    __shared__ double rsq[kBlock / kWave];
    __shared__ uint64_t hscan[kBlock];

    BfsState &st = *b.st;
    if (st.done) return;
    const uint64_t now = wall_clock64();
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    // The step kernel that ran the level stamped t_step; none did when the
    // level's variant is of a kind this launch left out.  Then the level is
    // still pending: the state stays as it is, and the host relaunches it
    // with both kinds.
    if (st.t_step == 0) {
        if (tid == 0) st.stalled = 1;
        return;
    }
    LaneModel lm;
    if (wid == 0) load_lane_model(p, lm);
    uint64_t os = 0;
    double sq = 0.0;
    // Four partials per thread in flight, added in the order of a plain
    // stride loop.
    for (uint32_t i0 = tid; i0 < nparts; i0 += 4 * kBlock) {
        uint64_t pa[4], pq[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t i = i0 + j * kBlock;
            pa[j] = i < nparts ? b.part[2 * i] : 0;
            pq[j] = i < nparts ? b.part[2 * i + 1] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (i0 + j * kBlock >= nparts) break;
            os += pa[j];
            sq += __longlong_as_double(static_cast<long long>(pq[j]));
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        os += __shfl_xor(os, o, kWave);
        sq += __shfl_xor(sq, o, kWave);
    }
    if (lane == 0) {
        rout[wid] = os;
        rsq[wid] = sq;
    }
    const uint32_t d = st.depth;
    const uint32_t nhub = min(st.hubout, b.hub_cap);
    const uint32_t nnew = nhub + st.qout;
    const bool deferred = defer_depths(st, b, g.V);  // this level's, before the state moves on
    __syncthreads();

    // Hub degree prefix of the next frontier, read by the next top-down
    // level: rounds of kBlock consecutive hubs, one per thread, each an
    // exclusive block scan added to the previous rounds' total (a level
    // with few hubs takes one round of two dependent loads).
    {
        const uint32_t nq = (d & 1) ^ 1;
        // The degrees of the first kPreRounds rounds are loaded before any
        // scan, so those rounds wait on memory once.
        constexpr int kPreRounds = 16;
        uint32_t ids[kPreRounds];
        uint64_t degs[kPreRounds];
#pragma unroll
        for (int r = 0; r < kPreRounds; r++) {
            const uint32_t h = r * kBlock + tid;
            ids[r] = h < nhub ? b.hubq[nq][h] : 0u;
        }
#pragma unroll
        for (int r = 0; r < kPreRounds; r++)
            degs[r] = r * kBlock + tid < nhub ? out_degree(g, ids[r]) : 0;
        uint64_t carry = 0;
        auto round = [&](uint32_t r0, uint64_t deg) {
            const uint32_t h = r0 + tid;
            uint64_t incl = deg;
            for (int o = 1; o < kWave; o <<= 1) {
                const uint64_t y = __shfl_up(incl, static_cast<unsigned>(o), kWave);
                if (lane >= o) incl += y;
            }
            if (lane == kWave - 1) hscan[wid] = incl;
            __syncthreads();
            uint64_t before = carry, total = carry;
            for (int w = 0; w < kBlock / kWave; w++) {
                if (w < wid) before += hscan[w];
                total += hscan[w];
            }
            if (h < nhub) b.hub_prefix[h] = before + incl - deg;
            carry = total;
            __syncthreads();  // hscan is rewritten by the next round
        };
#pragma unroll
        for (int r = 0; r < kPreRounds; r++)
            if (r * kBlock < nhub) round(r * kBlock, degs[r]);
        for (uint32_t r0 = kPreRounds * kBlock; r0 < nhub; r0 += kBlock) {
            const uint32_t h = r0 + tid;
            round(r0, h < nhub ? out_degree(g, b.hubq[nq][h]) : 0);
        }
        if (tid == 0) b.hub_prefix[nhub] = carry;
    }

    if (tid == 0) {
        uint64_t a = 0;
        double q = 0.0;
        for (int w = 0; w < kBlock / kWave; w++) {
            a += rout[w];
            q += rsq[w];
        }
        // Level time: from the start of the step kernel's first workgroup
        // to the start of this epilogue.
        const float ns = static_cast<float>(now - st.t_step) * b.ns_per_tick;
        st.t_step = 0;  // read above by every thread, before the barriers
        if (d < b.max_levels) {
            b.counts[d] = nnew;
            b.times[d] = static_cast<uint64_t>(ns);
            b.choices[d] = st.choice;
        }
        st.num_levels = d + 1;
        if (nnew == 0 || d + 1 >= b.max_levels) {
            st.done = 1;
        } else {
            st.prev_choice = st.choice;
            st.prev_time_ns = ns;
            st.prev_frontier = static_cast<uint64_t>(st.qcount) + st.hubcount;
            st.depth = d + 1;
            st.qcount = st.qout;
            st.hubcount = nhub;
            st.qout = 0;
            st.hubout = 0;
            st.visited += nnew;
            st.explored_edges += a;
            st.frontier_edges = a;
            st.frontier_sq = q;
            st.fill_levels = deferred ? 1u : 0u;
        }
    }
    if (wid == 0) choose_next(p, st, g, lm);
}

// Level 0: the roots (roots[0..nroots), or just `root` when roots is NULL)
// are the whole frontier; a repeated root counts once.
__global__ __launch_bounds__(kBlock) void bfs_engine_init(Graph g, Buffers b, PolicyArgs p, uint32_t root,
                                                          const uint32_t *roots, uint32_t nroots)
{
    BfsState &st = *b.st;
    LaneModel lm;
    if (threadIdx.x < kWave) load_lane_model(p, lm);
    if (threadIdx.x == 0) {
        uint32_t nq = 0, nh = 0;
        uint64_t visited = 0, edges = 0;
        double sq = 0.0;
        b.hub_prefix[0] = 0;
        for (uint32_t i = 0; i < (roots ? nroots : 1u); i++) {
            const uint32_t r = roots ? roots[i] : root;
            if (b.levels[r] == 0) continue;
            const uint32_t dr = out_degree(g, r);
            b.levels[r] = 0;
            b.bits[0][r >> 5] |= 1u << (r & 31);
            b.vis[r >> 5] |= 1u << (r & 31);
            visited++;
            edges += dr;
            sq += static_cast<double>(dr) * dr;
            if (dr > b.hub_degree && nh < b.hub_cap) {
                b.hubq[0][nh] = r;
                b.hub_prefix[nh + 1] = b.hub_prefix[nh] + dr;
                nh++;
            } else {
                b.q[0][nq++] = r;
            }
        }
        st.depth = 0;
        st.done = 0;
        st.qout = 0;
        st.hubout = 0;
        st.num_levels = 0;
        st.visited = visited;
        st.explored_edges = edges;
        st.prev_frontier = 0;
        st.frontier_edges = edges;
        st.frontier_sq = sq;
        st.prev_choice = -1;
        st.prev_time_ns = 0.0f;
        st.t_step = 0;
        st.stalled = 0;
        st.fill_levels = 0;
        st.qcount = nq;
        st.hubcount = nh;
        st.choice = p.fixed;
    }
    if (threadIdx.x < kWave) choose_next(p, st, g, lm);
}

__global__ void set_choice(BfsState *st, int32_t c) { st->choice = c; }
__global__ void copy_state(BfsState *dst, const BfsState *src) { *dst = *src; }

// Likely parent of each vertex: among its first kParentScan in-edges, the
// in-neighbour of largest out-degree (the first one on ties), which a
// bottom-up level probes before the in-edges.  Only an order of probes:
// a vertex is discovered iff some in-neighbour is in the frontier.
__global__ __launch_bounds__(kBlock) void likely_parent_kernel(Graph g, uint32_t *parent)
{
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; v < g.V;
         v += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint32_t s = g.roff[v], e = min(g.roff[v + 1], s + kParentScan);
        uint32_t best = kNoParent, bestDeg = 0;
        for (uint32_t j = s; j < e; j++) {
            const uint32_t u = g.rcol[j];
            const uint32_t du = out_degree(g, u);
            if (best == kNoParent || du > bestDeg) {
                best = u;
                bestDeg = du;
            }
        }
        parent[v] = best;
    }
}

// Bitmap of the vertices without in-edges: no search can reach them except
// as a root, so each search starts with them marked visited and bottom-up
// levels never scan them.  A wave covers 64 consecutive vertices, i.e. two
// whole words.
__global__ __launch_bounds__(kBlock) void no_in_edges_kernel(Graph g, uint32_t *bits)
{
    const int lane = __lane_id();
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
         v - lane < g.V; v += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const bool none = v < g.V && g.roff[v] == g.roff[v + 1];
        const uint64_t m = __ballot(none);
        const uint64_t w = (v - lane) >> 5;
        if (lane == 0) bits[w] = static_cast<uint32_t>(m);
        if (lane == 32 && ((v - lane) + 32) < g.V) bits[w + 1] = static_cast<uint32_t>(m >> 32);
    }
}

}  // namespace
}  // namespace bw

using namespace bw;

struct bw_bfs_engine {
    Graph g{};
    Buffers b{};
    uint32_t grid = 0;
    int32_t *own_levels = nullptr;
    // Snapshot of levels, queues, bitmaps and state, for timing every
    // variant on the same level.
    int32_t *snap_levels = nullptr;
    uint32_t *snap_q = nullptr, *snap_hub = nullptr, *snap_bits = nullptr;
    uint64_t *snap_prefix = nullptr;
    BfsState *snap_st = nullptr;
    float *mlp = nullptr;
    int32_t mlp_out = 0;
    int32_t *schedule = nullptr;
    int32_t schedule_cap = 0;
    uint32_t schedule_kinds = 1;  // level kernels the schedule needs (see policy_kinds)
    uint32_t *roots = nullptr;
    uint32_t roots_cap = 0;
    uint32_t *no_in = nullptr;  // vertices without in-edges (bitmap)
    uint32_t *parent = nullptr;  // likely parent per vertex
    hipStream_t own_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t exec_key[4] = {-1, -1, -1, -1};  // policy, fixed, schedule_len, levels per batch
    void *exec_stream = nullptr;

    void drop_exec()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }

    ~bw_bfs_engine()
    {
        drop_exec();
        void *bufs[] = {own_levels, b.q[0], b.q[1], b.hubq[0], b.hubq[1], b.hub_prefix, b.bits[0], b.bits[1],
                        b.bits[2], b.vis, b.st, b.part, b.counts, b.times, b.choices, snap_levels, snap_q, snap_hub,
                        snap_bits, snap_prefix, snap_st, mlp, schedule, roots, no_in, parent};
        for (void *p : bufs)
            if (p) (void)hipFree(p);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

namespace {

PolicyArgs make_policy(const bw_bfs_engine *e, int32_t policy, int32_t fixed, int32_t schedule_len)
{
    PolicyArgs p;
    p.policy = policy;
    p.fixed = fixed;
    p.schedule = e->schedule;
    p.schedule_len = std::max(schedule_len, 1);
    p.mlp = e->mlp;
    p.nout = e->mlp_out;
    return p;
}

int check_variant(const bw_bfs_engine *e, int32_t v)
{
    if (v < 0 || v >= kVariants) return fail(BW_HIP_ERR_ARG, "unknown BFS engine variant " + std::to_string(v));
    if (!variant_available(e->g, v))
        return fail(BW_HIP_ERR_ARG, "BFS engine variant " + std::to_string(v) +
                                        " needs an edge-list representation the engine was not given");
    return BW_HIP_OK;
}

int check_policy(const bw_bfs_engine *e, int32_t policy, int32_t fixed, int32_t schedule_len)
{
    BW_TRY(check_variant(e, fixed));
    if (policy < FIXED || policy > BEAMER) return fail(BW_HIP_ERR_ARG, "unknown policy " + std::to_string(policy));
    if (policy == LEARNED && !e->mlp) return fail(BW_HIP_ERR_ARG, "no model set for the learned policy");
    if (policy == SCHEDULE && (!e->schedule || schedule_len < 1 || schedule_len > e->schedule_cap))
        return fail(BW_HIP_ERR_ARG, "no schedule of that length set");
    return BW_HIP_OK;
}

// Which level kernels a policy may need: bit 0 the native one, bit 1 the
// reference families' (a level launches each it may need; the one whose
// kind the level's variant is not returns at once).
uint32_t policy_kinds(const bw_bfs_engine *e, const PolicyArgs &p)
{
    switch (p.policy) {
    case FIXED: return p.fixed < kRef ? 1u : 2u;
    case SCHEDULE: return e->schedule_kinds;
    case LEARNED: return p.nout > kRef ? 3u : 1u;
    case BEAMER: return 1u;
    default: return 3u;  // HOST: the variant is set on the device
    }
}

int launch_step(bw_bfs_engine *e, bool ref, hipStream_t s)
{
    const size_t lds = sizeof(uint32_t) * e->b.lds_queue;
    if (ref) hipLaunchKernelGGL(bfs_engine_step<true>, dim3(e->grid), dim3(kBlock), lds, s, e->g, e->b);
    else hipLaunchKernelGGL(bfs_engine_step<false>, dim3(e->grid), dim3(kBlock), lds, s, e->g, e->b);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

int launch_level(bw_bfs_engine *e, const PolicyArgs &p, hipStream_t s, uint32_t kinds)
{
    if (kinds & 1u) BW_TRY(launch_step(e, false, s));
    if (kinds & 2u) BW_TRY(launch_step(e, true, s));
    hipLaunchKernelGGL(bfs_engine_epilogue, dim3(1), dim3(kBlock), 0, s, e->g, e->b, p, e->grid);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

}  // namespace

extern "C" int bw_bfs_engine_create(bw_bfs_engine **out, const bw_graph_dev *g, int32_t *levels, uint32_t max_levels)
{
    if (!out || !g || max_levels == 0) return fail(BW_HIP_ERR_ARG, "bw_bfs_engine_create: invalid argument");
    *out = nullptr;
    if (g->offset_bytes != 4 || !g->offsets || !g->rev_offsets || (g->edge_count && (!g->cols || !g->rev_cols)))
        return fail(BW_HIP_ERR_ARG, "the BFS engine needs both CSR directions with uint32 offsets");
    if (g->vertex_count == 0 || g->vertex_count >= 0x7FFFFFFFull)
        return fail(BW_HIP_ERR_ARG, "the BFS engine needs 0 < V < 2^31");
    int dev = 0, rate_khz = 0;
    BW_HIP_CHECK(hipGetDevice(&dev));
    BW_HIP_CHECK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
    if (rate_khz <= 0) return fail(BW_HIP_ERR_RUNTIME, "device reports no wall clock rate");
    auto e = new bw_bfs_engine();
    const uint64_t V = g->vertex_count;
    e->g = Graph{V,       g->edge_count, static_cast<const uint32_t *>(g->offsets), g->cols,
                 static_cast<const uint32_t *>(g->rev_offsets), g->rev_cols, g->el_in, g->el_out, g->rel_in,
                 g->rel_out, g->sel, g->rsel};
    e->grid = static_cast<uint32_t>(device_cus()) * 8;
    const uint32_t words = static_cast<uint32_t>((V + 31) / 32);
    bool ok = true;
    auto alloc = [&](auto **p, size_t bytes) {
        if (ok && hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(bytes, 8)) != hipSuccess) ok = false;
    };
    if (!levels) {
        alloc(&e->own_levels, sizeof(int32_t) * V);
        levels = e->own_levels;
    }
    e->b.levels = levels;
    alloc(&e->b.q[0], sizeof(uint32_t) * V);
    alloc(&e->b.q[1], sizeof(uint32_t) * V);
    alloc(&e->b.hubq[0], sizeof(uint32_t) * kHubCapMax);
    alloc(&e->b.hubq[1], sizeof(uint32_t) * kHubCapMax);
    alloc(&e->b.hub_prefix, sizeof(uint64_t) * (kHubCapMax + 1));
    for (auto &bm : e->b.bits) alloc(&bm, sizeof(uint32_t) * words);
    alloc(&e->b.vis, sizeof(uint32_t) * words);
    alloc(&e->no_in, sizeof(uint32_t) * words);
    alloc(&e->parent, sizeof(uint32_t) * V);
    alloc(&e->b.st, sizeof(BfsState));
    alloc(&e->b.part, sizeof(uint64_t) * 2 * e->grid);
    alloc(&e->b.counts, sizeof(uint32_t) * max_levels);
    alloc(&e->b.times, sizeof(uint64_t) * max_levels);
    alloc(&e->b.choices, sizeof(int32_t) * max_levels);
    e->b.max_levels = max_levels;
    e->b.words = words;
    e->b.ns_per_tick = 1e6f / static_cast<float>(rate_khz);
    e->b.hub_degree = BW_BFS_DEFAULT_HUB_DEGREE;
    e->b.wave_degree = BW_BFS_DEFAULT_WAVE_DEGREE;
    e->b.lds_queue = BW_BFS_DEFAULT_LDS_QUEUE;
    e->b.hub_cap = BW_BFS_DEFAULT_HUB_CAP;
    if (!ok) {
        delete e;
        return fail(BW_HIP_ERR_NOMEM, "device allocation of the BFS engine failed");
    }
    e->b.parent = e->parent;
    hipLaunchKernelGGL(no_in_edges_kernel, dim3(e->grid), dim3(kBlock), 0, nullptr, e->g, e->no_in);
    hipLaunchKernelGGL(likely_parent_kernel, dim3(e->grid), dim3(kBlock), 0, nullptr, e->g, e->parent);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        delete e;
        return fail(BW_HIP_ERR_RUNTIME, "building the in-edge bitmap failed");
    }
    *out = e;
    return BW_HIP_OK;
}

extern "C" int bw_bfs_engine_destroy(bw_bfs_engine *e)
{
    delete e;
    return BW_HIP_OK;
}

extern "C" int bw_bfs_engine_levels(const bw_bfs_engine *e, int32_t **levels)
{
    if (!e || !levels) return fail(BW_HIP_ERR_ARG, "invalid argument");
    *levels = e->b.levels;
    return BW_HIP_OK;
}

extern "C" int32_t bw_bfs_engine_mlp_len(int32_t outputs) { return mlp_len(outputs); }

extern "C" int bw_bfs_engine_set_model(bw_bfs_engine *e, const float *weights, int32_t outputs)
{
    if (!e || !weights || (outputs != BW_BFS_NATIVE_VARIANTS && outputs != BW_BFS_ENGINE_VARIANTS))
        return fail(BW_HIP_ERR_ARG, "bw_bfs_engine_set_model: the model needs one output per native variant, or "
                                    "per engine variant");
    if (!e->mlp) BW_HIP_CHECK(hipMalloc(&e->mlp, sizeof(float) * mlp_len(kMaxOut)));
    if (outputs != e->mlp_out) e->drop_exec();  // the captured batch holds the output count
    BW_HIP_CHECK(hipMemcpy(e->mlp, weights, sizeof(float) * mlp_len(outputs), hipMemcpyHostToDevice));
    e->mlp_out = outputs;
    return BW_HIP_OK;
}

extern "C" int bw_bfs_engine_set_schedule(bw_bfs_engine *e, const int32_t *schedule, int32_t len)
{
    if (!e || !schedule || len < 1) return fail(BW_HIP_ERR_ARG, "invalid schedule");
    uint32_t kinds = 0;
    for (int32_t i = 0; i < len; i++) {
        BW_TRY(check_variant(e, schedule[i]));
        kinds |= schedule[i] < kRef ? 1u : 2u;
    }
    if (kinds != e->schedule_kinds) e->drop_exec();  // the captured batch launches per the kinds
    e->schedule_kinds = kinds;
    if (len > e->schedule_cap) {
        e->drop_exec();  // the captured batch holds the old schedule pointer
        if (e->schedule) (void)hipFree(e->schedule);
        e->schedule = nullptr;
        e->schedule_cap = 0;
        BW_HIP_CHECK(hipMalloc(&e->schedule, sizeof(int32_t) * len));
        e->schedule_cap = len;
    }
    BW_HIP_CHECK(hipMemcpy(e->schedule, schedule, sizeof(int32_t) * len, hipMemcpyHostToDevice));
    return BW_HIP_OK;
}

extern "C" int bw_bfs_engine_set_tuning(bw_bfs_engine *e, uint32_t hub_degree, uint32_t wave_degree)
{
    if (!e) return fail(BW_HIP_ERR_ARG, "engine is NULL");
    if (hub_degree < 256) return fail(BW_HIP_ERR_ARG, "hub_degree must be at least 256");
    e->drop_exec();  // the captured batch holds the old thresholds
    e->b.hub_degree = hub_degree;
    e->b.wave_degree = wave_degree;
    return BW_HIP_OK;
}

extern "C" int bw_bfs_engine_set_queues(bw_bfs_engine *e, uint32_t lds_queue, uint32_t hub_cap)
{
    if (!e) return fail(BW_HIP_ERR_ARG, "engine is NULL");
    if (lds_queue < kBlock || lds_queue > kLdsQueueMax || lds_queue % kWave)
        return fail(BW_HIP_ERR_ARG, "lds_queue must be a multiple of 64 in [256, " + std::to_string(kLdsQueueMax) +
                                        "]");
    if (hub_cap < kBlock || hub_cap > kHubCapMax || hub_cap % kBlock)
        return fail(BW_HIP_ERR_ARG, "hub_cap must be a multiple of 256 in [256, " + std::to_string(kHubCapMax) + "]");
    e->drop_exec();  // the captured batch holds the old sizes
    e->b.lds_queue = lds_queue;
    e->b.hub_cap = hub_cap;
    return BW_HIP_OK;
}

namespace {

// Resets levels, bitmaps and state; roots_dev (device, n entries) or the
// single `root` form the depth-0 frontier.
int begin_impl(bw_bfs_engine *e, uint32_t root, const uint32_t *roots_dev, uint32_t nroots, int32_t policy,
               int32_t fixed, int32_t schedule_len, hipStream_t s)
{
    BW_TRY(check_policy(e, policy, fixed, schedule_len));
    BW_TRY(bw_hip_fill_u32(e->b.levels, static_cast<uint32_t>(kUnvisited), e->g.V, s));
    for (uint32_t *bm : e->b.bits) BW_TRY(bw_hip_fill_u32(bm, 0u, e->b.words, s));
    BW_HIP_CHECK(hipMemcpyAsync(e->b.vis, e->no_in, sizeof(uint32_t) * e->b.words, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(bfs_engine_init, dim3(1), dim3(kBlock), 0, s, e->g, e->b,
                       make_policy(e, policy, fixed, schedule_len), root, roots_dev, nroots);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

// Copies host roots into the engine's device buffer (checked in range).
int upload_roots(bw_bfs_engine *e, const uint32_t *roots, uint32_t n, hipStream_t s)
{
    if (!roots || n == 0 || n > BW_BFS_MAX_ROOTS)
        return fail(BW_HIP_ERR_ARG, "roots must hold 1.." + std::to_string(BW_BFS_MAX_ROOTS) + " vertices");
    for (uint32_t i = 0; i < n; i++)
        if (roots[i] >= e->g.V) return fail(BW_HIP_ERR_ARG, "root " + std::to_string(roots[i]) + " out of range");
    if (n > e->roots_cap) {
        if (e->roots) (void)hipFree(e->roots);
        e->roots = nullptr;
        e->roots_cap = 0;
        BW_HIP_CHECK(hipMalloc(&e->roots, sizeof(uint32_t) * n));
        e->roots_cap = n;
    }
    BW_HIP_CHECK(hipMemcpyAsync(e->roots, roots, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    BW_HIP_CHECK(hipStreamSynchronize(s));
    return BW_HIP_OK;
}

}  // namespace

// Resets levels, bitmaps and state for a BFS from `root`; the level-0
// variant is chosen by the policy.
extern "C" int bw_bfs_engine_begin(bw_bfs_engine *e, uint32_t root, int32_t policy, int32_t fixed,
                                   int32_t schedule_len, void *stream)
{
    if (!e) return fail(BW_HIP_ERR_ARG, "engine is NULL");
    if (root >= e->g.V) return fail(BW_HIP_ERR_ARG, "root out of range");
    return begin_impl(e, root, nullptr, 0, policy, fixed, schedule_len, static_cast<hipStream_t>(stream));
}

extern "C" int bw_bfs_engine_begin_multi(bw_bfs_engine *e, const uint32_t *roots_host, uint32_t nroots,
                                         int32_t policy, int32_t fixed, int32_t schedule_len, void *stream)
{
    if (!e) return fail(BW_HIP_ERR_ARG, "engine is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    BW_TRY(upload_roots(e, roots_host, nroots, s));
    return begin_impl(e, 0, e->roots, nroots, policy, fixed, schedule_len, s);
}

// One level (step kernel + epilogue) with the given policy choosing the
// variant of the level after it.
extern "C" int bw_bfs_engine_level(bw_bfs_engine *e, int32_t policy, int32_t fixed, int32_t schedule_len,
                                   void *stream)
{
    if (!e) return fail(BW_HIP_ERR_ARG, "engine is NULL");
    BW_TRY(check_policy(e, policy, fixed, schedule_len));
    const PolicyArgs p = make_policy(e, policy, fixed, schedule_len);
    return launch_level(e, p, static_cast<hipStream_t>(stream), policy_kinds(e, p));
}

// Overrides the variant of the pending level.
extern "C" int bw_bfs_engine_force(bw_bfs_engine *e, int32_t variant, void *stream)
{
    if (!e) return fail(BW_HIP_ERR_ARG, "engine is NULL");
    BW_TRY(check_variant(e, variant));
    hipLaunchKernelGGL(set_choice, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), e->b.st, variant);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

// Host view of the state (synchronises the stream): done flag, depth,
// levels finished, the pending variant and the 8 features of the pending
// level.
extern "C" int bw_bfs_engine_status(bw_bfs_engine *e, uint32_t *done, uint32_t *depth, uint32_t *num_levels,
                                    int32_t *choice, float *feat, void *stream)
{
    if (!e) return fail(BW_HIP_ERR_ARG, "engine is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    BfsState st;
    BW_HIP_CHECK(hipMemcpyAsync(&st, e->b.st, sizeof st, hipMemcpyDeviceToHost, s));
    BW_HIP_CHECK(hipStreamSynchronize(s));
    if (done) *done = st.done;
    if (depth) *depth = st.depth;
    if (num_levels) *num_levels = st.num_levels;
    if (choice) *choice = st.choice;
    if (feat) features(st, e->g.V, e->g.E, feat);
    return BW_HIP_OK;
}

// restore == 0 saves levels, frontier queues, bitmaps and state; restore
// != 0 puts the saved copy back.  Synchronises the stream.
extern "C" int bw_bfs_engine_snapshot(bw_bfs_engine *e, int32_t restore, void *stream)
{
    if (!e) return fail(BW_HIP_ERR_ARG, "engine is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t V = e->g.V;
    const size_t words = e->b.words;
    if (!e->snap_levels) {
        if (restore) return fail(BW_HIP_ERR_ARG, "no snapshot to restore");
        BW_HIP_CHECK(hipMalloc(&e->snap_levels, sizeof(int32_t) * V));
        BW_HIP_CHECK(hipMalloc(&e->snap_q, sizeof(uint32_t) * V));
        BW_HIP_CHECK(hipMalloc(&e->snap_hub, sizeof(uint32_t) * kHubCapMax));
        BW_HIP_CHECK(hipMalloc(&e->snap_bits, sizeof(uint32_t) * 4 * words));
        BW_HIP_CHECK(hipMalloc(&e->snap_prefix, sizeof(uint64_t) * (kHubCapMax + 1)));
        BW_HIP_CHECK(hipMalloc(&e->snap_st, sizeof(BfsState)));
    }
    BfsState st;
    BW_HIP_CHECK(hipMemcpyAsync(&st, restore ? e->snap_st : e->b.st, sizeof st, hipMemcpyDeviceToHost, s));
    BW_HIP_CHECK(hipStreamSynchronize(s));
    const uint32_t cur = st.depth & 1;
    const size_t nq = std::max<size_t>(st.qcount, 1), nh = std::max<size_t>(std::min(st.hubcount, e->b.hub_cap), 1);
    auto copy = [&](void *saved, void *live, size_t bytes) {
        return restore ? hipMemcpyAsync(live, saved, bytes, hipMemcpyDeviceToDevice, s)
                       : hipMemcpyAsync(saved, live, bytes, hipMemcpyDeviceToDevice, s);
    };
    BW_HIP_CHECK(copy(e->snap_levels, e->b.levels, sizeof(int32_t) * V));
    BW_HIP_CHECK(copy(e->snap_q, e->b.q[cur], sizeof(uint32_t) * nq));
    BW_HIP_CHECK(copy(e->snap_hub, e->b.hubq[cur], sizeof(uint32_t) * nh));
    BW_HIP_CHECK(copy(e->snap_prefix, e->b.hub_prefix, sizeof(uint64_t) * (e->b.hub_cap + 1)));
    for (int i = 0; i < 3; i++) BW_HIP_CHECK(copy(e->snap_bits + i * words, e->b.bits[i], sizeof(uint32_t) * words));
    BW_HIP_CHECK(copy(e->snap_bits + 3 * words, e->b.vis, sizeof(uint32_t) * words));
    if (restore) hipLaunchKernelGGL(copy_state, dim3(1), dim3(1), 0, s, e->b.st, e->snap_st);
    else hipLaunchKernelGGL(copy_state, dim3(1), dim3(1), 0, s, e->snap_st, e->b.st);
    BW_LAUNCH_CHECK();
    BW_HIP_CHECK(hipStreamSynchronize(s));
    return BW_HIP_OK;
}

// Host-driven level with a forced variant, timed with events around the
// step kernel; for exhaustive per-level timing.  Synchronises the stream.
extern "C" int bw_bfs_engine_step_forced(bw_bfs_engine *e, int32_t variant, float *ms, uint32_t *count,
                                         uint32_t *done, void *stream)
{
    if (!e) return fail(BW_HIP_ERR_ARG, "engine is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    BfsState before;
    BW_HIP_CHECK(hipMemcpyAsync(&before, e->b.st, sizeof before, hipMemcpyDeviceToHost, s));
    BW_HIP_CHECK(hipStreamSynchronize(s));
    if (before.done) return fail(BW_HIP_ERR_ARG, "the search is finished");
    BW_TRY(bw_bfs_engine_force(e, variant, stream));
    hipEvent_t ev0, ev1;
    BW_HIP_CHECK(hipEventCreate(&ev0));
    BW_HIP_CHECK(hipEventCreate(&ev1));
    int st = BW_HIP_OK;
    if (hipEventRecord(ev0, s) != hipSuccess) st = fail(BW_HIP_ERR_RUNTIME, "event record failed");
    if (st == BW_HIP_OK) st = launch_step(e, variant >= kRef, s);
    if (st == BW_HIP_OK && hipEventRecord(ev1, s) != hipSuccess) st = fail(BW_HIP_ERR_RUNTIME, "event record failed");
    if (st == BW_HIP_OK) {
        hipLaunchKernelGGL(bfs_engine_epilogue, dim3(1), dim3(kBlock), 0, s, e->g, e->b,
                           make_policy(e, HOST, variant, 1), e->grid);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) st = hip_fail(le, "bfs_engine_epilogue", __FILE__, __LINE__);
    }
    float t = 0;
    if (st == BW_HIP_OK && (hipStreamSynchronize(s) != hipSuccess || hipEventElapsedTime(&t, ev0, ev1) != hipSuccess))
        st = fail(BW_HIP_ERR_RUNTIME, "timing the level failed");
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (st != BW_HIP_OK) return st;
    BfsState after;
    BW_HIP_CHECK(hipMemcpy(&after, e->b.st, sizeof after, hipMemcpyDeviceToHost));
    if (ms) *ms = t;
    if (count) BW_HIP_CHECK(hipMemcpy(count, e->b.counts + before.depth, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (done) *done = after.done;
    return BW_HIP_OK;
}

// The whole search on the device: levels are launched in batches of
// `levels_per_sync` (replayed as one captured hipGraph when use_graph != 0)
// and the host reads only the done flag between batches.  Per-level counts,
// device times and variants are copied out at the end when requested.
namespace {

int run_impl(bw_bfs_engine *e, uint32_t root, const uint32_t *roots_host, uint32_t nroots, int32_t policy,
             int32_t fixed, int32_t schedule_len, int32_t levels_per_sync, int32_t use_graph, uint32_t *counts_host,
             uint64_t *times_ns_host, int32_t *choices_host, uint32_t *num_levels, float *total_ms, void *stream)
{
    if (!e || !num_levels || levels_per_sync < 1) return fail(BW_HIP_ERR_ARG, "bw_bfs_engine_run: invalid argument");
    BW_TRY(check_policy(e, policy, fixed, schedule_len));
    if (!roots_host && root >= e->g.V) return fail(BW_HIP_ERR_ARG, "root out of range");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (use_graph && !s) {  // the null stream cannot be captured
        if (!e->own_stream) BW_HIP_CHECK(hipStreamCreateWithFlags(&e->own_stream, hipStreamNonBlocking));
        s = e->own_stream;
    }
    if (roots_host) BW_TRY(upload_roots(e, roots_host, nroots, s));
    const PolicyArgs p = make_policy(e, policy, fixed, schedule_len);
    // A learned policy over both kinds launches only the native level kernel
    // (the reference families' kernel would mostly return at once, a fixed
    // cost per level) until a level picks a reference family.  That level's
    // epilogue leaves it pending and flags the state as stalled, and from the
    // next batch on the search launches both kinds.
    const uint32_t all_kinds = policy_kinds(e, p);
    const uint32_t graph_kinds = (policy == LEARNED && all_kinds == 3u) ? 1u : all_kinds;
    uint32_t kinds = graph_kinds;
    if (use_graph) {
        const int32_t key[4] = {policy, fixed, std::max(schedule_len, 1), levels_per_sync};
        if (e->exec && (std::memcmp(key, e->exec_key, sizeof key) != 0 || e->exec_stream != s)) {
            (void)hipGraphExecDestroy(e->exec);
            e->exec = nullptr;
        }
        if (!e->exec) {
            // Kernels only: no memset nodes in the captured batch.
            hipGraph_t graph = nullptr;
            BW_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int cst = BW_HIP_OK;
            for (int32_t i = 0; cst == BW_HIP_OK && i < levels_per_sync; i++)
                cst = launch_level(e, p, s, graph_kinds);
            hipError_t ce = hipStreamEndCapture(s, &graph);
            if (cst != BW_HIP_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return cst;
            }
            if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture", __FILE__, __LINE__);
            hipError_t ie = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                e->exec = nullptr;
                return hip_fail(ie, "hipGraphInstantiate", __FILE__, __LINE__);
            }
            std::memcpy(e->exec_key, key, sizeof key);
            e->exec_stream = s;
        }
    }
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    BW_HIP_CHECK(hipEventCreate(&ev0));
    BW_HIP_CHECK(hipEventCreate(&ev1));
    int st = BW_HIP_OK;
    if (hipEventRecord(ev0, s) != hipSuccess) st = fail(BW_HIP_ERR_RUNTIME, "event record failed");
    if (st == BW_HIP_OK)
        st = begin_impl(e, root, roots_host ? e->roots : nullptr, nroots, policy, fixed, schedule_len, s);
    BfsState bs{};
    uint32_t batches = 0;
    // One more batch than the levels need: the one a stall leaves unfinished.
    const uint32_t max_batches = e->b.max_levels / static_cast<uint32_t>(levels_per_sync) + 3;
    while (st == BW_HIP_OK && !bs.done) {
        if (batches++ > max_batches) {
            st = fail(BW_HIP_ERR_LIMIT, "BFS engine exceeded its level bound");
            break;
        }
        if (use_graph && kinds == graph_kinds) {
            hipError_t le = hipGraphLaunch(e->exec, s);
            if (le != hipSuccess) st = hip_fail(le, "hipGraphLaunch", __FILE__, __LINE__);
        } else {
            for (int32_t i = 0; st == BW_HIP_OK && i < levels_per_sync; i++) st = launch_level(e, p, s, kinds);
        }
        if (st != BW_HIP_OK) break;
        if (hipMemcpyAsync(&bs, e->b.st, sizeof bs, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            st = fail(BW_HIP_ERR_RUNTIME, "reading the BFS state failed");
        else if (bs.stalled)
            kinds = all_kinds;
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
    BfsState hs;
    BW_HIP_CHECK(hipMemcpy(&hs, e->b.st, sizeof hs, hipMemcpyDeviceToHost));
    *num_levels = hs.num_levels;
    if (hs.num_levels >= e->b.max_levels) {
        uint32_t last = 0;
        BW_HIP_CHECK(hipMemcpy(&last, e->b.counts + e->b.max_levels - 1, sizeof last, hipMemcpyDeviceToHost));
        if (last != 0)
            return fail(BW_HIP_ERR_LIMIT, "BFS reached max_levels = " + std::to_string(e->b.max_levels) +
                                              " with a non-empty frontier");
    }
    const size_t n = std::min<uint32_t>(hs.num_levels, e->b.max_levels);
    if (counts_host && n) BW_HIP_CHECK(hipMemcpy(counts_host, e->b.counts, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (times_ns_host && n)
        BW_HIP_CHECK(hipMemcpy(times_ns_host, e->b.times, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    if (choices_host && n)
        BW_HIP_CHECK(hipMemcpy(choices_host, e->b.choices, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return BW_HIP_OK;
}

}  // namespace

extern "C" int bw_bfs_engine_run(bw_bfs_engine *e, uint32_t root, int32_t policy, int32_t fixed,
                                 int32_t schedule_len, int32_t levels_per_sync, int32_t use_graph,
                                 uint32_t *counts_host, uint64_t *times_ns_host, int32_t *choices_host,
                                 uint32_t *num_levels, float *total_ms, void *stream)
{
    return run_impl(e, root, nullptr, 0, policy, fixed, schedule_len, levels_per_sync, use_graph, counts_host,
                    times_ns_host, choices_host, num_levels, total_ms, stream);
}

extern "C" int bw_bfs_engine_run_multi(bw_bfs_engine *e, const uint32_t *roots_host, uint32_t nroots,
                                       int32_t policy, int32_t fixed, int32_t schedule_len, int32_t levels_per_sync,
                                       int32_t use_graph, uint32_t *counts_host, uint64_t *times_ns_host,
                                       int32_t *choices_host, uint32_t *num_levels, float *total_ms, void *stream)
{
    if (!roots_host) return fail(BW_HIP_ERR_ARG, "roots is NULL");
    return run_impl(e, 0, roots_host, nroots, policy, fixed, schedule_len, levels_per_sync, use_graph, counts_host,
                    times_ns_host, choices_host, num_levels, total_ms, stream);
}
