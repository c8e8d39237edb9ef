// BFS level kernels for gfx950 and their C ABI entry points.
//
// Each kernel family restates the semantics of one reference kernel
// (bfs/*.cu of merijn/Belewitte) -- same depth encoding (int32, INT_MAX =
// unvisited), same frontier count (newly discovered vertices) -- with a
// wave64 mapping: 16-byte vector loads on the edge streams, an atomic only
// when a plain load shows the neighbour still unvisited, wave64 ballots in
// place of __any_sync, and the four frontier reductions of bfs/bfs.hpp over
// 64-lane waves.

#include "common.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace bw {
namespace {

// ------------------------------------------------------------ edge lists

// One edge (in, out): forward tests levels[in] and relaxes out; reverse
// (REV) tests levels[out] and relaxes in (bfs/rev_edgelist.cu:14-18).
template<int R, bool REV>
__device__ __forceinline__ void visit_edge(uint32_t in, uint32_t out, int32_t *levels,
                                           int32_t depth, FrontierCount<R> &fc,
                                           uint32_t *frontier)
{
    const uint32_t test = REV ? out : in;
    const uint32_t relax = REV ? in : out;
    if (levels[test] == depth && relax_depth(levels, relax, depth + 1)) fc.update(frontier);
}

// SoA EdgeList, 4 edges per thread step through uint4 loads of in[]/out[]
// (the arrays come from hipMalloc or bw_build_edge_lists and are 16-byte
// aligned; the host falls back to the scalar tail kernel otherwise).
template<int R, bool REV>
__global__ __launch_bounds__(1024) void bfs_edge_list_kernel(
    uint64_t E, const uint32_t *__restrict__ in, const uint32_t *__restrict__ out,
    int32_t *levels, int32_t depth, uint32_t *frontier, int vec)
{
    FrontierCount<R> fc;
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t done = 0;
    if (vec) {
        const uint64_t E4 = E / 4;
        const uint4 *in4 = reinterpret_cast<const uint4 *>(in);
        const uint4 *out4 = reinterpret_cast<const uint4 *>(out);
        for (uint64_t i = gid; i < E4; i += stride) {
            const uint4 a = in4[i], b = out4[i];
            visit_edge<R, REV>(a.x, b.x, levels, depth, fc, frontier);
            visit_edge<R, REV>(a.y, b.y, levels, depth, fc, frontier);
            visit_edge<R, REV>(a.z, b.z, levels, depth, fc, frontier);
            visit_edge<R, REV>(a.w, b.w, levels, depth, fc, frontier);
        }
        done = E4 * 4;
    }
    for (uint64_t i = done + gid; i < E; i += stride)
        visit_edge<R, REV>(in[i], out[i], levels, depth, fc, frontier);
    fc.finalise(frontier);
}

// AoS StructEdgeList: (in, out) pairs, two edges per uint4.
template<int R, bool REV>
__global__ __launch_bounds__(1024) void bfs_struct_edge_list_kernel(
    uint64_t E, const uint32_t *__restrict__ edges, int32_t *levels, int32_t depth,
    uint32_t *frontier, int vec)
{
    FrontierCount<R> fc;
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t done = 0;
    if (vec) {
        const uint64_t E2 = E / 2;
        const uint4 *e4 = reinterpret_cast<const uint4 *>(edges);
        for (uint64_t i = gid; i < E2; i += stride) {
            const uint4 p = e4[i];
            visit_edge<R, REV>(p.x, p.y, levels, depth, fc, frontier);
            visit_edge<R, REV>(p.z, p.w, levels, depth, fc, frontier);
        }
        done = E2 * 2;
    }
    for (uint64_t i = done + gid; i < E; i += stride)
        visit_edge<R, REV>(edges[2 * i], edges[2 * i + 1], levels, depth, fc, frontier);
    fc.finalise(frontier);
}

// ---------------------------------------------------------- vertex forms

// Out-degree (in-degree for pull) above which a vertex's edges are taken
// by its whole wave instead of its own thread.
constexpr uint64_t kWaveDegree = 64;
constexpr int kPushBatch = 4;                 // edges per lane in flight in the push kernels
constexpr uint32_t kNoVertex = 0xFFFFFFFFu;   // an empty slot of such a batch

// The 64-bit offset of lane `l` broadcast to the wave.
__device__ __forceinline__ uint64_t read_lane_u64(uint64_t x, int l)
{
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(x), l);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(x >> 32), l);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// bfs/vertex_push.cu: a thread per vertex expands the out-edges of its
// depth-d vertex; a vertex with more than kWaveDegree out-edges is expanded
// by the whole wave (the lanes stride over its edges), one such vertex at a
// time in ballot order, so a hub no longer serialises one thread.  Every
// out-edge of every frontier vertex is relaxed once, as in the reference,
// so depths and counts are the same.
template<int R, typename Off>
__global__ __launch_bounds__(1024) void bfs_vertex_push_kernel(
    uint64_t V, const Off *__restrict__ off, const uint32_t *__restrict__ cols,
    int32_t *levels, int32_t depth, uint32_t *frontier)
{
    FrontierCount<R> fc;
    const int32_t nd = depth + 1;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    // The loop runs while any lane of the wave has a vertex, so the wave
    // stays converged for its ballots.
    const uint64_t first = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t waveFirst = first - lane;
    for (uint64_t base = waveFirst; base < V; base += stride) {
        const uint64_t v = base + lane;
        const bool active = v < V && levels[v] == depth;
        uint64_t s = 0, e = 0;
        if (active) {
            s = load_off(off, v);
            e = load_off(off, v + 1);
        }
        const bool big = active && e - s > kWaveDegree;
        // kPushBatch edges per lane in flight: their column loads are issued
        // before any depth is probed.  A vertex is counted by the atomic that
        // lowers it, whatever the order, so depths and counts are unchanged.
        for (uint64_t m = __ballot(big); m; m &= m - 1) {
            const int l = __builtin_ctzll(m);
            const uint64_t hs = read_lane_u64(s, l), he = read_lane_u64(e, l);
            for (uint64_t j0 = hs + lane; j0 < he; j0 += kWave * kPushBatch) {
                uint32_t w[kPushBatch];
#pragma unroll
                for (int u = 0; u < kPushBatch; u++) {
                    const uint64_t j = j0 + u * kWave;
                    w[u] = j < he ? cols[j] : kNoVertex;
                }
#pragma unroll
                for (int u = 0; u < kPushBatch; u++)
                    if (w[u] != kNoVertex && relax_depth(levels, w[u], nd)) fc.update(frontier);
            }
        }
        if (active && !big)
            for (uint64_t j0 = s; j0 < e; j0 += kPushBatch) {
                uint32_t w[kPushBatch];
#pragma unroll
                for (int u = 0; u < kPushBatch; u++) w[u] = j0 + u < e ? cols[j0 + u] : kNoVertex;
#pragma unroll
                for (int u = 0; u < kPushBatch; u++)
                    if (w[u] != kNoVertex && relax_depth(levels, w[u], nd)) fc.update(frontier);
            }
    }
    fc.finalise(frontier);
}

// bfs/vertex_pull.cu: a thread per unvisited vertex scans its in-edges
// and the first in-neighbour at depth d sets it to d+1 (a plain store:
// only this vertex's thread or wave writes levels[v]).  A vertex with more
// than kWaveDegree in-edges is probed by the whole wave, 64 in-edges per
// step, stopping at the first step where any lane finds one (a ballot);
// which in-neighbour is found does not change the depth or the count.
template<int R, typename Off>
__global__ __launch_bounds__(1024) void bfs_vertex_pull_kernel(
    uint64_t V, const Off *__restrict__ roff, const uint32_t *__restrict__ rcols,
    int32_t *levels, int32_t depth, uint32_t *frontier)
{
    FrontierCount<R> fc;
    const int32_t nd = depth + 1;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    const uint64_t waveFirst = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x - lane;
    for (uint64_t base = waveFirst; base < V; base += stride) {
        const uint64_t v = base + lane;
        const bool open = v < V && levels[v] > nd;
        uint64_t s = 0, e = 0;
        if (open) {
            s = load_off(roff, v);
            e = load_off(roff, v + 1);
        }
        const bool big = open && e - s > kWaveDegree;
        for (uint64_t m = __ballot(big); m; m &= m - 1) {
            const int l = __builtin_ctzll(m);
            const uint64_t hs = read_lane_u64(s, l), he = read_lane_u64(e, l);
            bool found = false;
            for (uint64_t j0 = hs; j0 < he && !found; j0 += kWave) {
                const uint64_t j = j0 + lane;
                found = __ballot(j < he && levels[rcols[j]] == depth) != 0;
            }
            if (found && lane == static_cast<uint32_t>(l)) {
                levels[v] = nd;
                fc.update(frontier);
            }
        }
        if (open && !big) {
            for (uint64_t j = s; j < e; j++) {
                if (levels[rcols[j]] == depth) {
                    levels[v] = nd;
                    fc.update(frontier);
                    break;
                }
            }
        }
    }
    fc.finalise(frontier);
}

// Dynamic LDS layout of the _warp kernels: per virtual warp `chunk` levels
// then `chunk + 1` offsets (the chunkMemory lambda of bfs/bfs.cpp:187-189),
// offsets 8-byte aligned.
template<typename Off>
__host__ __device__ inline uint32_t warp_lds_levels_bytes(uint32_t warps, uint32_t chunk)
{ return (warps * chunk * 4u + 7u) & ~7u; }

template<typename Off>
__host__ __device__ inline uint32_t warp_lds_bytes(uint32_t warps, uint32_t chunk)
{ return warp_lds_levels_bytes<Off>(warps, chunk) + warps * (chunk + 1) * static_cast<uint32_t>(sizeof(Off)); }

// bfs/vertex_push_warp.cu: a virtual warp of W lanes (W | 64) stages a
// chunk's levels and offsets in LDS (memcpy_SIMD) and strides over each
// depth-d vertex's neighbours.
template<int R, typename Off>
__global__ __launch_bounds__(1024) void bfs_vertex_push_warp_kernel(
    uint32_t W, uint32_t C, uint64_t V, const Off *__restrict__ off,
    const uint32_t *__restrict__ cols, int32_t *levels, int32_t depth, uint32_t *frontier)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    FrontierCount<R> fc;
    const int32_t nd = depth + 1;
    const uint32_t warps = blockDim.x / W;
    const uint32_t bw = threadIdx.x / W, lane = threadIdx.x % W;
    int32_t *myLevels = reinterpret_cast<int32_t *>(lds_raw) + bw * C;
    Off *myOff = reinterpret_cast<Off *>(lds_raw + warp_lds_levels_bytes<Off>(warps, C)) + bw * (C + 1);
    const uint64_t warpId = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / W;
    const uint64_t totalWarps = static_cast<uint64_t>(warps) * gridDim.x;

    for (uint64_t chunk = warpId; chunk * C < V; chunk += totalWarps) {
        const uint64_t base = chunk * C;
        const uint32_t end = static_cast<uint32_t>(std::min<uint64_t>(C, V - base));
        for (uint32_t i = lane; i < end; i += W) myLevels[i] = levels[base + i];
        for (uint32_t i = lane; i <= end; i += W) myOff[i] = off[base + i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint32_t v = 0; v < end; v++) {
            if (myLevels[v] != depth) continue;
            const uint64_t s = myOff[v], n = static_cast<uint64_t>(myOff[v + 1]) - s;
            for (uint64_t i = lane; i < n; i += W)
                if (relax_depth(levels, cols[s + i], nd)) fc.update(frontier);
        }
        __builtin_amdgcn_wave_barrier();
    }
    fc.finalise(frontier);
}

// bfs/vertex_pull_warp.cu: a virtual warp probes an unvisited vertex's
// in-neighbours; the lanes' hits are combined with a wave64 ballot masked to
// the virtual warp (the __any_sync of :14) and the probe stops at the first
// stride that finds one.  The chunk's new depths are counted and written
// back from LDS.
template<int R, typename Off>
__global__ __launch_bounds__(1024) void bfs_vertex_pull_warp_kernel(
    uint32_t W, uint32_t C, uint64_t V, const Off *__restrict__ roff,
    const uint32_t *__restrict__ rcols, int32_t *levels, int32_t depth, uint32_t *frontier)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    FrontierCount<R> fc;
    const int32_t nd = depth + 1;
    const uint32_t warps = blockDim.x / W;
    const uint32_t bw = threadIdx.x / W, lane = threadIdx.x % W;
    int32_t *myLevels = reinterpret_cast<int32_t *>(lds_raw) + bw * C;
    Off *myOff = reinterpret_cast<Off *>(lds_raw + warp_lds_levels_bytes<Off>(warps, C)) + bw * (C + 1);
    const uint64_t warpId = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / W;
    const uint64_t totalWarps = static_cast<uint64_t>(warps) * gridDim.x;
    const uint32_t waveLane = threadIdx.x & (kWave - 1);
    const uint32_t subBase = (waveLane / W) * W;
    const uint64_t subMask = (W >= 64) ? ~0ull : (((1ull << W) - 1ull) << subBase);

    for (uint64_t chunk = warpId; chunk * C < V; chunk += totalWarps) {
        const uint64_t base = chunk * C;
        const uint32_t end = static_cast<uint32_t>(std::min<uint64_t>(C, V - base));
        for (uint32_t i = lane; i < end; i += W) myLevels[i] = levels[base + i];
        for (uint32_t i = lane; i <= end; i += W) myOff[i] = roff[base + i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint32_t v = 0; v < end; v++) {
            if (myLevels[v] <= depth) continue;
            const uint64_t s = myOff[v], n = static_cast<uint64_t>(myOff[v + 1]) - s;
            bool found = false;
            for (uint64_t i0 = 0; i0 < n && !found; i0 += W) {
                const uint64_t i = i0 + lane;
                const bool hit = i < n && levels[rcols[s + i]] == depth;
                found = (__ballot(hit) & subMask) != 0;
            }
            if (found && lane == 0) myLevels[v] = nd;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        for (uint32_t i = lane; i < end; i += W) {
            if (myLevels[i] == nd) {
                fc.update(frontier);
                levels[base + i] = nd;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    fc.finalise(frontier);
}

// ------------------------------------------------------- host dispatch

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template<template<int> class Launch, typename... A>
int by_reduction(int32_t reduction, A &&...args)
{
    switch (reduction) {
    case BW_RED_NORMAL: return Launch<BW_RED_NORMAL>::run(args...);
    case BW_RED_BULK: return Launch<BW_RED_BULK>::run(args...);
    case BW_RED_WAVE: return Launch<BW_RED_WAVE>::run(args...);
    case BW_RED_BLOCK: return Launch<BW_RED_BLOCK>::run(args...);
    default: return fail(BW_HIP_ERR_ARG, "unknown frontier reduction " + std::to_string(reduction));
    }
}

int check_common(const int32_t *levels, const uint32_t *frontier)
{
    if (!levels || !frontier) return fail(BW_HIP_ERR_ARG, "levels and frontier must be device pointers");
    return BW_HIP_OK;
}

template<bool REV>
struct EdgeListLaunch {
    template<int R>
    struct L {
        static int run(uint64_t E, const uint32_t *in, const uint32_t *out, int32_t *levels,
                       int32_t depth, uint32_t *frontier, const bw_launch *cfg)
        {
            uint32_t block, grid;
            const int vec = aligned16(in) && aligned16(out);
            BW_TRY(resolve_geometry(cfg, vec ? (E + 3) / 4 : E, block, grid));
            hipLaunchKernelGGL((bfs_edge_list_kernel<R, REV>), dim3(grid), dim3(block), 0,
                               stream_of(cfg), E, in, out, levels, depth, frontier, vec);
            BW_LAUNCH_CHECK();
            return BW_HIP_OK;
        }
    };
};

template<bool REV>
struct StructLaunch {
    template<int R>
    struct L {
        static int run(uint64_t E, const uint32_t *edges, int32_t *levels, int32_t depth,
                       uint32_t *frontier, const bw_launch *cfg)
        {
            uint32_t block, grid;
            const int vec = aligned16(edges);
            BW_TRY(resolve_geometry(cfg, vec ? (E + 1) / 2 : E, block, grid));
            hipLaunchKernelGGL((bfs_struct_edge_list_kernel<R, REV>), dim3(grid), dim3(block), 0,
                               stream_of(cfg), E, edges, levels, depth, frontier, vec);
            BW_LAUNCH_CHECK();
            return BW_HIP_OK;
        }
    };
};

enum class VertexKind { Push, Pull, PushWarp, PullWarp };

template<VertexKind K, typename Off>
struct VertexLaunch {
    template<int R>
    struct L {
        static int run(uint64_t V, const Off *off, const uint32_t *cols, int32_t *levels,
                       int32_t depth, uint32_t *frontier, const bw_launch *cfg)
        {
            uint32_t block, grid;
            hipStream_t s = stream_of(cfg);
            if constexpr (K == VertexKind::Push || K == VertexKind::Pull) {
                BW_TRY(resolve_geometry(cfg, V, block, grid));
                if constexpr (K == VertexKind::Push)
                    hipLaunchKernelGGL((bfs_vertex_push_kernel<R, Off>), dim3(grid), dim3(block), 0, s,
                                       V, off, cols, levels, depth, frontier);
                else
                    hipLaunchKernelGGL((bfs_vertex_pull_kernel<R, Off>), dim3(grid), dim3(block), 0, s,
                                       V, off, cols, levels, depth, frontier);
            } else {
                const uint32_t W = (cfg && cfg->warp) ? cfg->warp : 32;
                const uint32_t C = (cfg && cfg->chunk) ? cfg->chunk : 32;
                if (W == 0 || W > 64 || (W & (W - 1)))
                    return fail(BW_HIP_ERR_ARG, "Warp size should be a power of 2 no larger than 64!");
                if (C % W != 0) return fail(BW_HIP_ERR_ARG, "Chunk size should be a multiple of warp size!");
                // A grid sized by chunks of vertices, one virtual warp each.
                const uint64_t chunks = (V + C - 1) / C;
                BW_TRY(resolve_geometry(cfg, chunks * W, block, grid));
                const uint32_t lds = warp_lds_bytes<Off>(block / W, C);
                if (lds > 64 * 1024)
                    return fail(BW_HIP_ERR_ARG, "Insufficient shared memory, " + std::to_string(lds)
                                                + " requested, 65536 available.");
                if constexpr (K == VertexKind::PushWarp)
                    hipLaunchKernelGGL((bfs_vertex_push_warp_kernel<R, Off>), dim3(grid), dim3(block), lds, s,
                                       W, C, V, off, cols, levels, depth, frontier);
                else
                    hipLaunchKernelGGL((bfs_vertex_pull_warp_kernel<R, Off>), dim3(grid), dim3(block), lds, s,
                                       W, C, V, off, cols, levels, depth, frontier);
            }
            BW_LAUNCH_CHECK();
            return BW_HIP_OK;
        }
    };
};

template<VertexKind K>
int vertex_entry(uint64_t V, const void *off, int32_t offset_bytes, const uint32_t *cols,
                 int32_t *levels, int32_t depth, uint32_t *frontier, int32_t reduction,
                 const bw_launch *cfg)
{
    BW_TRY(check_common(levels, frontier));
    if (V == 0) return BW_HIP_OK;
    if (!off || !cols) return fail(BW_HIP_ERR_ARG, "CSR arrays must be device pointers");
    if (offset_bytes == 4)
        return by_reduction<VertexLaunch<K, uint32_t>::template L>(
            reduction, V, static_cast<const uint32_t *>(off), cols, levels, depth, frontier, cfg);
    if (offset_bytes == 8)
        return by_reduction<VertexLaunch<K, uint64_t>::template L>(
            reduction, V, static_cast<const uint64_t *>(off), cols, levels, depth, frontier, cfg);
    return fail(BW_HIP_ERR_ARG, "offset_bytes must be 4 or 8");
}

// ----------------------------------------------------- edge-list builder

// Edge lists from a CSR, edge-parallel: a workgroup takes kListEdges
// consecutive edges, finds the rows of the first and last by binary search
// over the offsets, and each thread finds the row of each of its edges by
// binary search within that range (a few steps over cached offsets; none
// inside one long row).  Stores are coalesced across the workgroup.
constexpr uint64_t kListEdges = 4096;

// The row of edge e: the largest r in [lo, hi] with off[r] <= e.
template<typename Off>
__device__ __forceinline__ uint64_t row_of(const Off *off, uint64_t lo, uint64_t hi, uint64_t e)
{
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if (static_cast<uint64_t>(off[mid]) <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template<typename Off>
__global__ __launch_bounds__(256) void build_edge_lists_kernel(uint64_t V, uint64_t E, const Off *__restrict__ off,
                                                               const uint32_t *__restrict__ cols, uint32_t *in,
                                                               uint32_t *out, uint2 *pairs)
{
    __shared__ uint64_t range[2];
    for (uint64_t e0 = static_cast<uint64_t>(blockIdx.x) * kListEdges; e0 < E;
         e0 += static_cast<uint64_t>(gridDim.x) * kListEdges) {
        const uint64_t e1 = e0 + kListEdges < E ? e0 + kListEdges : E;
        __syncthreads();  // the previous step has read the range
        if (threadIdx.x == 0) {
            const uint64_t r0 = row_of(off, 0, V - 1, e0);
            range[0] = r0;
            range[1] = row_of(off, r0, V - 1, e1 - 1);
        }
        __syncthreads();
        const uint64_t r0 = range[0], r1 = range[1];
        for (uint64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
            const uint32_t r = static_cast<uint32_t>(row_of(off, r0, r1, e));
            const uint32_t c = cols[e];
            if (in) in[e] = r;
            if (out) out[e] = c;
            if (pairs) pairs[e] = make_uint2(r, c);
        }
    }
}

__global__ void set_root_kernel(int32_t *levels, uint32_t root) { levels[root] = 0; }

}  // namespace
}  // namespace bw

using namespace bw;

extern "C" int bw_bfs_edge_list(uint64_t E, const uint32_t *in, const uint32_t *out, int32_t *levels,
                                int32_t depth, uint32_t *frontier, int32_t reduction, const bw_launch *cfg)
{
    BW_TRY(check_common(levels, frontier));
    if (E == 0) return BW_HIP_OK;
    if (!in || !out) return fail(BW_HIP_ERR_ARG, "EdgeList arrays must be device pointers");
    return by_reduction<EdgeListLaunch<false>::L>(reduction, E, in, out, levels, depth, frontier, cfg);
}

extern "C" int bw_bfs_rev_edge_list(uint64_t E, const uint32_t *in, const uint32_t *out, int32_t *levels,
                                    int32_t depth, uint32_t *frontier, int32_t reduction, const bw_launch *cfg)
{
    BW_TRY(check_common(levels, frontier));
    if (E == 0) return BW_HIP_OK;
    if (!in || !out) return fail(BW_HIP_ERR_ARG, "EdgeList arrays must be device pointers");
    return by_reduction<EdgeListLaunch<true>::L>(reduction, E, in, out, levels, depth, frontier, cfg);
}

extern "C" int bw_bfs_struct_edge_list(uint64_t E, const uint32_t *edges, int32_t *levels, int32_t depth,
                                       uint32_t *frontier, int32_t reduction, const bw_launch *cfg)
{
    BW_TRY(check_common(levels, frontier));
    if (E == 0) return BW_HIP_OK;
    if (!edges) return fail(BW_HIP_ERR_ARG, "StructEdgeList must be a device pointer");
    return by_reduction<StructLaunch<false>::L>(reduction, E, edges, levels, depth, frontier, cfg);
}

extern "C" int bw_bfs_rev_struct_edge_list(uint64_t E, const uint32_t *edges, int32_t *levels,
                                           int32_t depth, uint32_t *frontier, int32_t reduction,
                                           const bw_launch *cfg)
{
    BW_TRY(check_common(levels, frontier));
    if (E == 0) return BW_HIP_OK;
    if (!edges) return fail(BW_HIP_ERR_ARG, "StructEdgeList must be a device pointer");
    return by_reduction<StructLaunch<true>::L>(reduction, E, edges, levels, depth, frontier, cfg);
}

extern "C" int bw_bfs_vertex_push(uint64_t V, const void *off, int32_t ob, const uint32_t *cols,
                                  int32_t *levels, int32_t depth, uint32_t *frontier, int32_t reduction,
                                  const bw_launch *cfg)
{ return vertex_entry<VertexKind::Push>(V, off, ob, cols, levels, depth, frontier, reduction, cfg); }

extern "C" int bw_bfs_vertex_pull(uint64_t V, const void *off, int32_t ob, const uint32_t *cols,
                                  int32_t *levels, int32_t depth, uint32_t *frontier, int32_t reduction,
                                  const bw_launch *cfg)
{ return vertex_entry<VertexKind::Pull>(V, off, ob, cols, levels, depth, frontier, reduction, cfg); }

extern "C" int bw_bfs_vertex_push_warp(uint64_t V, const void *off, int32_t ob, const uint32_t *cols,
                                       int32_t *levels, int32_t depth, uint32_t *frontier,
                                       int32_t reduction, const bw_launch *cfg)
{ return vertex_entry<VertexKind::PushWarp>(V, off, ob, cols, levels, depth, frontier, reduction, cfg); }

extern "C" int bw_bfs_vertex_pull_warp(uint64_t V, const void *off, int32_t ob, const uint32_t *cols,
                                       int32_t *levels, int32_t depth, uint32_t *frontier,
                                       int32_t reduction, const bw_launch *cfg)
{ return vertex_entry<VertexKind::PullWarp>(V, off, ob, cols, levels, depth, frontier, reduction, cfg); }

extern "C" int bw_bfs_level(const bw_graph_dev *g, int32_t variant, int32_t reduction, int32_t *levels,
                            int32_t depth, uint32_t *frontier, const bw_launch *cfg)
{
    if (!g) return fail(BW_HIP_ERR_ARG, "graph is NULL");
    auto need = [&](const void *p, const char *what) {
        if (!p && g->edge_count) {
            fail(BW_HIP_ERR_ARG, std::string("variant needs the ") + what + " representation");
            return false;
        }
        return true;
    };
    switch (variant) {
    case BW_BFS_EDGE_LIST:
        if (!need(g->el_in, "EdgeList") || !need(g->el_out, "EdgeList")) return BW_HIP_ERR_ARG;
        return bw_bfs_edge_list(g->edge_count, g->el_in, g->el_out, levels, depth, frontier, reduction, cfg);
    case BW_BFS_REV_EDGE_LIST:
        if (!need(g->rel_in, "reverse EdgeList") || !need(g->rel_out, "reverse EdgeList")) return BW_HIP_ERR_ARG;
        return bw_bfs_rev_edge_list(g->edge_count, g->rel_in, g->rel_out, levels, depth, frontier, reduction, cfg);
    case BW_BFS_STRUCT_EDGE_LIST:
        if (!need(g->sel, "StructEdgeList")) return BW_HIP_ERR_ARG;
        return bw_bfs_struct_edge_list(g->edge_count, g->sel, levels, depth, frontier, reduction, cfg);
    case BW_BFS_REV_STRUCT_EDGE_LIST:
        if (!need(g->rsel, "reverse StructEdgeList")) return BW_HIP_ERR_ARG;
        return bw_bfs_rev_struct_edge_list(g->edge_count, g->rsel, levels, depth, frontier, reduction, cfg);
    case BW_BFS_VERTEX_PUSH:
        return bw_bfs_vertex_push(g->vertex_count, g->offsets, g->offset_bytes, g->cols, levels, depth,
                                  frontier, reduction, cfg);
    case BW_BFS_VERTEX_PULL:
        return bw_bfs_vertex_pull(g->vertex_count, g->rev_offsets, g->offset_bytes, g->rev_cols, levels,
                                  depth, frontier, reduction, cfg);
    case BW_BFS_VERTEX_PUSH_WARP:
        return bw_bfs_vertex_push_warp(g->vertex_count, g->offsets, g->offset_bytes, g->cols, levels, depth,
                                       frontier, reduction, cfg);
    case BW_BFS_VERTEX_PULL_WARP:
        return bw_bfs_vertex_pull_warp(g->vertex_count, g->rev_offsets, g->offset_bytes, g->rev_cols,
                                       levels, depth, frontier, reduction, cfg);
    default:
        return fail(BW_HIP_ERR_ARG, "unknown BFS variant " + std::to_string(variant));
    }
}

extern "C" int bw_build_edge_lists(uint64_t V, uint64_t E, const void *off, int32_t ob, const uint32_t *cols,
                                   uint32_t *in, uint32_t *out, uint32_t *pairs, void *stream)
{
    if (V == 0 || E == 0) return BW_HIP_OK;
    if (!off || !cols) return fail(BW_HIP_ERR_ARG, "CSR arrays must be device pointers");
    if (reinterpret_cast<uintptr_t>(pairs) % 8) return fail(BW_HIP_ERR_ARG, "pairs must be 8-byte aligned");
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((E + kListEdges - 1) / kListEdges, 65536));
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint2 *p2 = reinterpret_cast<uint2 *>(pairs);  // device allocations are 8-byte aligned
    if (ob == 4)
        hipLaunchKernelGGL(build_edge_lists_kernel<uint32_t>, dim3(grid), dim3(256), 0, s, V, E,
                           static_cast<const uint32_t *>(off), cols, in, out, p2);
    else if (ob == 8)
        hipLaunchKernelGGL(build_edge_lists_kernel<uint64_t>, dim3(grid), dim3(256), 0, s, V, E,
                           static_cast<const uint64_t *>(off), cols, in, out, p2);
    else
        return fail(BW_HIP_ERR_ARG, "offset_bytes must be 4 or 8");
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_bfs_run(const bw_graph_dev *g, const int32_t *schedule, int32_t schedule_len,
                          int32_t reduction, uint32_t root, int32_t *levels, uint32_t *counts_host,
                          float *level_ms_host, int64_t max_levels, int64_t *num_levels,
                          const bw_launch *cfg)
{
    if (!g || !schedule || schedule_len <= 0 || !levels || !num_levels)
        return fail(BW_HIP_ERR_ARG, "bw_bfs_run: invalid argument");
    *num_levels = 0;
    if (root >= g->vertex_count) return BW_HIP_OK;  // bfs/bfs.cpp:53
    hipStream_t s = stream_of(cfg);

    uint32_t *frontier = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int status = BW_HIP_OK;
    auto cleanup = [&]() {
        if (frontier) (void)hipFree(frontier);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    auto check = [&](hipError_t e, const char *what) {
        if (e != hipSuccess) status = hip_fail(e, what, __FILE__, __LINE__);
        return e == hipSuccess;
    };
    if (!check(hipMalloc(&frontier, sizeof(uint32_t)), "hipMalloc") ||
        !check(hipEventCreate(&ev[0]), "hipEventCreate") || !check(hipEventCreate(&ev[1]), "hipEventCreate")) {
        cleanup();
        return status;
    }
    status = bw_hip_fill_u32(levels, static_cast<uint32_t>(kUnvisited), g->vertex_count, s);
    if (status == BW_HIP_OK) {
        hipLaunchKernelGGL(set_root_kernel, dim3(1), dim3(1), 0, s, levels, root);
        check(hipGetLastError(), "set_root_kernel");
    }
    int64_t curr = 0;
    uint32_t count = 1;
    while (status == BW_HIP_OK && count != 0) {
        if (curr >= max_levels) {
            status = fail(BW_HIP_ERR_LIMIT, "BFS did not finish within max_levels launches");
            break;
        }
        if ((status = bw_hip_fill_u32(frontier, 0, 1, s)) != BW_HIP_OK) break;
        if (!check(hipEventRecord(ev[0], s), "hipEventRecord")) break;
        const int32_t variant = schedule[curr % schedule_len];
        if ((status = bw_bfs_level(g, variant, reduction, levels, static_cast<int32_t>(curr), frontier, cfg)) !=
            BW_HIP_OK)
            break;
        if (!check(hipEventRecord(ev[1], s), "hipEventRecord")) break;
        // getFrontier (bfs/bfs.cu:13-18): the synchronous read-back ends the level.
        if (!check(hipMemcpyAsync(&count, frontier, sizeof count, hipMemcpyDeviceToHost, s), "hipMemcpyAsync") ||
            !check(hipStreamSynchronize(s), "hipStreamSynchronize"))
            break;
        if (counts_host) counts_host[curr] = count;
        if (level_ms_host) {
            float ms = 0;
            if (!check(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime")) break;
            level_ms_host[curr] = ms;
        }
        curr++;
    }
    *num_levels = curr;
    cleanup();
    return status;
}

// ----------------------------------------------- allocation-free level loop

// Scratch of bw_bfs_run_static: one frontier counter per level, so that a
// level's count is never reused before it is read.  Each device holds its
// own copy; its address comes from hipGetSymbolAddress, never from the
// array's name, which is only a host-side stand-in in host code.
static __device__ __align__(256) uint32_t scratch[BW_BFS_STATIC_MAX_LEVELS];

namespace {

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
uint32_t *g_scratch[kMaxDevices] = {};

// Device address of the current device's scratch, looked up once per device.
int static_scratch(uint32_t **out)
{
    int dev = 0;
    BW_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(BW_HIP_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        BW_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<uint32_t *>(p);
    }
    *out = g_scratch[dev];
    return BW_HIP_OK;
}

}  // namespace

extern "C" int bw_bfs_run_static(const bw_graph_dev *g, const int32_t *schedule, int32_t schedule_len,
                                 int32_t reduction, uint32_t root, int32_t *levels, uint32_t *counts_host,
                                 int64_t max_levels, int64_t *num_levels, const bw_launch *cfg)
{
    if (!g || !schedule || schedule_len <= 0 || !levels || !num_levels)
        return fail(BW_HIP_ERR_ARG, "bw_bfs_run_static: invalid argument");
    if (max_levels < 0 || max_levels > BW_BFS_STATIC_MAX_LEVELS)
        return fail(BW_HIP_ERR_ARG, "bw_bfs_run_static: max_levels " + std::to_string(max_levels) +
                                        " exceeds the " + std::to_string(BW_BFS_STATIC_MAX_LEVELS) +
                                        " frontier counters of the static scratch");
    *num_levels = 0;
    if (root >= g->vertex_count) return BW_HIP_OK;  // bfs/bfs.cpp:53
    hipStream_t s = stream_of(cfg);
    uint32_t *counters = nullptr;
    BW_TRY(static_scratch(&counters));
    // The counters of this search start at zero (a fill kernel, no memset).
    if (max_levels > 0) BW_TRY(bw_hip_fill_u32(counters, 0u, static_cast<uint64_t>(max_levels), s));
    BW_TRY(bw_hip_fill_u32(levels, static_cast<uint32_t>(kUnvisited), g->vertex_count, s));
    hipLaunchKernelGGL(set_root_kernel, dim3(1), dim3(1), 0, s, levels, root);
    BW_LAUNCH_CHECK();
    int64_t curr = 0;
    uint32_t count = 1;
    while (count != 0) {
        if (curr >= max_levels) {
            *num_levels = curr;
            return fail(BW_HIP_ERR_LIMIT, "BFS did not finish within max_levels launches");
        }
        const int32_t variant = schedule[curr % schedule_len];
        const int st = bw_bfs_level(g, variant, reduction, levels, static_cast<int32_t>(curr), counters + curr, cfg);
        if (st != BW_HIP_OK) {
            *num_levels = curr;
            return st;
        }
        // getFrontier (bfs/bfs.cu:13-18): the read-back ends the level.
        BW_HIP_CHECK(hipMemcpyAsync(&count, counters + curr, sizeof count, hipMemcpyDeviceToHost, s));
        BW_HIP_CHECK(hipStreamSynchronize(s));
        if (counts_host) counts_host[curr] = count;
        curr++;
    }
    *num_levels = curr;
    return BW_HIP_OK;
}
