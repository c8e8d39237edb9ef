// Propagation blocking for the push forms of PageRank on gfx950: the
// iteration kernels of pagerank/edgelist.cu:3-23, struct_edgelist.cu:3-24,
// vertex_push.cu:3-31 and vertex_push_warp.cu:4-49 (merijn/Belewitte) add
// every edge's contribution to new_pr[destination] with one global FP32
// atomic.  On MI355X a device-scope float atomic to a random address runs
// at about 21 G/s, and at 4.6-7.3 G/s when many edges of an R-MAT graph hit
// the same hub, so the edge loop cannot get near the HBM roofline that way.
//
// Here one iteration is two launches over a layout built once per run:
//   1. bin: each workgroup streams a fixed range of 8192 edges in their
//      stored order (source-grouped), computes each edge's contribution,
//      sorts (destination row within its partition, contribution) by
//      destination partition of 2^13 rows in LDS, and writes each
//      partition's run to its bin with consecutive stores.  A workgroup's
//      slots in each bin are fixed by the layout, so only LDS atomics hand
//      them out.
//   2. accumulate: each chunk of a partition's bin is summed into an FP64
//      row vector in LDS and written once: stored when the chunk is the
//      whole bin, otherwise added atomically (one add per row and chunk).
// The sums' order differs from the reference's, whose atomics have no
// fixed order either; each row's FP64 partial is rounded once to FP32.

#include "common.hpp"

#include <algorithm>
#include <memory>
#include <vector>

namespace bw {
namespace {

constexpr uint32_t kRowsLog = 13;
constexpr uint32_t kRows = 1u << kRowsLog;
constexpr uint32_t kMaxParts = 4096;        // LDS cursors of the bin kernel
constexpr uint32_t kBinThreads = 256;
constexpr uint32_t kBinBaseEdges = 8192;    // edges per bin workgroup (scaled up for big graphs)
constexpr uint32_t kAccThreads = 512;
constexpr uint32_t kChunkEdges = 1u << 17;  // most edges one accumulate workgroup sums
constexpr uint64_t kMaxLayout = 1ull << 25; // partitions x bin workgroups

// -------------------------------------------------------- edge sources

// The three stored forms the push variants read, each yielding (source,
// destination) for edge e and the contribution a source sends along each
// of its out-edges.
struct EdgeListSrc {
    const uint32_t *in, *out;
    __device__ __forceinline__ void edge(uint64_t e, uint32_t &s, uint32_t &d) const { s = in[e], d = out[e]; }
    // pagerank/edgelist.cu:15-21: pr[src] / deg[src], zero without out-edges.
    __device__ __forceinline__ float value(uint32_t s, const uint32_t *__restrict__ deg,
                                          const float *__restrict__ pr) const
    {
        const uint32_t d = deg[s];
        return d != 0 ? pr[s] / static_cast<float>(d) : 0.0f;
    }
};
struct StructSrc {
    const uint32_t *pairs;
    __device__ __forceinline__ void edge(uint64_t e, uint32_t &s, uint32_t &d) const
    {
        s = pairs[2 * e], d = pairs[2 * e + 1];
    }
    __device__ __forceinline__ float value(uint32_t s, const uint32_t *__restrict__ deg,
                                          const float *__restrict__ pr) const
    {
        const uint32_t d = deg[s];
        return d != 0 ? pr[s] / static_cast<float>(d) : 0.0f;
    }
};
// A forward CSR with each edge's source expanded once per layout
// (src[e] = the row holding edge e): vertex_push.cu:16-30 sends
// pr[v] / (off[v+1] - off[v]) along each out-edge of v.
template<typename Off>
struct CsrSrc {
    const Off *off;
    const uint32_t *cols, *src;
    __device__ __forceinline__ void edge(uint64_t e, uint32_t &s, uint32_t &d) const { s = src[e], d = cols[e]; }
    __device__ __forceinline__ float value(uint32_t s, const uint32_t *__restrict__, const float *__restrict__ pr) const
    {
        const uint64_t n = static_cast<uint64_t>(off[s + 1]) - static_cast<uint64_t>(off[s]);
        return pr[s] / static_cast<float>(static_cast<uint32_t>(n));
    }
};

// src[e] = the CSR row holding edge e, one wave per row.
template<typename Off>
__global__ void expand_sources_kernel(uint64_t V, const Off *__restrict__ off, uint32_t *src)
{
    const uint64_t waves = static_cast<uint64_t>(blockDim.x) * gridDim.x / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    for (uint64_t row = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave; row < V;
         row += waves)
        for (uint64_t j = static_cast<uint64_t>(off[row]) + lane; j < static_cast<uint64_t>(off[row + 1]); j += kWave)
            src[j] = static_cast<uint32_t>(row);
}

// ----------------------------------------------------------- layout

// Per bin workgroup, the count of its edges per destination partition
// (partition-major: counts[p * nb + blk]).
template<typename Src>
__global__ __launch_bounds__(kBinThreads) void scatter_count_kernel(Src src, uint64_t E, uint32_t per_block,
                                                                    uint32_t parts, uint32_t nb, uint32_t *counts)
{
    __shared__ uint32_t hist[kMaxParts];
    for (uint32_t p = threadIdx.x; p < parts; p += kBinThreads) hist[p] = 0;
    __syncthreads();
    const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * per_block;
    const uint64_t b1 = b0 + per_block < E ? b0 + per_block : E;
    for (uint64_t e = b0 + threadIdx.x; e < b1; e += kBinThreads) {
        uint32_t s, d;
        src.edge(e, s, d);
        atomicAdd(&hist[d >> kRowsLog], 1u);
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < parts; p += kBinThreads)
        counts[static_cast<uint64_t>(p) * nb + blockIdx.x] = hist[p];
}

// Exclusive scan of n uint32 counts into uint32 positions, in three passes:
// per-tile totals, a one-workgroup scan of the totals, per-tile scans.
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kScanTile = kScanThreads * 4;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds, uint32_t &total)
{
    // Wave inclusive scan, then the wave totals.
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, kWave);
        if (static_cast<int>(lane) >= off) x += y;
    }
    if (lane == kWave - 1) lds[wv] = x;
    __syncthreads();
    if (wv == 0) {
        const uint32_t nw = blockDim.x / kWave;
        uint32_t w = lane < nw ? lds[lane] : 0u;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(w, off, kWave);
            if (static_cast<int>(lane) >= off) w += y;
        }
        if (lane < nw) lds[lane] = w;
    }
    __syncthreads();
    total = lds[blockDim.x / kWave - 1];
    const uint32_t before = wv ? lds[wv - 1] : 0u;
    __syncthreads();
    return before + x - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_tiles_kernel(const uint32_t *__restrict__ in, uint64_t n,
                                                                  uint32_t *tile_sums)
{
    __shared__ uint32_t lds[kScanThreads / kWave];
    const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * 4;
    uint32_t v = 0;
    for (uint32_t k = 0; k < 4; k++)
        if (t0 + k < n) v += in[t0 + k];
    uint32_t total;
    (void)block_exclusive_scan(v, lds, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void scan_sums_kernel(uint32_t *sums, uint32_t n)
{
    __shared__ uint32_t lds[kScanThreads / kWave];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += kScanThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, lds, total);
        if (i < n) sums[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t *__restrict__ in, uint64_t n,
                                                                  const uint32_t *__restrict__ tile_sums,
                                                                  uint32_t *out)
{
    __shared__ uint32_t lds[kScanThreads / kWave];
    const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * 4;
    uint32_t v[4], s = 0;
    for (uint32_t k = 0; k < 4; k++) {
        v[k] = t0 + k < n ? in[t0 + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = tile_sums[blockIdx.x] + block_exclusive_scan(s, lds, total);
    for (uint32_t k = 0; k < 4; k++) {
        if (t0 + k < n) out[t0 + k] = run;
        run += v[k];
    }
}

// base[blk * parts + p] = pos[p * nb + blk]: the bin kernel reads its
// workgroup's slot bases contiguously.
__global__ void transpose_kernel(const uint32_t *__restrict__ pos, uint32_t parts, uint32_t nb, uint32_t *base)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<uint64_t>(parts) * nb) return;
    const uint32_t p = static_cast<uint32_t>(i / nb), blk = static_cast<uint32_t>(i % nb);
    base[static_cast<uint64_t>(blk) * parts + p] = pos[i];
}

// starts[p] = pos[p * nb]: the first slot of each partition's bin.
__global__ void bin_starts_kernel(const uint32_t *__restrict__ pos, uint32_t parts, uint32_t nb, uint32_t *starts)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < parts) starts[p] = pos[static_cast<uint64_t>(p) * nb];
}

// -------------------------------------------------------- iteration

struct BinOut {
    const uint32_t *base;  // [nb][parts] first slot of each (workgroup, partition)
    uint16_t *row;         // [E] destination row within its partition
    float *val;            // [E] contribution
    uint32_t parts;
    uint32_t per_block;
};

__device__ __forceinline__ void bin_setup(const BinOut &o, uint32_t *cursor)
{
    const uint32_t *b = o.base + static_cast<uint64_t>(blockIdx.x) * o.parts;
    for (uint32_t p = threadIdx.x; p < o.parts; p += kBinThreads) cursor[p] = b[p];
    __syncthreads();
}

__device__ __forceinline__ void bin_put(const BinOut &o, uint32_t *cursor, uint32_t d, float c)
{
    const uint32_t slot = atomicAdd(&cursor[d >> kRowsLog], 1u);
    o.row[slot] = static_cast<uint16_t>(d & (kRows - 1));
    o.val[slot] = c;
}

// Edge lists: thread t takes edges t, t + 256, ... of the workgroup's
// range, kBinBatch at a time with every load of a batch in flight together
// (indices past the range repeat its last edge and are not binned).
constexpr uint32_t kBinBatch = 8;

template<typename Src>
__global__ __launch_bounds__(kBinThreads) void scatter_bin_kernel(Src src, uint64_t E,
                                                                  const uint32_t *__restrict__ deg,
                                                                  const float *__restrict__ pr, BinOut o)
{
    __shared__ uint32_t cursor[kMaxParts];
    bin_setup(o, cursor);
    const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * o.per_block;
    const uint64_t b1 = b0 + o.per_block < E ? b0 + o.per_block : E;
    for (uint64_t e0 = b0 + threadIdx.x; e0 < b1; e0 += kBinBatch * kBinThreads) {
        uint32_t s[kBinBatch], d[kBinBatch];
        float c[kBinBatch];
#pragma unroll
        for (uint32_t k = 0; k < kBinBatch; k++) {
            const uint64_t e = e0 + k * kBinThreads;
            src.edge(e < b1 ? e : b1 - 1, s[k], d[k]);
        }
#pragma unroll
        for (uint32_t k = 0; k < kBinBatch; k++) c[k] = src.value(s[k], deg, pr);
#pragma unroll
        for (uint32_t k = 0; k < kBinBatch; k++)
            if (e0 + k * kBinThreads < b1) bin_put(o, cursor, d[k], c[k]);
    }
}

// The same with the workgroup's output sorted by partition in LDS first:
// each partition's run of the workgroup is then written with consecutive
// stores, where the direct form's lanes each store 2 and 4 bytes to a
// different run (one L2 request per lane and store).  Used when the
// workgroup's edge range is the base 8192 edges.
constexpr uint32_t kStagedThreads = 1024;
constexpr uint32_t kStagedPerThread = kBinBaseEdges / kStagedThreads;

template<typename Src>
__global__ __launch_bounds__(kStagedThreads) void scatter_bin_staged_kernel(Src src, uint64_t E,
                                                                           const uint32_t *__restrict__ deg,
                                                                           const float *__restrict__ pr, BinOut o)
{
    __shared__ uint32_t counts[kMaxParts], starts[kMaxParts], gbase[kMaxParts];
    __shared__ uint16_t stRow[kBinBaseEdges], stPart[kBinBaseEdges];
    __shared__ float stVal[kBinBaseEdges];
    __shared__ uint32_t scan[kStagedThreads / kWave];
    const uint32_t t = threadIdx.x;
    const uint32_t *b = o.base + static_cast<uint64_t>(blockIdx.x) * o.parts;
    for (uint32_t p = t; p < o.parts; p += kStagedThreads) {
        counts[p] = 0;
        gbase[p] = b[p];
    }
    __syncthreads();
    const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * kBinBaseEdges;
    const uint64_t b1 = b0 + kBinBaseEdges < E ? b0 + kBinBaseEdges : E;
    uint32_t sv[kStagedPerThread], d[kStagedPerThread], rank[kStagedPerThread];
    float c[kStagedPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kStagedPerThread; k++) {
        const uint64_t e = b0 + k * kStagedThreads + t;
        src.edge(e < b1 ? e : b1 - 1, sv[k], d[k]);
    }
#pragma unroll
    for (uint32_t k = 0; k < kStagedPerThread; k++) c[k] = src.value(sv[k], deg, pr);
#pragma unroll
    for (uint32_t k = 0; k < kStagedPerThread; k++)
        rank[k] = b0 + k * kStagedThreads + t < b1 ? atomicAdd(&counts[d[k] >> kRowsLog], 1u) : 0u;
    __syncthreads();
    // starts = exclusive scan of counts; each thread scans 4 consecutive.
    constexpr uint32_t kPer = kMaxParts / kStagedThreads;
    uint32_t part[kPer], sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < kPer; q++) {
        const uint32_t p = t * kPer + q;
        part[q] = p < o.parts ? counts[p] : 0u;
        sum += part[q];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, scan, total);
#pragma unroll
    for (uint32_t q = 0; q < kPer; q++) {
        const uint32_t p = t * kPer + q;
        if (p < o.parts) starts[p] = run;
        run += part[q];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kStagedPerThread; k++) {
        if (b0 + k * kStagedThreads + t >= b1) continue;
        const uint32_t p = d[k] >> kRowsLog;
        const uint32_t pos = starts[p] + rank[k];
        stRow[pos] = static_cast<uint16_t>(d[k] & (kRows - 1));
        stPart[pos] = static_cast<uint16_t>(p);
        stVal[pos] = c[k];
    }
    __syncthreads();
    const uint32_t n = static_cast<uint32_t>(b1 - b0);
    for (uint32_t i = t; i < n; i += kStagedThreads) {
        const uint32_t p = stPart[i];
        const uint32_t slot = gbase[p] + (i - starts[p]);
        o.row[slot] = stRow[i];
        o.val[slot] = stVal[i];
    }
}

// One chunk of one partition's bin: (partition, first slot, end slot,
// whole = the chunk is the partition's whole bin).
struct Chunk {
    uint32_t part, begin, end, whole;
};

__global__ __launch_bounds__(kAccThreads) void scatter_accumulate_kernel(const Chunk *__restrict__ chunks,
                                                                         const uint16_t *__restrict__ row,
                                                                         const float *__restrict__ val, uint64_t V,
                                                                         float *np)
{
    __shared__ double acc[kRows];
    const Chunk ch = chunks[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < kRows; i += kAccThreads) acc[i] = 0.0;
    __syncthreads();
    for (uint32_t j = ch.begin + threadIdx.x; j < ch.end; j += kAccThreads)
        atomicAdd(&acc[row[j]], static_cast<double>(val[j]));
    __syncthreads();
    const uint64_t r0 = static_cast<uint64_t>(ch.part) << kRowsLog;
    const uint32_t n = static_cast<uint32_t>(V - r0 < kRows ? V - r0 : kRows);
    for (uint32_t i = threadIdx.x; i < n; i += kAccThreads) {
        const float x = static_cast<float>(acc[i]);
        if (ch.whole) np[r0 + i] = x;
        else if (x != 0.0f) unsafeAtomicAdd(&np[r0 + i], x);
    }
}

}  // namespace
}  // namespace bw

using namespace bw;

struct bw_pr_scatter {
    int32_t kind = 0;
    uint64_t V = 0, E = 0;
    const void *a = nullptr, *b = nullptr;
    int32_t ob = 4;
    uint32_t parts = 0, nb = 0, per_block = 0, nchunks = 0;
    uint32_t *base = nullptr, *src = nullptr;
    uint16_t *row = nullptr;
    float *val = nullptr;
    Chunk *chunks = nullptr;
    ~bw_pr_scatter()
    {
        for (void *p : {static_cast<void *>(base), static_cast<void *>(src), static_cast<void *>(row),
                        static_cast<void *>(val), static_cast<void *>(chunks)})
            if (p) (void)hipFree(p);
    }
};

namespace {

template<typename F>
int with_off(const void *off, int32_t ob, F &&f)
{
    if (ob == 4) return f(static_cast<const uint32_t *>(off));
    if (ob == 8) return f(static_cast<const uint64_t *>(off));
    return fail(BW_HIP_ERR_ARG, "offset_bytes must be 4 or 8");
}

// Temporary device buffers of the layout build.
struct Tmp {
    std::vector<void *> p;
    ~Tmp()
    {
        for (void *x : p) (void)hipFree(x);
    }
    template<typename T>
    int alloc(T *&out, uint64_t n)
    {
        void *q = nullptr;
        BW_HIP_CHECK(hipMalloc(&q, sizeof(T) * std::max<uint64_t>(n, 1)));
        p.push_back(q);
        out = static_cast<T *>(q);
        return BW_HIP_OK;
    }
};

int scan_exclusive(const uint32_t *in, uint64_t n, uint32_t *out, hipStream_t s)
{
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles > 0xFFFFFFFFull) return fail(BW_HIP_ERR_ARG, "scan too long");
    Tmp t;
    uint32_t *sums = nullptr;
    BW_TRY(t.alloc(sums, tiles));
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(static_cast<uint32_t>(tiles)), dim3(kScanThreads), 0, s, in, n, sums);
    BW_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanThreads), 0, s, sums, static_cast<uint32_t>(tiles));
    BW_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_apply_kernel, dim3(static_cast<uint32_t>(tiles)), dim3(kScanThreads), 0, s, in, n, sums,
                       out);
    BW_LAUNCH_CHECK();
    BW_HIP_CHECK(hipStreamSynchronize(s));
    return BW_HIP_OK;
}

}  // namespace

extern "C" int bw_pr_scatter_create(bw_pr_scatter **out, int32_t kind, uint64_t V, uint64_t E, const void *a,
                                    const void *b, int32_t offset_bytes, void *stream)
{
    if (!out) return fail(BW_HIP_ERR_ARG, "bw_pr_scatter_create: out is NULL");
    *out = nullptr;
    if (kind != BW_PR_SCATTER_EDGE_LIST && kind != BW_PR_SCATTER_STRUCT && kind != BW_PR_SCATTER_CSR)
        return fail(BW_HIP_ERR_ARG, "unknown scatter source kind");
    if (V == 0 || E == 0) return fail(BW_HIP_ERR_ARG, "the scatter layout needs vertices and edges");
    if (!a || (kind != BW_PR_SCATTER_STRUCT && !b)) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    if (kind == BW_PR_SCATTER_CSR && offset_bytes != 4 && offset_bytes != 8)
        return fail(BW_HIP_ERR_ARG, "offset_bytes must be 4 or 8");
    const uint64_t parts = (V + kRows - 1) >> kRowsLog;
    if (parts > kMaxParts || E > 0xFFFFFFFFull)
        return fail(BW_HIP_ERR_ARG, "graph too large for the scatter layout (at most 2^25 vertices and "
                                    "2^32 - 1 edges)");
    // Edges per bin workgroup: a multiple of 8192 keeping the layout table
    // (partitions x workgroups) at most 2^25 entries.
    uint64_t per = kBinBaseEdges;
    while (parts * ((E + per - 1) / per) > kMaxLayout) per *= 2;
    const uint64_t nb = (E + per - 1) / per;
    hipStream_t s = static_cast<hipStream_t>(stream);

    auto sc = std::make_unique<bw_pr_scatter>();
    sc->kind = kind, sc->V = V, sc->E = E, sc->a = a, sc->b = b, sc->ob = offset_bytes;
    sc->parts = static_cast<uint32_t>(parts), sc->nb = static_cast<uint32_t>(nb);
    sc->per_block = static_cast<uint32_t>(per);
    const uint64_t cells = parts * nb;
    Tmp tmp;
    uint32_t *counts = nullptr, *pos = nullptr;
    BW_TRY(tmp.alloc(counts, cells));
    BW_TRY(tmp.alloc(pos, cells));
    BW_HIP_CHECK(hipMalloc(&sc->base, sizeof(uint32_t) * cells));
    BW_HIP_CHECK(hipMalloc(&sc->row, sizeof(uint16_t) * E));
    BW_HIP_CHECK(hipMalloc(&sc->val, sizeof(float) * E));

    if (kind == BW_PR_SCATTER_EDGE_LIST) {
        EdgeListSrc src{static_cast<const uint32_t *>(a), static_cast<const uint32_t *>(b)};
        hipLaunchKernelGGL(scatter_count_kernel<EdgeListSrc>, dim3(sc->nb), dim3(kBinThreads), 0, s, src, E,
                           sc->per_block, sc->parts, sc->nb, counts);
    } else if (kind == BW_PR_SCATTER_STRUCT) {
        StructSrc src{static_cast<const uint32_t *>(a)};
        hipLaunchKernelGGL(scatter_count_kernel<StructSrc>, dim3(sc->nb), dim3(kBinThreads), 0, s, src, E,
                           sc->per_block, sc->parts, sc->nb, counts);
    } else {
        BW_HIP_CHECK(hipMalloc(&sc->src, sizeof(uint32_t) * E));
        BW_TRY(with_off(a, offset_bytes, [&](auto *o) {
            using Off = std::remove_const_t<std::remove_pointer_t<decltype(o)>>;
            uint32_t block = 256, grid = 0;
            BW_TRY(resolve_geometry(nullptr, V * kWave, block, grid));
            hipLaunchKernelGGL(expand_sources_kernel<Off>, dim3(grid), dim3(block), 0, s, V, o, sc->src);
            BW_LAUNCH_CHECK();
            CsrSrc<Off> src{o, static_cast<const uint32_t *>(b), sc->src};
            hipLaunchKernelGGL(scatter_count_kernel<CsrSrc<Off>>, dim3(sc->nb), dim3(kBinThreads), 0, s, src, E,
                               sc->per_block, sc->parts, sc->nb, counts);
            return BW_HIP_OK;
        }));
    }
    BW_LAUNCH_CHECK();
    BW_TRY(scan_exclusive(counts, cells, pos, s));
    hipLaunchKernelGGL(transpose_kernel, dim3(static_cast<uint32_t>((cells + 255) / 256)), dim3(256), 0, s, pos,
                       sc->parts, sc->nb, sc->base);
    BW_LAUNCH_CHECK();

    // Partition p's bin is [pos[p * nb], pos[(p + 1) * nb]); chunks of at
    // most kChunkEdges, the whole bin in one chunk when it fits.
    std::vector<uint32_t> starts(parts + 1);
    uint32_t *dstarts = nullptr;
    BW_TRY(tmp.alloc(dstarts, parts));
    hipLaunchKernelGGL(bin_starts_kernel, dim3(static_cast<uint32_t>((parts + 255) / 256)), dim3(256), 0, s, pos,
                       sc->parts, sc->nb, dstarts);
    BW_LAUNCH_CHECK();
    BW_HIP_CHECK(hipMemcpyAsync(starts.data(), dstarts, sizeof(uint32_t) * parts, hipMemcpyDeviceToHost, s));
    BW_HIP_CHECK(hipStreamSynchronize(s));
    starts[parts] = static_cast<uint32_t>(E);
    std::vector<Chunk> chunks;
    for (uint32_t p = 0; p < parts; p++) {
        const uint32_t b0 = starts[p], b1 = starts[p + 1];
        if (b0 == b1) continue;
        const uint32_t whole = (b1 - b0) <= kChunkEdges ? 1u : 0u;
        for (uint32_t c = b0; c < b1; c += kChunkEdges)
            chunks.push_back(Chunk{p, c, std::min(b1, c + kChunkEdges), whole});
    }
    sc->nchunks = static_cast<uint32_t>(chunks.size());
    BW_HIP_CHECK(hipMalloc(&sc->chunks, sizeof(Chunk) * std::max<size_t>(chunks.size(), 1)));
    if (!chunks.empty())
        BW_HIP_CHECK(hipMemcpy(sc->chunks, chunks.data(), sizeof(Chunk) * chunks.size(), hipMemcpyHostToDevice));
    *out = sc.release();
    return BW_HIP_OK;
}

extern "C" int bw_pr_scatter_destroy(bw_pr_scatter *sc)
{
    delete sc;
    return BW_HIP_OK;
}

extern "C" int bw_pr_scatter_iteration(const bw_pr_scatter *sc, const uint32_t *deg, const float *pr, float *np,
                                       void *stream)
{
    if (!sc || !pr || !np) return fail(BW_HIP_ERR_ARG, "bw_pr_scatter_iteration: invalid argument");
    if (sc->kind != BW_PR_SCATTER_CSR && !deg) return fail(BW_HIP_ERR_ARG, "edge-list scatter needs the degrees");
    hipStream_t s = static_cast<hipStream_t>(stream);
    BinOut o{sc->base, sc->row, sc->val, sc->parts, sc->per_block};
    auto bin = [&](auto src) {
        using Src = decltype(src);
        if (sc->per_block == kBinBaseEdges)
            hipLaunchKernelGGL(scatter_bin_staged_kernel<Src>, dim3(sc->nb), dim3(kStagedThreads), 0, s, src, sc->E,
                               deg, pr, o);
        else
            hipLaunchKernelGGL(scatter_bin_kernel<Src>, dim3(sc->nb), dim3(kBinThreads), 0, s, src, sc->E, deg, pr,
                               o);
        return BW_HIP_OK;
    };
    if (sc->kind == BW_PR_SCATTER_EDGE_LIST) {
        bin(EdgeListSrc{static_cast<const uint32_t *>(sc->a), static_cast<const uint32_t *>(sc->b)});
    } else if (sc->kind == BW_PR_SCATTER_STRUCT) {
        bin(StructSrc{static_cast<const uint32_t *>(sc->a)});
    } else {
        BW_TRY(with_off(sc->a, sc->ob, [&](auto *off) {
            using Off = std::remove_const_t<std::remove_pointer_t<decltype(off)>>;
            return bin(CsrSrc<Off>{off, static_cast<const uint32_t *>(sc->b), sc->src});
        }));
    }
    BW_LAUNCH_CHECK();
    if (sc->nchunks) {
        hipLaunchKernelGGL(scatter_accumulate_kernel, dim3(sc->nchunks), dim3(kAccThreads), 0, s, sc->chunks, sc->row,
                           sc->val, sc->V, np);
        BW_LAUNCH_CHECK();
    }
    return BW_HIP_OK;
}
