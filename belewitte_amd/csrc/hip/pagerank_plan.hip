// The binned PageRank plan, built on the device.
//
// The layout is the one of pagerank_bins_plan.hpp (propagation blocking:
// partitions of 2^L rows, one value per (source, partition) pair, bins in
// (partition, source) order, edge codes in (partition, source, row) order,
// scatter blocks of consecutive vertices), and the arrays come out equal to
// the host builder's bit for bit.  What the host does in sequential passes
// is done here with stable radix sorts and scans:
//   1. the forward lists: the reverse edges stably sorted by source, so each
//      source's rows come in increasing order;
//   2. the pairs: the first edge of each (source, partition) run, compacted
//      by an exclusive scan of the new-pair flags;
//   3. the scatter blocks: the host's greedy cut, one wave per partition
//      walking its vertices with wave-uniform state, then each block's runs
//      (distinct target partitions) counted in an LDS bitmap, one
//      workgroup per block;
//   4. the bin index of each pair: a stable sort of the pairs by partition;
//      the scatter (flat) order: a stable sort by (block, partition);
//   5. the edge codes: the forward edges stably sorted by partition.
// Only K-sized tables (partition starts, chunk bases, block bases) and the
// chunk and task records (a few thousand) are assembled on the host.

#include <cstring>  // rocPRIM's headers use memset without including it

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "pagerank_bins_plan.hpp"
#include "pagerank_plan.hpp"

#include <algorithm>
#include <vector>

namespace bw {
namespace bins {
namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kEdgesPerStep = 4096;  // edges per workgroup step of the row search
constexpr uint32_t kTouchedWords = kMaxRuns / 32;

uint32_t grid_for(uint64_t n, uint32_t per = kThreads)
{
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((n + per - 1) / per, 65536)));
}

uint32_t bits_for(uint64_t max_value)
{
    uint32_t b = 1;
    while (b < 32 && (max_value >> b) != 0) b++;
    return b;
}

// The row of reverse edge e: the largest r in [lo, hi] with roff[r] <= e.
__device__ __forceinline__ uint32_t row_of(const uint32_t *roff, uint32_t lo, uint32_t hi, uint32_t e)
{
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (roff[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// rows[e] = row of reverse edge e; bad = 1 if a column is out of range.
__global__ __launch_bounds__(kThreads) void plan_rows_kernel(uint32_t V, uint32_t E, const uint32_t *roff,
                                                             const uint32_t *rcols, uint32_t *rows, uint32_t *bad)
{
    __shared__ uint32_t range[2];
    for (uint64_t e0 = static_cast<uint64_t>(blockIdx.x) * kEdgesPerStep; e0 < E;
         e0 += static_cast<uint64_t>(gridDim.x) * kEdgesPerStep) {
        const uint32_t a = static_cast<uint32_t>(e0);
        const uint32_t b = static_cast<uint32_t>(std::min<uint64_t>(e0 + kEdgesPerStep, E));
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t r0 = row_of(roff, 0, V - 1, a);
            range[0] = r0;
            range[1] = row_of(roff, r0, V - 1, b - 1);
        }
        __syncthreads();
        const uint32_t r0 = range[0], r1 = range[1];
        for (uint32_t e = a + threadIdx.x; e < b; e += kThreads) {
            rows[e] = row_of(roff, r0, r1, e);
            if (rcols[e] >= V) *bad = 1u;
        }
    }
}

// foff[s] = first index of the sorted sources holding a value >= s, s <= V.
__global__ __launch_bounds__(kThreads) void plan_foff_kernel(uint32_t V, uint32_t E, const uint32_t *fsrc,
                                                             uint32_t *foff)
{
    for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; s <= V;
         s += static_cast<uint64_t>(gridDim.x) * kThreads) {
        uint32_t lo = 0, hi = E;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (fsrc[mid] < s) lo = mid + 1;
            else hi = mid;
        }
        foff[s] = lo;
    }
}

// flag[e] = 1 when forward edge e opens a (source, partition) pair; flag[E] = 0.
__global__ __launch_bounds__(kThreads) void plan_fresh_kernel(uint32_t E, uint32_t L, const uint32_t *fsrc,
                                                              const uint32_t *frow, uint32_t *flag)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i <= E;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        const uint32_t e = static_cast<uint32_t>(i);
        uint32_t f = 0;
        if (e < E)
            f = e == 0 || fsrc[e] != fsrc[e - 1] || (frow[e] >> L) != (frow[e - 1] >> L) ? 1u : 0u;
        flag[e] = f;
    }
}

// Pair p = pscan[e] of each opening edge e: its source and partition.
__global__ __launch_bounds__(kThreads) void plan_pairs_kernel(uint32_t E, uint32_t L, const uint32_t *fsrc,
                                                              const uint32_t *frow, const uint32_t *flag,
                                                              const uint32_t *pscan, uint32_t *pair_src,
                                                              uint32_t *pair_part)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < E;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        if (!flag[i]) continue;
        const uint32_t p = pscan[i];
        pair_src[p] = fsrc[i];
        pair_part[p] = frow[i] >> L;
    }
}

// pbase[s] = first pair of source s (pbase[V] = the pair count).
__global__ __launch_bounds__(kThreads) void plan_pbase_kernel(uint32_t V, const uint32_t *foff,
                                                              const uint32_t *pscan, uint32_t *pbase)
{
    for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; s <= V;
         s += static_cast<uint64_t>(gridDim.x) * kThreads)
        pbase[s] = pscan[foff[s]];
}

// Pairs per partition (K <= kMaxRuns): an LDS histogram per workgroup.
__global__ __launch_bounds__(kThreads) void plan_hist_kernel(uint32_t U, uint32_t K, const uint32_t *pair_part,
                                                             uint32_t *pairsK)
{
    __shared__ uint32_t h[kMaxRuns];
    for (uint32_t k = threadIdx.x; k < K; k += kThreads) h[k] = 0;
    __syncthreads();
    for (uint64_t p = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; p < U;
         p += static_cast<uint64_t>(gridDim.x) * kThreads)
        atomicAdd(&h[pair_part[p]], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += kThreads)
        if (h[k]) atomicAdd(&pairsK[k], h[k]);
}

// edge_bound[k] = roff[min(k << L, V)], k <= K.
__global__ void plan_bounds_kernel(uint32_t V, uint32_t L, uint32_t K, const uint32_t *roff, uint32_t *edge_bound)
{
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= K; k += gridDim.x * blockDim.x)
        edge_bound[k] = roff[std::min<uint64_t>(static_cast<uint64_t>(k) << L, V)];
}

// The scatter blocks of partition k = blockIdx.x, cut like the host
// builder: consecutive vertices, a new block when the current one holds
// kBlockVerts vertices or when adding the vertex would take the block's
// work (vertices x chunks of the partition + pairs) past kBlockWork.  (The
// host builder also closes a block whose distinct target partitions would
// pass kMaxRuns; that never happens, since a plan has at most kMaxRuns
// partitions.)  One wave: lane j holds the pair count of vertex base + j,
// and the greedy walks them with wave-uniform state.  Writes each vertex's
// block (local to the partition) and, at k << L + b, block b's (first
// vertex, end vertex, pairs, 0); plan_block_runs_kernel fills in the runs.
__global__ __launch_bounds__(64) void plan_cut_kernel(uint32_t V, uint32_t L, const uint32_t *pbase,
                                                      const uint32_t *chunk_base, uint32_t *block_local,
                                                      uint4 *pblocks, uint32_t *nblocks)
{
    const uint32_t k = blockIdx.x, lane = threadIdx.x;
    const uint32_t vBeg = k << L;
    const uint32_t vEnd = static_cast<uint32_t>(std::min<uint64_t>(static_cast<uint64_t>(vBeg) + (1u << L), V));
    const uint64_t perVertex = std::max<uint32_t>(1, chunk_base[k + 1] - chunk_base[k]);
    uint32_t nb = 0, v0 = vBeg, npairs = 0;
    uint64_t work = 0;
    auto close_block = [&](uint32_t end) {
        if (lane == 0) pblocks[vBeg + nb] = make_uint4(v0, end, npairs, 0u);
        nb++;
        v0 = end;
        work = 0;
        npairs = 0;
    };
    for (uint32_t base = vBeg; base < vEnd; base += 64) {
        const uint32_t s = base + lane, cnt = std::min(64u, vEnd - base);
        const uint32_t n = s < vEnd ? pbase[s + 1] - pbase[s] : 0u;
        uint32_t mine = 0;  // the block of vertex s
        for (uint32_t j = 0; j < cnt; j++) {
            const uint32_t sj = base + j, nj = __builtin_amdgcn_readlane(n, j);
            if (sj - v0 >= kBlockVerts) close_block(sj);
            const uint64_t w = perVertex + nj;
            if (sj > v0 && work + w > kBlockWork) close_block(sj);
            npairs += nj;
            work += w;
            if (lane == j) mine = nb;
        }
        if (s < vEnd) block_local[s] = mine;
    }
    close_block(vEnd);
    if (lane == 0) nblocks[k] = nb;
}

// Runs of scatter block blockIdx.x: the distinct partitions its pairs
// point into, counted in an LDS bitmap (a plan has at most kMaxRuns
// partitions).  Written to the block record and to bruns.
__global__ __launch_bounds__(kThreads) void plan_block_runs_kernel(const uint32_t *pbase, const uint32_t *pair_part,
                                                                   uint4 *blocks, uint32_t *bruns)
{
    __shared__ uint32_t touched[kTouchedWords];
    __shared__ uint32_t total;
    const uint32_t b = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < kTouchedWords; i += kThreads) touched[i] = 0;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    const uint4 r = blocks[b];
    const uint32_t p1 = pbase[r.y];
    for (uint32_t p = pbase[r.x] + threadIdx.x; p < p1; p += kThreads) {
        const uint32_t kk = pair_part[p];
        atomicOr(&touched[kk >> 5], 1u << (kk & 31));
    }
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < kTouchedWords; i += kThreads) c += static_cast<uint32_t>(__popc(touched[i]));
    if (c) atomicAdd(&total, c);
    __syncthreads();
    if (threadIdx.x == 0) {
        blocks[b].w = total;
        bruns[b] = total;
    }
}

// Global block records from the per-partition ones: (first vertex, end
// vertex, run begin, run count), plus each block's pair and run counts for
// the scans.
__global__ __launch_bounds__(kThreads) void plan_gather_blocks_kernel(uint32_t K, uint32_t L, const uint32_t *block_base,
                                                                      const uint4 *pblocks, uint4 *blocks,
                                                                      uint32_t *bpairs, uint32_t *bruns)
{
    for (uint32_t k = blockIdx.x; k < K; k += gridDim.x) {
        const uint32_t g0 = block_base[k], n = block_base[k + 1] - g0;
        for (uint32_t b = threadIdx.x; b < n; b += kThreads) {
            const uint4 r = pblocks[(k << L) + b];
            blocks[g0 + b] = make_uint4(r.x, r.y, 0, r.w);
            bpairs[g0 + b] = r.z;
            bruns[g0 + b] = r.w;
        }
    }
}

// Run begins: the runs of the blocks before, plus one terminator each.
__global__ __launch_bounds__(kThreads) void plan_run_begin_kernel(uint32_t nb, const uint32_t *run_scan, uint4 *blocks)
{
    for (uint64_t b = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; b < nb;
         b += static_cast<uint64_t>(gridDim.x) * kThreads)
        blocks[b].z = run_scan[b] + static_cast<uint32_t>(b);
}

// Sort keys of the pairs: partition, and (global block, partition).
__global__ __launch_bounds__(kThreads) void plan_pair_keys_kernel(uint32_t U, uint32_t L, uint32_t K,
                                                                  const uint32_t *pair_src, const uint32_t *pair_part,
                                                                  const uint32_t *block_local,
                                                                  const uint32_t *block_base, uint32_t *iota,
                                                                  uint32_t *flat_key, uint32_t *pair_block)
{
    for (uint64_t p = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; p < U;
         p += static_cast<uint64_t>(gridDim.x) * kThreads) {
        const uint32_t s = pair_src[p];
        const uint32_t b = block_base[s >> L] + block_local[s];
        iota[p] = static_cast<uint32_t>(p);
        pair_block[p] = b;
        flat_key[p] = b * K + pair_part[p];
    }
}

__global__ __launch_bounds__(kThreads) void plan_scatter_index_kernel(uint32_t n, const uint32_t *order,
                                                                      uint32_t *index)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * kThreads)
        index[order[i]] = static_cast<uint32_t>(i);
}

// run_start[i] = 1 when flat position i opens a (block, partition) run.
__global__ __launch_bounds__(kThreads) void plan_run_flags_kernel(uint32_t U, const uint32_t *sorted_key,
                                                                  uint32_t *run_start)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i <= U;
         i += static_cast<uint64_t>(gridDim.x) * kThreads)
        run_start[i] = i < U && (i == 0 || sorted_key[i] != sorted_key[i - 1]) ? 1u : 0u;
}

// psrc in flat order, and each run's first flat index and bin index.
__global__ __launch_bounds__(kThreads) void plan_runs_kernel(uint32_t U, const uint32_t *flat_order,
                                                             const uint32_t *pair_src, const uint32_t *pair_block,
                                                             const uint4 *blocks, const uint32_t *bin_index,
                                                             const uint32_t *run_start, const uint32_t *run_scan,
                                                             uint16_t *psrc, uint32_t *run_flat, uint32_t *run_bin)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < U;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        const uint32_t p = flat_order[i], b = pair_block[p];
        psrc[i] = static_cast<uint16_t>(pair_src[p] - blocks[b].x);
        if (run_start[i]) {
            const uint32_t pos = run_scan[i] + b;
            run_flat[pos] = static_cast<uint32_t>(i);
            run_bin[pos] = bin_index[p];
        }
    }
}

// Each block's terminator: the flat index past its last pair, bin 0.
__global__ __launch_bounds__(kThreads) void plan_terminators_kernel(uint32_t nb, const uint4 *blocks,
                                                                    const uint32_t *pair_scan, const uint32_t *bpairs,
                                                                    uint32_t *run_flat, uint32_t *run_bin)
{
    for (uint64_t b = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; b < nb;
         b += static_cast<uint64_t>(gridDim.x) * kThreads) {
        const uint4 r = blocks[b];
        run_flat[r.z + r.w] = pair_scan[b] + bpairs[b];
        run_bin[r.z + r.w] = 0;
    }
}

// Partition key and code of each forward edge.
__global__ __launch_bounds__(kThreads) void plan_code_keys_kernel(uint32_t E, uint32_t L, const uint32_t *frow,
                                                                  const uint32_t *flag, uint32_t *part,
                                                                  uint16_t *code)
{
    const uint32_t mask = (1u << L) - 1u;
    for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; e < E;
         e += static_cast<uint64_t>(gridDim.x) * kThreads) {
        const uint32_t r = frow[e];
        part[e] = r >> L;
        code[e] = static_cast<uint16_t>((r & mask) | (flag[e] ? kNewPair : 0));
    }
}

// codes[edge_start[k] + i - ustart[k]] = i-th code in partition order; the
// new-pair flags of the sorted codes for the task pair bases.
__global__ __launch_bounds__(kThreads) void plan_place_codes_kernel(uint32_t E, const uint32_t *sorted_part,
                                                                    const uint16_t *sorted_code,
                                                                    const uint32_t *edge_start,
                                                                    const uint32_t *ustart, uint16_t *codes,
                                                                    uint32_t *code_flag)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i <= E;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        if (i == E) {
            code_flag[i] = 0;
            continue;
        }
        const uint32_t k = sorted_part[i];
        const uint16_t c = sorted_code[i];
        codes[edge_start[k] + (static_cast<uint32_t>(i) - ustart[k])] = c;
        code_flag[i] = (c & kNewPair) ? 1u : 0u;
    }
}

__global__ __launch_bounds__(kThreads) void plan_gather_kernel(uint32_t n, const uint32_t *src, const uint32_t *idx,
                                                               uint32_t *dst)
{
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) dst[i] = src[idx[i]];
}

__global__ __launch_bounds__(kThreads) void plan_zero16_kernel(uint64_t n, uint16_t *p)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * kThreads)
        p[i] = 0;
}

// Device allocations of one build, freed on every path out.
struct Scratch {
    std::vector<void *> ptrs;
    ~Scratch()
    {
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
    template<typename T>
    T *get(uint64_t n)
    {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(T) * std::max<uint64_t>(n, 1)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return static_cast<T *>(p);
    }
    void give_up(void *p)  // ownership moves to the caller
    {
        for (void *&q : ptrs)
            if (q == p) q = nullptr;
    }
};

template<typename K, typename Vv>
hipError_t sort_pairs(Scratch &sc, const K *kin, K *kout, const Vv *vin, Vv *vout, uint64_t n, uint32_t bits,
                      hipStream_t s)
{
    size_t bytes = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0, bits, s);
    if (e != hipSuccess) return e;
    void *tmp = sc.get<uint8_t>(bytes);
    if (!tmp) return hipErrorOutOfMemory;
    return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, n, 0, bits, s);
}

hipError_t exclusive_scan(Scratch &sc, const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s)
{
    size_t bytes = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
    if (e != hipSuccess) return e;
    void *tmp = sc.get<uint8_t>(bytes);
    if (!tmp) return hipErrorOutOfMemory;
    return rocprim::exclusive_scan(tmp, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
}

}  // namespace

void release(DeviceBins &b)
{
    for (void *p : {static_cast<void *>(b.codes), static_cast<void *>(b.tasks), static_cast<void *>(b.chunks),
                    static_cast<void *>(b.blocks), static_cast<void *>(b.chunk_base), static_cast<void *>(b.run_flat),
                    static_cast<void *>(b.run_bin), static_cast<void *>(b.psrc)})
        if (p) (void)hipFree(p);
    b = DeviceBins();
}

int build_plan_device(const uint32_t *roff, const uint32_t *rcols, uint32_t V, uint32_t E, uint32_t L,
                      uint32_t chunk_tasks, hipStream_t s, DeviceBins &out, std::string &err)
{
    out = DeviceBins();
    if (L < 8 || L > 15) {
        err = "rows_log2 must be in [8, 15]";
        return BW_HIP_ERR_ARG;
    }
    if (chunk_tasks == 0) {
        err = "chunk_tasks must be positive";
        return BW_HIP_ERR_ARG;
    }
    if (V == 0) {
        err = "the plan needs at least one vertex";
        return BW_HIP_ERR_ARG;
    }
    const uint32_t D = 1u << L;
    const uint32_t K = static_cast<uint32_t>((static_cast<uint64_t>(V) + D - 1) >> L);
    if (K > kMaxRuns) {
        err = "too many partitions for the binned layout: raise rows_log2";
        return BW_HIP_ERR_ARG;
    }
    Scratch sc;
#define PLAN_TRY(x)                                                                                     \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) {                                                                         \
            err = std::string(#x) + ": " + hipGetErrorString(e_);                                       \
            return e_ == hipErrorOutOfMemory ? BW_HIP_ERR_NOMEM : BW_HIP_ERR_RUNTIME;                   \
        }                                                                                               \
    } while (0)
#define PLAN_ALLOC(var, T, n)                                                                           \
    T *var = sc.get<T>(n);                                                                              \
    if (!var) {                                                                                         \
        err = "device allocation of the plan builder failed";                                           \
        return BW_HIP_ERR_NOMEM;                                                                        \
    }

    // 1. Forward lists.
    PLAN_ALLOC(rows, uint32_t, E);
    PLAN_ALLOC(bad, uint32_t, 1);
    PLAN_ALLOC(fsrc, uint32_t, E);
    PLAN_ALLOC(frow, uint32_t, E);
    PLAN_ALLOC(foff, uint32_t, static_cast<uint64_t>(V) + 1);
    PLAN_TRY(hipMemsetAsync(bad, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(plan_rows_kernel, dim3(grid_for(E, kEdgesPerStep)), dim3(kThreads), 0, s, V, E, roff, rcols,
                       rows, bad);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(sort_pairs(sc, rcols, fsrc, static_cast<const uint32_t *>(rows), frow, E, bits_for(V - 1), s));
    hipLaunchKernelGGL(plan_foff_kernel, dim3(grid_for(static_cast<uint64_t>(V) + 1)), dim3(kThreads), 0, s, V, E,
                       fsrc, foff);
    PLAN_TRY(hipGetLastError());

    // 2. Pairs.
    PLAN_ALLOC(flag, uint32_t, static_cast<uint64_t>(E) + 1);
    PLAN_ALLOC(pscan, uint32_t, static_cast<uint64_t>(E) + 1);
    hipLaunchKernelGGL(plan_fresh_kernel, dim3(grid_for(static_cast<uint64_t>(E) + 1)), dim3(kThreads), 0, s, E, L,
                       fsrc, frow, flag);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(exclusive_scan(sc, flag, pscan, static_cast<uint64_t>(E) + 1, s));
    uint32_t hbad = 0, U = 0;
    PLAN_TRY(hipMemcpyAsync(&hbad, bad, sizeof hbad, hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipMemcpyAsync(&U, pscan + E, sizeof U, hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipStreamSynchronize(s));
    if (hbad) {
        err = "column id out of range";
        return BW_HIP_ERR_ARG;
    }
    PLAN_ALLOC(pair_src, uint32_t, U);
    PLAN_ALLOC(pair_part, uint32_t, U);
    PLAN_ALLOC(pbase, uint32_t, static_cast<uint64_t>(V) + 1);
    PLAN_ALLOC(pairsK, uint32_t, K);
    PLAN_ALLOC(edge_bound, uint32_t, static_cast<uint64_t>(K) + 1);
    hipLaunchKernelGGL(plan_pairs_kernel, dim3(grid_for(E)), dim3(kThreads), 0, s, E, L, fsrc, frow, flag, pscan,
                       pair_src, pair_part);
    PLAN_TRY(hipGetLastError());
    hipLaunchKernelGGL(plan_pbase_kernel, dim3(grid_for(static_cast<uint64_t>(V) + 1)), dim3(kThreads), 0, s, V, foff,
                       pscan, pbase);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(hipMemsetAsync(pairsK, 0, sizeof(uint32_t) * K, s));
    hipLaunchKernelGGL(plan_hist_kernel, dim3(std::min<uint32_t>(grid_for(U), 1024)), dim3(kThreads), 0, s, U, K,
                       pair_part, pairsK);
    PLAN_TRY(hipGetLastError());
    hipLaunchKernelGGL(plan_bounds_kernel, dim3(1), dim3(kThreads), 0, s, V, L, K, roff, edge_bound);
    PLAN_TRY(hipGetLastError());
    std::vector<uint32_t> hpairsK(K), hbound(K + 1);
    PLAN_TRY(hipMemcpyAsync(hpairsK.data(), pairsK, sizeof(uint32_t) * K, hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipMemcpyAsync(hbound.data(), edge_bound, sizeof(uint32_t) * (K + 1), hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipStreamSynchronize(s));

    // Partition tables (as the host builder).
    std::vector<uint32_t> pair_start(K + 1), edge_start(K + 1), edgesK(K), ustart(K + 1), chunk_base(K + 1, 0);
    uint64_t u = 0, ep = 0, us = 0;
    for (uint32_t k = 0; k < K; k++) {
        pair_start[k] = static_cast<uint32_t>(u);
        edge_start[k] = static_cast<uint32_t>(ep);
        ustart[k] = static_cast<uint32_t>(us);
        edgesK[k] = hbound[k + 1] - hbound[k];
        u += hpairsK[k];
        ep += (static_cast<uint64_t>(edgesK[k]) + 63) & ~63ull;
        us += edgesK[k];
        if (u >= (1ull << 30) || ep >= (1ull << 31) - 8) {
            err = "graph too large for the binned layout";
            return BW_HIP_ERR_ARG;
        }
    }
    pair_start[K] = static_cast<uint32_t>(u);
    edge_start[K] = static_cast<uint32_t>(ep);
    ustart[K] = static_cast<uint32_t>(us);
    if (u != U || us != E) {
        err = "internal error: pair or edge count";
        return BW_HIP_ERR_RUNTIME;
    }
    const uint64_t chunkEdges = static_cast<uint64_t>(chunk_tasks) * kTaskEdges;
    for (uint32_t k = 0; k < K; k++)
        chunk_base[k + 1] = chunk_base[k] + static_cast<uint32_t>((edgesK[k] + chunkEdges - 1) / chunkEdges);
    const uint32_t slots = chunk_base[K];

    // 3. Scatter blocks.
    PLAN_ALLOC(dchunk_base, uint32_t, static_cast<uint64_t>(K) + 1);
    PLAN_ALLOC(block_local, uint32_t, V);
    PLAN_ALLOC(pblocks, uint4, V);
    PLAN_ALLOC(nblocksK, uint32_t, K);
    PLAN_TRY(hipMemcpyAsync(dchunk_base, chunk_base.data(), sizeof(uint32_t) * (K + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(plan_cut_kernel, dim3(K), dim3(64), 0, s, V, L, pbase, dchunk_base, block_local, pblocks,
                       nblocksK);
    PLAN_TRY(hipGetLastError());
    std::vector<uint32_t> hnb(K), block_base(K + 1, 0);
    PLAN_TRY(hipMemcpyAsync(hnb.data(), nblocksK, sizeof(uint32_t) * K, hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipStreamSynchronize(s));
    for (uint32_t k = 0; k < K; k++) block_base[k + 1] = block_base[k] + hnb[k];
    const uint32_t NB = block_base[K];
    if (static_cast<uint64_t>(NB) * K >= (1ull << 32)) {
        err = "too many scatter blocks for the device plan's sort keys";
        return BW_HIP_ERR_ARG;
    }
    PLAN_ALLOC(dblock_base, uint32_t, static_cast<uint64_t>(K) + 1);
    PLAN_ALLOC(blocks, uint4, NB);
    PLAN_ALLOC(bpairs, uint32_t, static_cast<uint64_t>(NB) + 1);
    PLAN_ALLOC(bruns, uint32_t, static_cast<uint64_t>(NB) + 1);
    PLAN_ALLOC(pair_scan_b, uint32_t, static_cast<uint64_t>(NB) + 1);
    PLAN_ALLOC(run_scan_b, uint32_t, static_cast<uint64_t>(NB) + 1);
    PLAN_TRY(hipMemcpyAsync(dblock_base, block_base.data(), sizeof(uint32_t) * (K + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(plan_gather_blocks_kernel, dim3(std::min<uint32_t>(K, 1024)), dim3(kThreads), 0, s, K, L,
                       dblock_base, pblocks, blocks, bpairs, bruns);
    PLAN_TRY(hipGetLastError());
    if (NB) {
        hipLaunchKernelGGL(plan_block_runs_kernel, dim3(NB), dim3(kThreads), 0, s, pbase, pair_part, blocks, bruns);
        PLAN_TRY(hipGetLastError());
    }
    PLAN_TRY(hipMemsetAsync(bpairs + NB, 0, sizeof(uint32_t), s));
    PLAN_TRY(hipMemsetAsync(bruns + NB, 0, sizeof(uint32_t), s));
    PLAN_TRY(exclusive_scan(sc, bpairs, pair_scan_b, static_cast<uint64_t>(NB) + 1, s));
    PLAN_TRY(exclusive_scan(sc, bruns, run_scan_b, static_cast<uint64_t>(NB) + 1, s));
    hipLaunchKernelGGL(plan_run_begin_kernel, dim3(grid_for(NB)), dim3(kThreads), 0, s, NB, run_scan_b, blocks);
    PLAN_TRY(hipGetLastError());
    uint32_t hpairs_total = 0, nruns = 0;
    PLAN_TRY(hipMemcpyAsync(&hpairs_total, pair_scan_b + NB, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipMemcpyAsync(&nruns, run_scan_b + NB, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipStreamSynchronize(s));
    if (hpairs_total != U) {
        err = "internal error: pair count";
        return BW_HIP_ERR_RUNTIME;
    }

    // 4. Bin index and flat (scatter) order of the pairs.
    PLAN_ALLOC(iota, uint32_t, U);
    PLAN_ALLOC(flat_key, uint32_t, U);
    PLAN_ALLOC(pair_block, uint32_t, U);
    PLAN_ALLOC(sorted_key, uint32_t, U);
    PLAN_ALLOC(order, uint32_t, U);
    PLAN_ALLOC(bin_index, uint32_t, U);
    PLAN_ALLOC(run_start, uint32_t, static_cast<uint64_t>(U) + 1);
    PLAN_ALLOC(run_scan, uint32_t, static_cast<uint64_t>(U) + 1);
    hipLaunchKernelGGL(plan_pair_keys_kernel, dim3(grid_for(U)), dim3(kThreads), 0, s, U, L, K, pair_src, pair_part,
                       block_local, dblock_base, iota, flat_key, pair_block);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(sort_pairs(sc, static_cast<const uint32_t *>(pair_part), sorted_key, static_cast<const uint32_t *>(iota),
                        order, U, bits_for(K - 1), s));
    hipLaunchKernelGGL(plan_scatter_index_kernel, dim3(grid_for(U)), dim3(kThreads), 0, s, U, order, bin_index);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(sort_pairs(sc, static_cast<const uint32_t *>(flat_key), sorted_key, static_cast<const uint32_t *>(iota),
                        order, U, bits_for(static_cast<uint64_t>(NB) * K - 1), s));
    hipLaunchKernelGGL(plan_run_flags_kernel, dim3(grid_for(static_cast<uint64_t>(U) + 1)), dim3(kThreads), 0, s, U,
                       sorted_key, run_start);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(exclusive_scan(sc, run_start, run_scan, static_cast<uint64_t>(U) + 1, s));
    uint32_t runs_sorted = 0;
    PLAN_TRY(hipMemcpyAsync(&runs_sorted, run_scan + U, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipStreamSynchronize(s));
    if (runs_sorted != nruns) {  // the run entries below are placed by both counts
        err = "internal error: run count";
        return BW_HIP_ERR_RUNTIME;
    }
    uint16_t *psrc = sc.get<uint16_t>(U);
    uint32_t *run_flat = sc.get<uint32_t>(static_cast<uint64_t>(nruns) + NB);
    uint32_t *run_bin = sc.get<uint32_t>(static_cast<uint64_t>(nruns) + NB);
    if (!psrc || !run_flat || !run_bin) {
        err = "device allocation of the plan failed";
        return BW_HIP_ERR_NOMEM;
    }
    hipLaunchKernelGGL(plan_runs_kernel, dim3(grid_for(U)), dim3(kThreads), 0, s, U, order, pair_src, pair_block,
                       blocks, bin_index, run_start, run_scan, psrc, run_flat, run_bin);
    PLAN_TRY(hipGetLastError());
    hipLaunchKernelGGL(plan_terminators_kernel, dim3(grid_for(NB)), dim3(kThreads), 0, s, NB, blocks, pair_scan_b,
                       bpairs, run_flat, run_bin);
    PLAN_TRY(hipGetLastError());

    // 5. Edge codes in (partition, source, row) order.
    PLAN_ALLOC(epart, uint32_t, E);
    PLAN_ALLOC(ecode, uint16_t, E);
    PLAN_ALLOC(spart, uint32_t, E);
    PLAN_ALLOC(scode, uint16_t, E);
    PLAN_ALLOC(code_flag, uint32_t, static_cast<uint64_t>(E) + 1);
    PLAN_ALLOC(code_scan, uint32_t, static_cast<uint64_t>(E) + 1);
    PLAN_ALLOC(dedge_start, uint32_t, static_cast<uint64_t>(K) + 1);
    PLAN_ALLOC(dustart, uint32_t, static_cast<uint64_t>(K) + 1);
    const uint64_t ncodes = std::max<uint64_t>(ep, 8);
    uint16_t *codes = sc.get<uint16_t>(ncodes);
    if (!codes) {
        err = "device allocation of the plan failed";
        return BW_HIP_ERR_NOMEM;
    }
    PLAN_TRY(hipMemcpyAsync(dedge_start, edge_start.data(), sizeof(uint32_t) * (K + 1), hipMemcpyHostToDevice, s));
    PLAN_TRY(hipMemcpyAsync(dustart, ustart.data(), sizeof(uint32_t) * (K + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(plan_zero16_kernel, dim3(grid_for(ncodes)), dim3(kThreads), 0, s, ncodes, codes);
    PLAN_TRY(hipGetLastError());
    hipLaunchKernelGGL(plan_code_keys_kernel, dim3(grid_for(E)), dim3(kThreads), 0, s, E, L, frow, flag, epart, ecode);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(sort_pairs(sc, static_cast<const uint32_t *>(epart), spart, static_cast<const uint16_t *>(ecode), scode, E,
                        bits_for(K - 1), s));
    hipLaunchKernelGGL(plan_place_codes_kernel, dim3(grid_for(static_cast<uint64_t>(E) + 1)), dim3(kThreads), 0, s, E,
                       spart, scode, dedge_start, dustart, codes, code_flag);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(exclusive_scan(sc, code_flag, code_scan, static_cast<uint64_t>(E) + 1, s));

    // Tasks and chunks (as the host builder): each partition's edges split
    // evenly into chunks (boundaries rounded down to 64 edges), each chunk's
    // evenly into kChunkWaves tasks, empty ones dropped.  A task's pair base
    // needs the new-pair flags before its first edge: gathered from the scan.
    struct Cand {
        uint32_t k, c, tb, te;
    };
    std::vector<Cand> cand;
    std::vector<uint32_t> query;
    for (uint32_t k = 0; k < K; k++) {
        const uint64_t n = edgesK[k], nch = chunk_base[k + 1] - chunk_base[k];
        auto round64 = [](uint64_t x) { return static_cast<uint32_t>(x & ~63ull); };
        for (uint64_t c = 0; c < nch; c++) {
            const uint32_t cb = round64(n * c / nch);
            const uint32_t ce = c + 1 == nch ? static_cast<uint32_t>(n) : round64(n * (c + 1) / nch);
            for (uint32_t w = 0; w < kChunkWaves; w++) {
                const uint32_t tb = cb + round64(static_cast<uint64_t>(ce - cb) * w / kChunkWaves);
                const uint32_t te =
                    w + 1 == kChunkWaves ? ce : cb + round64(static_cast<uint64_t>(ce - cb) * (w + 1) / kChunkWaves);
                cand.push_back({k, static_cast<uint32_t>(c), tb, te});
                if (te > tb) query.push_back(ustart[k] + tb);
            }
        }
    }
    for (uint32_t k = 0; k < K; k++) query.push_back(ustart[k]);
    PLAN_ALLOC(dquery, uint32_t, query.size());
    PLAN_ALLOC(danswer, uint32_t, query.size());
    std::vector<uint32_t> answer(query.size());
    PLAN_TRY(hipMemcpyAsync(dquery, query.data(), sizeof(uint32_t) * query.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(plan_gather_kernel, dim3(grid_for(query.size())), dim3(kThreads), 0, s,
                       static_cast<uint32_t>(query.size()), code_scan, dquery, danswer);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(hipMemcpyAsync(answer.data(), danswer, sizeof(uint32_t) * answer.size(), hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipStreamSynchronize(s));
    const size_t kbase = query.size() - K;  // answers at each partition's first edge
    std::vector<Quad> tasks, chunks;
    {
        size_t qi = 0, ci = 0;
        for (uint32_t k = 0; k < K; k++) {
            const uint32_t nch = chunk_base[k + 1] - chunk_base[k];
            for (uint32_t c = 0; c < nch; c++) {
                const uint32_t t0 = static_cast<uint32_t>(tasks.size());
                for (uint32_t w = 0; w < kChunkWaves; w++, ci++) {
                    const Cand &q = cand[ci];
                    if (q.te <= q.tb) continue;
                    const uint32_t flags = answer[qi++] - answer[kbase + k];
                    tasks.push_back({edge_start[k] + q.tb, edge_start[k] + q.te, pair_start[k] - 1u + flags, k});
                }
                chunks.push_back({t0, static_cast<uint32_t>(tasks.size()), k, chunk_base[k] + c});
            }
        }
    }
    std::stable_sort(chunks.begin(), chunks.end(), [](const Quad &a, const Quad &b) { return a.y - a.x > b.y - b.x; });
    static_assert(sizeof(Quad) == sizeof(uint4), "Quad and uint4 share a layout");
    uint4 *dtasks = sc.get<uint4>(tasks.size());
    uint4 *dchunks = sc.get<uint4>(chunks.size());
    if (!dtasks || !dchunks) {
        err = "device allocation of the plan failed";
        return BW_HIP_ERR_NOMEM;
    }
    if (!tasks.empty())
        PLAN_TRY(hipMemcpyAsync(dtasks, tasks.data(), sizeof(uint4) * tasks.size(), hipMemcpyHostToDevice, s));
    if (!chunks.empty())
        PLAN_TRY(hipMemcpyAsync(dchunks, chunks.data(), sizeof(uint4) * chunks.size(), hipMemcpyHostToDevice, s));
    PLAN_TRY(hipStreamSynchronize(s));
#undef PLAN_TRY
#undef PLAN_ALLOC

    out.codes = codes;
    out.tasks = dtasks;
    out.chunks = dchunks;
    out.blocks = blocks;
    out.chunk_base = dchunk_base;
    out.run_flat = run_flat;
    out.run_bin = run_bin;
    out.psrc = psrc;
    for (void *p : {static_cast<void *>(codes), static_cast<void *>(dtasks), static_cast<void *>(dchunks),
                    static_cast<void *>(blocks), static_cast<void *>(dchunk_base), static_cast<void *>(run_flat),
                    static_cast<void *>(run_bin), static_cast<void *>(psrc)})
        sc.give_up(p);
    out.K = K;
    out.pairs = U;
    out.slots = slots;
    out.nchunks = static_cast<uint32_t>(chunks.size());
    out.nblocks = NB;
    out.ntasks = static_cast<uint32_t>(tasks.size());
    out.nruns = nruns + NB;
    out.ncodes = ncodes;
    return BW_HIP_OK;
}

// ------------------------------------------------ relabel by out-degree

namespace {

// key = ~out-degree, so an ascending stable sort puts larger degrees first
// and keeps equal degrees in id order.
__global__ __launch_bounds__(kThreads) void relabel_keys_kernel(uint32_t V, const uint32_t *off, uint32_t *key,
                                                                uint32_t *iota)
{
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; v < V;
         v += static_cast<uint64_t>(gridDim.x) * kThreads) {
        key[v] = ~(off[v + 1] - off[v]);
        iota[v] = static_cast<uint32_t>(v);
    }
}

__global__ __launch_bounds__(kThreads) void relabel_perm_kernel(uint32_t V, const uint32_t *order, uint32_t *perm)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < V;
         i += static_cast<uint64_t>(gridDim.x) * kThreads)
        perm[order[i]] = static_cast<uint32_t>(i);
}

// ndeg[i] = degree of old row order[i]; ndeg[V] = 0 (for the scan).
__global__ __launch_bounds__(kThreads) void relabel_degrees_kernel(uint32_t V, const uint32_t *off,
                                                                   const uint32_t *order, uint32_t *ndeg)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i <= V;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        uint32_t d = 0;
        if (i < V) {
            const uint32_t o = order[i];
            d = off[o + 1] - off[o];
        }
        ndeg[i] = d;
    }
}

// key[j] = (new row << 32) | new column, for new edge j (rows found by
// binary search over the new offsets within each workgroup's range).
__global__ __launch_bounds__(kThreads) void relabel_edges_kernel(uint32_t V, uint32_t E, const uint32_t *off,
                                                                 const uint32_t *cols, const uint32_t *order,
                                                                 const uint32_t *perm, const uint32_t *noff,
                                                                 uint64_t *key)
{
    __shared__ uint32_t range[2];
    for (uint64_t e0 = static_cast<uint64_t>(blockIdx.x) * kEdgesPerStep; e0 < E;
         e0 += static_cast<uint64_t>(gridDim.x) * kEdgesPerStep) {
        const uint32_t a = static_cast<uint32_t>(e0);
        const uint32_t b = static_cast<uint32_t>(std::min<uint64_t>(e0 + kEdgesPerStep, E));
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t r0 = row_of(noff, 0, V - 1, a);
            range[0] = r0;
            range[1] = row_of(noff, r0, V - 1, b - 1);
        }
        __syncthreads();
        const uint32_t r0 = range[0], r1 = range[1];
        for (uint32_t j = a + threadIdx.x; j < b; j += kThreads) {
            const uint32_t i = row_of(noff, r0, r1, j);
            const uint32_t old = off[order[i]] + (j - noff[i]);
            key[j] = (static_cast<uint64_t>(i) << 32) | perm[cols[old]];
        }
    }
}

__global__ __launch_bounds__(kThreads) void relabel_cols_kernel(uint32_t E, const uint64_t *key, uint32_t *ncols)
{
    for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; j < E;
         j += static_cast<uint64_t>(gridDim.x) * kThreads)
        ncols[j] = static_cast<uint32_t>(key[j]);
}

// One stored direction: rows permuted by `order`, columns mapped by `perm`
// and sorted within each row.
int relabel_direction(Scratch &sc, uint32_t V, uint32_t E, const uint32_t *off, const uint32_t *cols,
                      const uint32_t *order, const uint32_t *perm, uint32_t *noff, uint32_t *ncols, hipStream_t s,
                      std::string &err)
{
    uint32_t *ndeg = sc.get<uint32_t>(static_cast<uint64_t>(V) + 1);
    uint64_t *key = sc.get<uint64_t>(E);
    uint64_t *skey = sc.get<uint64_t>(E);
    if (!ndeg || !key || !skey) {
        err = "device allocation of the relabel failed";
        return BW_HIP_ERR_NOMEM;
    }
    hipLaunchKernelGGL(relabel_degrees_kernel, dim3(grid_for(static_cast<uint64_t>(V) + 1)), dim3(kThreads), 0, s, V,
                       off, order, ndeg);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = exclusive_scan(sc, ndeg, noff, static_cast<uint64_t>(V) + 1, s);
    if (e == hipSuccess && E) {
        hipLaunchKernelGGL(relabel_edges_kernel, dim3(grid_for(E, kEdgesPerStep)), dim3(kThreads), 0, s, V, E, off,
                           cols, order, perm, noff, key);
        e = hipGetLastError();
        size_t bytes = 0;
        const uint32_t bits = 32 + bits_for(V - 1);
        if (e == hipSuccess) e = rocprim::radix_sort_keys(nullptr, bytes, key, skey, E, 0, bits, s);
        void *tmp = e == hipSuccess ? sc.get<uint8_t>(bytes) : nullptr;
        if (e == hipSuccess && !tmp) e = hipErrorOutOfMemory;
        if (e == hipSuccess) e = rocprim::radix_sort_keys(tmp, bytes, key, skey, E, 0, bits, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(relabel_cols_kernel, dim3(grid_for(E)), dim3(kThreads), 0, s, E, skey, ncols);
            e = hipGetLastError();
        }
    }
    if (e != hipSuccess) {
        err = std::string("relabel: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? BW_HIP_ERR_NOMEM : BW_HIP_ERR_RUNTIME;
    }
    return BW_HIP_OK;
}

}  // namespace

namespace {

__global__ __launch_bounds__(kThreads) void check_cols_kernel(uint32_t V, uint32_t E, const uint32_t *cols,
                                                              uint32_t *bad)
{
    for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; j < E;
         j += static_cast<uint64_t>(gridDim.x) * kThreads)
        if (cols[j] >= V) *bad = 1u;
}

}  // namespace

int relabel_by_degree_device(uint32_t V, uint32_t E, const uint32_t *off, const uint32_t *cols, const uint32_t *roff,
                             const uint32_t *rcols, uint32_t *noff, uint32_t *ncols, uint32_t *nroff, uint32_t *nrcols,
                             uint32_t *perm_out, hipStream_t s, std::string &err)
{
    Scratch sc;
    if (E) {  // every column must name a vertex before it indexes the permutation
        uint32_t *bad = sc.get<uint32_t>(1), hbad = 0;
        if (!bad || hipMemsetAsync(bad, 0, sizeof(uint32_t), s) != hipSuccess) {
            err = "device allocation of the relabel failed";
            return BW_HIP_ERR_NOMEM;
        }
        hipLaunchKernelGGL(check_cols_kernel, dim3(grid_for(E)), dim3(kThreads), 0, s, V, E, cols, bad);
        if (roff) hipLaunchKernelGGL(check_cols_kernel, dim3(grid_for(E)), dim3(kThreads), 0, s, V, E, rcols, bad);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&hbad, bad, sizeof hbad, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            err = "relabel: column check failed on the device";
            return BW_HIP_ERR_RUNTIME;
        }
        if (hbad) {
            err = "column id out of range";
            return BW_HIP_ERR_ARG;
        }
    }
    uint32_t *key = sc.get<uint32_t>(V), *skey = sc.get<uint32_t>(V), *iota = sc.get<uint32_t>(V);
    uint32_t *order = sc.get<uint32_t>(V);
    uint32_t *perm = perm_out ? perm_out : sc.get<uint32_t>(V);
    if (!key || !skey || !iota || !order || !perm) {
        err = "device allocation of the relabel failed";
        return BW_HIP_ERR_NOMEM;
    }
    hipLaunchKernelGGL(relabel_keys_kernel, dim3(grid_for(V)), dim3(kThreads), 0, s, V, off, key, iota);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = sort_pairs(sc, static_cast<const uint32_t *>(key), skey, static_cast<const uint32_t *>(iota),
                                        order, V, 32, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(relabel_perm_kernel, dim3(grid_for(V)), dim3(kThreads), 0, s, V, order, perm);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        err = std::string("relabel: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? BW_HIP_ERR_NOMEM : BW_HIP_ERR_RUNTIME;
    }
    BW_TRY(relabel_direction(sc, V, E, off, cols, order, perm, noff, ncols, s, err));
    if (roff) BW_TRY(relabel_direction(sc, V, E, roff, rcols, order, perm, nroff, nrcols, s, err));
    if (hipStreamSynchronize(s) != hipSuccess) {
        err = "relabel failed on the device";
        return BW_HIP_ERR_RUNTIME;
    }
    return BW_HIP_OK;
}

}  // namespace bins
}  // namespace bw

using namespace bw;

extern "C" int bw_graph_relabel_by_degree(uint64_t V, uint64_t E, const uint32_t *offsets, const uint32_t *cols,
                                          const uint32_t *rev_offsets, const uint32_t *rev_cols, uint32_t *new_offsets,
                                          uint32_t *new_cols, uint32_t *new_rev_offsets, uint32_t *new_rev_cols,
                                          uint32_t *perm, void *stream)
{
    if (V == 0 || V >= 0x7FFFFFFFull || E >= 0xFFFFFFFFull)
        return fail(BW_HIP_ERR_ARG, "relabel needs 0 < V < 2^31 and E < 2^32 - 1");
    if (!offsets || !new_offsets || (E && (!cols || !new_cols)))
        return fail(BW_HIP_ERR_ARG, "bw_graph_relabel_by_degree: missing forward arrays");
    if ((rev_offsets == nullptr) != (new_rev_offsets == nullptr) ||
        (rev_offsets && E && (!rev_cols || !new_rev_cols)))
        return fail(BW_HIP_ERR_ARG, "bw_graph_relabel_by_degree: reverse arrays must be given in and out together");
    std::string err;
    const int st = bins::relabel_by_degree_device(static_cast<uint32_t>(V), static_cast<uint32_t>(E), offsets, cols,
                                                  rev_offsets, rev_cols, new_offsets, new_cols, new_rev_offsets,
                                                  new_rev_cols, perm, static_cast<hipStream_t>(stream), err);
    if (st != BW_HIP_OK) return fail(st, err);
    return BW_HIP_OK;
}
