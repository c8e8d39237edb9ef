// Fused pull PageRank for gfx950 ("vertex-pull-nodiv-fused").
//
// Same recurrence as pagerank/vertex_pull_nodiv.cu + consolidateRankNoDiv
// (merijn/Belewitte): contributions c[u] = rank[u]/outdeg(u) are gathered
// along the reverse CSR and the consolidate step is fused into the gather,
// so one iteration reads
//   offsets 4(V+1) + columns 4E + gathered contributions 4E
//   + out-degree 4V + own old contribution 4V, and writes 4V
// = 8E + 16V + 4 bytes, the HBM roofline of the iteration.
//
// Mapping (CSR-Adaptive, one 256-thread workgroup per work item), two
// launches per iteration:
//   1. pieces kernel: rows with more than kTileNnz in-edges are split into
//      pieces of kTileNnz edges; each workgroup writes the Kahan partial of
//      one piece.  One extra workgroup reduces the previous launch's
//      per-workgroup FP64 partials (dangling mass, L1 residual) in a fixed
//      order.
//   2. rows kernel: one workgroup per long row combines its piece partials
//      in piece order and finalises the row; one workgroup per row tile
//      (<= kRowsMax rows, <= kTileNnz edges) gathers the tile's
//      contributions into LDS -- each thread issues its kPerThread column
//      loads (streamed, non-temporal) and then its kPerThread gathers back
//      to back -- and a power-of-two group of lanes per row reduces each row
//      from LDS.
// The exchange between workgroups happens only across the kernel boundary:
// on gfx950 a device-scope fence inside a kernel writes back and
// invalidates the XCD's L2, which would evict the contributions every
// other workgroup is gathering.  Every sum is Kahan-compensated and
// combined in a fixed order, so ranks are bit-reproducible run to run.

#include "common.hpp"
#include "pagerank_bins_plan.hpp"
#include "pagerank_plan.hpp"

#include <algorithm>
#include <cmath>
#include <thread>
#include <type_traits>
#include <vector>

namespace bw {
namespace {

constexpr int kBlock = 256;
constexpr int kPerThread = 16;
constexpr uint32_t kTileNnz = kBlock * kPerThread;  // 4096
#ifndef BW_PR_ROWS_MAX
#define BW_PR_ROWS_MAX 1024
#endif
constexpr uint32_t kRowsMax = BW_PR_ROWS_MAX;
constexpr uint32_t kBigRow = 64;  // rows longer than this get a whole wave
constexpr float kDamp = 0.85f;
constexpr uint32_t kInitBlocks = 1024;

struct LongRow {
    uint32_t row, pieces, base, pad;
};

struct Kahan {
    float s = 0.0f, c = 0.0f;  // value = s - c
    __device__ __forceinline__ void add(float x)
    {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// (s1 - c1) + (s2 - c2) keeping the rounding error of the head sum (TwoSum).
__device__ __forceinline__ void kahan_merge(float &s, float &c, float s2, float c2)
{
    const float t = s + s2;
    const float bp = t - s;
    const float err = (s - (t - bp)) + (s2 - bp);
    s = t;
    c = (c + c2) - err;
}

struct IterArgs {
    uint32_t V;
    const uint32_t *roff;
    const uint32_t *rcols;
    const uint32_t *deg;
    const float *cin;
    float *cout;
    const uint4 *pieces;     // (row, partial index, edge begin, edge end)
    const LongRow *lrows;
    const uint4 *tiles;      // (first row, end row, edge begin, edge end)
    float2 *partials;        // one Kahan partial per piece
    uint32_t npieces, nlong, ntiles;
    const uint4 *windows;    // (first tile, tiles, edge begin, edge end), source-sorted mode
    uint32_t nwin;
    const uint32_t *scols;   // per window: column ids sorted by source
    const uint16_t *sperm;   // per window: CSR position (within the window) of each sorted column
    const double2 *prev;     // per-workgroup (residual, dangling) of the previous launch
    uint32_t nprev;
    double *resid_out;       // reduced residual of the previous launch
    double *dangl_out;       // reduced dangling mass of the previous launch
    const double *dangl;     // dangling mass of the ranks being gathered
    double2 *part;           // this launch's per-workgroup (residual, dangling)
    int mode;
    int not_last;
    int cin_divided;
};

// The new rank from the gathered sum: DROP follows consolidateRankNoDiv,
// pagerank/pagerank.cu:154 (float base, fma); UNIFORM adds the dangling mass
// D spread over all vertices, in FP64.
__device__ __forceinline__ float new_rank(int mode, uint32_t V, float sum, double D)
{
    if (mode == BW_PR_DANGLING_DROP) {
        const float baseF = (1.0f - kDamp) / static_cast<float>(V);
        return __builtin_fmaf(kDamp, sum, baseF);
    }
    const double d = static_cast<double>(kDamp);
    return static_cast<float>((1.0 - d) / static_cast<double>(V) +
                              d * (static_cast<double>(sum) + D / static_cast<double>(V)));
}

__device__ __forceinline__ void finalise(const IterArgs &a, uint32_t v, float sum, uint32_t dv, float old,
                                         double D, double &resid, double &dangl)
{
    const float nr = new_rank(a.mode, a.V, sum, D);
    double oldRank = static_cast<double>(old);
    if (a.cin_divided && dv != 0) oldRank *= static_cast<double>(dv);
    resid += fabs(static_cast<double>(nr) - oldRank);
    if (dv == 0) dangl += static_cast<double>(nr);
    a.cout[v] = (dv != 0 && a.not_last) ? nr / static_cast<float>(dv) : nr;
}

// Contributions of edges [e, e_end) that fall to this thread (edge
// e + tid + kBlock*k for k < kPerThread): all column loads, then all gathers.
// With kHot > 0, sources below kHot are read from the workgroup's LDS copy
// of their contributions (`hot`) instead of through the L2: on a graph
// relabelled by decreasing out-degree those are the most gathered sources.
// `t` is the thread's index within its group of kBlock threads.
template<uint32_t kHot>
__device__ __forceinline__ void gather_tile(const IterArgs &a, uint32_t e, uint32_t e_end, float (&val)[kPerThread],
                                            const float *hot, uint32_t t)
{
    uint32_t col[kPerThread];
#pragma unroll
    for (int k = 0; k < kPerThread; k++) {
        const uint32_t j = e + t + kBlock * k;
        col[k] = j < e_end ? __builtin_nontemporal_load(&a.rcols[j]) : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int k = 0; k < kPerThread; k++) {
        if (kHot > 0 && col[k] < kHot) val[k] = hot[col[k]];
        else val[k] = col[k] != 0xFFFFFFFFu ? a.cin[col[k]] : 0.0f;
    }
}

// Block-wide (sum, compensation) merge in a fixed order; valid in thread 0.
__device__ __forceinline__ void block_kahan(float &s, float &c, float (*wsum)[2])
{
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const float s2 = __shfl_xor(s, o, kWave), c2 = __shfl_xor(c, o, kWave);
        kahan_merge(s, c, s2, c2);
    }
    if (lane == 0) {
        wsum[wid][0] = s;
        wsum[wid][1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s = wsum[0][0];
        c = wsum[0][1];
        for (int w = 1; w < kBlock / kWave; w++) kahan_merge(s, c, wsum[w][0], wsum[w][1]);
    }
}

// Sum of n (residual, dangling) partials in a fixed order: thread t adds
// partials t, t + kThreads, ...; then a fixed tree.  Results in thread 0.
template<int kThreads = kBlock>
__device__ __forceinline__ void reduce_partials(const double2 *p, uint32_t n, double &r, double &g, double (*red)[2])
{
    r = 0.0;
    g = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += kThreads) {
        const double2 t = p[i];
        r += t.x;
        g += t.y;
    }
    r = wave_sum_f64(r);
    g = wave_sum_f64(g);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) {
        red[wid][0] = r;
        red[wid][1] = g;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        r = 0.0;
        g = 0.0;
        for (int w = 0; w < kThreads / kWave; w++) {
            r += red[w][0];
            g += red[w][1];
        }
    }
}

// The last workgroup reduces the previous launch's partials; the others
// take pieces b, b + (gridDim.x - 1), ...
template<uint32_t kHot>
__global__ __launch_bounds__(kBlock) void pr_pieces_kernel(IterArgs a)
{
    __shared__ float wsum[kBlock / kWave][2];
    __shared__ double red[kBlock / kWave][2];
    __shared__ float hot[kHot > 0 ? kHot : 1];
    if (blockIdx.x == gridDim.x - 1) {
        double r, g;
        reduce_partials(a.prev, a.nprev, r, g, red);
        if (threadIdx.x == 0) {
            *a.resid_out = r;
            *a.dangl_out = g;
        }
        return;
    }
    if (kHot > 0) {
        const uint32_t n = min(kHot, a.V);
        for (uint32_t i = threadIdx.x; i < n; i += kBlock) hot[i] = a.cin[i];
        __syncthreads();
    }
    for (uint32_t p = blockIdx.x; p < a.npieces; p += gridDim.x - 1) {
        const uint4 pc = a.pieces[p];
        float val[kPerThread];
        gather_tile<kHot>(a, pc.z, pc.w, val, hot, threadIdx.x);
        Kahan k;
#pragma unroll
        for (int q = 0; q < kPerThread; q++) k.add(val[q]);
        float s = k.s, c = k.c;
        block_kahan(s, c, wsum);
        if (threadIdx.x == 0) a.partials[pc.y] = make_float2(s, c);
        __syncthreads();  // wsum is reused by the next piece
    }
}

// LDS staging of one row tile besides its gathered values: row offsets
// relative to the tile's first edge (<= kTileNnz, so 16 bits) and the list
// of big rows.  Kept small so that more workgroups fit a CU.
struct TileMeta {
    uint16_t roffs[kRowsMax + 1];
    uint32_t bigRows[kTileNnz / kBigRow];
    uint32_t nbig;
};

// Row offsets relative to the tile's first edge e0; `tid` indexes the
// group's kBlock threads.
__device__ __forceinline__ void stage_rows(const IterArgs &a, TileMeta &M, uint32_t r0, uint32_t nrows, uint32_t e0,
                                           int tid)
{
    for (uint32_t i = tid; i <= nrows; i += kBlock) M.roffs[i] = static_cast<uint16_t>(a.roff[r0 + i] - e0);
}

// Rows with more than kBigRow edges, listed in row order by the group's
// first wave.
__device__ __forceinline__ void list_big_rows(TileMeta &M, uint32_t nrows, int lane)
{
    uint32_t count = 0;
    for (uint32_t b = 0; b < nrows; b += kWave) {
        const uint32_t row = b + lane;
        const bool big = row < nrows && M.roffs[row + 1] - M.roffs[row] > kBigRow;
        const uint64_t m = __ballot(big);
        if (big) M.bigRows[count + __popcll(m & ((1ull << lane) - 1ull))] = row;
        count += __popcll(m);
    }
    if (lane == 0) M.nbig = count;
}

// Finalises the tile's rows from the values in CSR order (`vals`): the big
// rows by a whole wave each, the others by groups of tpr lanes, so one
// long row never serialises on a single lane.  Each row's Kahan sum runs
// over its edges in CSR order whatever the gather order was.
__device__ __forceinline__ void reduce_rows(const IterArgs &a, const TileMeta &M, const float *vals, uint32_t r0,
                                            uint32_t nrows, int tid, double D, double &resid, double &dangl)
{
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    // Lanes per row: the largest power of two <= 64 with nrows groups
    // fitting kBlock threads; one lane per row (looping) beyond.
    uint32_t tpr = kWave;
    while (tpr > 1 && tpr * nrows > kBlock) tpr >>= 1;
    const uint32_t sub = tid % tpr;
    for (uint32_t row = tid / tpr; row < nrows; row += kBlock / tpr) {
        const uint32_t rs = M.roffs[row], re = M.roffs[row + 1];
        if (tpr < kWave && re - rs > kBigRow) continue;  // whole group skips
        Kahan k;
        for (uint32_t j = rs + sub; j < re; j += tpr) k.add(vals[j]);
        float ps = k.s, pc = k.c;
        for (uint32_t o = tpr / 2; o > 0; o >>= 1) {
            const float s2 = __shfl_xor(ps, static_cast<int>(o), kWave);
            const float c2 = __shfl_xor(pc, static_cast<int>(o), kWave);
            kahan_merge(ps, pc, s2, c2);
        }
        if (sub == 0) finalise(a, r0 + row, ps - pc, a.deg[r0 + row], a.cin[r0 + row], D, resid, dangl);
    }
    if (tpr < kWave) {
        for (uint32_t b = wid; b < M.nbig; b += kBlock / kWave) {
            const uint32_t row = M.bigRows[b];
            const uint32_t rs = M.roffs[row], re = M.roffs[row + 1];
            Kahan k;
            for (uint32_t j = rs + lane; j < re; j += kWave) k.add(vals[j]);
            float ps = k.s, pc = k.c;
            for (int o = kWave / 2; o > 0; o >>= 1) {
                const float s2 = __shfl_xor(ps, o, kWave), c2 = __shfl_xor(pc, o, kWave);
                kahan_merge(ps, pc, s2, c2);
            }
            if (lane == 0) finalise(a, r0 + row, ps - pc, a.deg[r0 + row], a.cin[r0 + row], D, resid, dangl);
        }
    }
}

// Long-row finaliser: combines the row's piece partials in piece order.
__device__ __forceinline__ void finalise_long(const IterArgs &a, uint32_t item, double D, double &resid,
                                              double &dangl)
{
    const LongRow lr = a.lrows[item];
    float s = 0.0f, c = 0.0f;
    for (uint32_t q = 0; q < lr.pieces; q++) {
        const float2 p = a.partials[lr.base + q];
        kahan_merge(s, c, p.x, p.y);
    }
    finalise(a, lr.row, s - c, a.deg[lr.row], a.cin[lr.row], D, resid, dangl);
}

// The workgroup's (residual, dangling) partial, fixed reduction order.
template<int kThreads>
__device__ __forceinline__ void write_partial(double2 *part, double resid, double dangl, double (*red)[2])
{
    resid = wave_sum_f64(resid);
    dangl = wave_sum_f64(dangl);
    const int lane = threadIdx.x & (kWave - 1);
    if (lane == 0) {
        red[threadIdx.x / kWave][0] = resid;
        red[threadIdx.x / kWave][1] = dangl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0, q = 0.0;
        for (int w = 0; w < kThreads / kWave; w++) {
            r += red[w][0];
            q += red[w][1];
        }
        part[blockIdx.x] = make_double2(r, q);
    }
}

// Work items: nlong long-row finalisers, then the row tiles.  A workgroup
// holds kSub groups of kBlock threads that take kSub consecutive items per
// step: workgroup b takes steps b, b + gridDim.x, ... (one step each when
// the grid covers all items; persistent otherwise).  The kHot table is
// loaded once per workgroup and shared by its groups.  Every group passes
// the same barriers whatever its item (or none), so the block-wide
// barriers stay matched.
template<uint32_t kHot, int kSub>
__global__ __launch_bounds__(kBlock * kSub) void pr_rows_kernel(IterArgs a)
{
    __shared__ float vals[kSub][kTileNnz];
    __shared__ TileMeta T[kSub];
    __shared__ double red[kBlock * kSub / kWave][2];
    __shared__ float hot[kHot > 0 ? kHot : 1];

    const int g = threadIdx.x / kBlock;   // group
    const int tid = threadIdx.x % kBlock;  // thread within the group
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    TileMeta &M = T[g];
    const double D = (a.mode == BW_PR_DANGLING_UNIFORM) ? *a.dangl : 0.0;
    double resid = 0.0, dangl = 0.0;
    if (kHot > 0) {
        const uint32_t n = min(kHot, a.V);
        for (uint32_t i = threadIdx.x; i < n; i += kBlock * kSub) hot[i] = a.cin[i];
        __syncthreads();
    }

    const uint32_t nitems = a.nlong + a.ntiles;
    for (uint32_t step = blockIdx.x; step * kSub < nitems; step += gridDim.x) {
        const uint32_t item = step * kSub + g;
        const bool isTile = item >= a.nlong && item < nitems;
        uint32_t r0 = 0, nrows = 0;
        if (item < a.nlong) {
            if (tid == 0) finalise_long(a, item, D, resid, dangl);
        } else if (isTile) {
            const uint4 it = a.tiles[item - a.nlong];
            r0 = it.x;
            nrows = it.y - it.x;
            float val[kPerThread];
            gather_tile<kHot>(a, it.z, it.w, val, hot, tid);
            stage_rows(a, M, r0, nrows, it.z, tid);  // while the gathers are in flight
#pragma unroll
            for (int k = 0; k < kPerThread; k++) vals[g][tid + kBlock * k] = val[k];
        }
        __syncthreads();
        if (isTile && wid == 0) list_big_rows(M, nrows, lane);
        __syncthreads();
        if (isTile) reduce_rows(a, M, vals[g], r0, nrows, tid, D, resid, dangl);
        __syncthreads();  // the next step reuses the LDS staging
    }
    write_partial<kBlock * kSub>(a.part, resid, dangl, red);
}

// Row tiles grouped into windows of up to kWinTiles consecutive tiles with
// contiguous edges (<= kWinTiles * kTileNnz edges).  Within a window the
// plan stores the column ids sorted by source (scols) and, for each, its
// position in the window in CSR order (sperm): the gathers then run in
// source order, so the lanes of one load hit far fewer distinct cache lines
// (hub sources share lines), and each value is stored back at its CSR
// position, so every row's sum is the same as in pr_rows_kernel.
template<int kWinTiles, int kWinBatch>
__global__ __launch_bounds__(kBlock * kWinTiles) void pr_windows_kernel(IterArgs a)
{
    constexpr uint32_t kWinNnz = kWinTiles * kTileNnz;
    __shared__ float vals[kWinNnz];
    __shared__ TileMeta T[kWinTiles];
    __shared__ double red[kBlock * kWinTiles / kWave][2];

    const int g = threadIdx.x / kBlock;
    const int tid = threadIdx.x % kBlock;
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    TileMeta &M = T[g];
    const double D = (a.mode == BW_PR_DANGLING_UNIFORM) ? *a.dangl : 0.0;
    double resid = 0.0, dangl = 0.0;

    const uint32_t nitems = a.nlong + a.nwin;
    for (uint32_t item = blockIdx.x; item < nitems; item += gridDim.x) {
        bool isTile = false;
        uint32_t r0 = 0, nrows = 0, voff = 0;
        if (item < a.nlong) {
            if (threadIdx.x == 0) finalise_long(a, item, D, resid, dangl);
        } else {
            const uint4 w = a.windows[item - a.nlong];  // first tile, tiles, edge begin, edge end
            if (g < static_cast<int>(w.y)) {
                const uint4 it = a.tiles[w.x + g];
                isTile = true;
                r0 = it.x;
                nrows = it.y - it.x;
                voff = it.z - w.z;
                stage_rows(a, M, r0, nrows, it.z, tid);
            }
            // kWinBatch gathers in flight per thread and round: a smaller
            // batch holds fewer registers, so more waves fit a CU.
#pragma unroll
            for (int r = 0; r < kPerThread / kWinBatch; r++) {
                uint32_t col[kWinBatch];
                uint16_t pos[kWinBatch];
#pragma unroll
                for (int k = 0; k < kWinBatch; k++) {
                    const uint32_t j = w.z + threadIdx.x + kBlock * kWinTiles * (r * kWinBatch + k);
                    const bool in = j < w.w;
                    col[k] = in ? __builtin_nontemporal_load(&a.scols[j]) : 0xFFFFFFFFu;
                    pos[k] = in ? __builtin_nontemporal_load(&a.sperm[j]) : 0;
                }
                float val[kWinBatch];
#pragma unroll
                for (int k = 0; k < kWinBatch; k++) val[k] = col[k] != 0xFFFFFFFFu ? a.cin[col[k]] : 0.0f;
#pragma unroll
                for (int k = 0; k < kWinBatch; k++)
                    if (col[k] != 0xFFFFFFFFu) vals[pos[k]] = val[k];
            }
        }
        __syncthreads();
        if (isTile && wid == 0) list_big_rows(M, nrows, lane);
        __syncthreads();
        if (isTile) reduce_rows(a, M, vals + voff, r0, nrows, tid, D, resid, dangl);
        __syncthreads();
    }
    write_partial<kBlock * kWinTiles>(a.part, resid, dangl, red);
}

// ------------------------------------------------ binned iteration
// Propagation blocking (the plan is described in pagerank_bins_plan.hpp):
// per iteration an accumulate launch (chunks of one partition's edges into
// an LDS row vector, written out as one partial per chunk) and a finalise
// launch (partials summed in chunk order into new ranks, whose
// contributions are scattered to the bins of the next iteration).
//
// The LDS row accumulators are 64-bit fixed point (ds_add_u64): a bin holds
// its contribution c (the FP32 rank / out-degree) scaled by 2^(S - 32), and
// the accumulate adds floor(c * 2^S) as an integer, split into the integer
// and fraction parts of the stored value.  Integer addition is exact and
// associative, so a chunk's row sums do not depend on the order of the
// atomics and the ranks repeat bit for bit; each sum is below the exact one
// by less than (edges of the row in the chunk) * 2^-S, and is rounded once
// to the FP32 partial.  S = 62 - ceil(log2 M) with M the bound on a row sum
// (the rank mass: 1, or 1 + E/V after the undivided first iteration of the
// nodiv quirk), so no sum reaches 2^63.  On gfx950 LDS atomics on random
// rows of a 2^14-row table run at 4.3 per CU and clock as ds_add_u64 (the
// rate of plain 64-bit stores), 2.5 as ds_add_f64 and 0.33 as ds_add_f32.

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4u __attribute__((ext_vector_type(4))) __attribute__((aligned(4)));  // 4-byte aligned


struct BinArgs {
    uint32_t V;
    const uint16_t *codes;      // per edge: local row | new-pair flag
    const uint4 *tasks;         // (edge begin, edge end, pair base, partition)
    const uint4 *chunks;        // (task begin, task end, partition, slot)
    uint32_t nchunks;
    uint32_t npairs;            // length of bins (at least 1)
    float *bins;                // one contribution per pair, times pair_scale
    float *partials;            // [slot][rows per partition]
    const uint32_t *chunk_base; // per partition: its first slot (K + 1)
    const uint4 *blocks;        // (first vertex, end vertex, run begin, run count)
    uint32_t nblocks;
    const uint32_t *run_flat;
    const uint32_t *run_bin;
    const uint16_t *psrc;
    const uint32_t *deg;
    float *rank;                // ranks, updated in place
    const double2 *prev;        // per-workgroup (residual, dangling) of the previous finalise
    uint32_t nprev;
    double *resid_out;
    double *dangl_out;
    const double *dangl;        // dangling mass of the ranks the bins hold
    double2 *part;
    int mode;
    int phase;    // 0: initial ranks 1/V; 1: finalise from the partials
    int scatter;  // write the bins for the next iteration
    int divide;   // phase 0: contributions divided by the out-degree
    float pair_scale;     // 2^(S - 32): a bin holds contribution * pair_scale
    float partial_scale;  // 2^-S: a fixed-point row sum times this is its value
};

// Fixed-point value floor(y * 2^32) of a stored bin value y (0 <= y < 2^32).
__device__ __forceinline__ unsigned long long bin_fixed(float y)
{
    const uint32_t hi = static_cast<uint32_t>(y);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_fractf(y) * 4294967296.0f);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// The fixed-point exponent S of the binned iteration: a row sum never
// exceeds the rank mass M (1; 1 + E/V when the first iteration is not
// divided by the out-degree), and S = 62 - ceil(log2 M) keeps every sum
// and every stored value y = c * 2^(S - 32) below 2^31.
inline int bins_fixed_shift(uint64_t V, uint64_t E, bool undivided_first)
{
    int lg = 0;
    if (undivided_first) {
        const double M = 1.0 + static_cast<double>(E) / static_cast<double>(V);
        while (std::ldexp(1.0, lg) < M) lg++;
    }
    return 62 - lg;
}

// v moved across lanes by the DPP control kCtrl in the rows of kRowMask;
// 0 in the other rows (and, with kBound, where the control has no source).
template<int kCtrl, int kRowMask, bool kBound>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), kCtrl, kRowMask, 0xF, kBound));
}

// Inclusive prefix sum over the wave's 64 lanes by DPP row shifts (within
// each row of 16 lanes) and row broadcasts (lane 15 into row 1, lane 47 into
// row 3, then lane 31 into rows 2 and 3): six VALU adds, where a prefix by
// ballots takes about six instructions per bit of the summed value.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x)
{
    x += dpp_u32<0x111, 0xF, true>(x);   // row_shr:1 (lanes without a source read 0)
    x += dpp_u32<0x112, 0xF, true>(x);   // row_shr:2
    x += dpp_u32<0x114, 0xF, true>(x);   // row_shr:4
    x += dpp_u32<0x118, 0xF, true>(x);   // row_shr:8
    x += dpp_u32<0x142, 0xA, false>(x);  // row_bcast:15 into rows 1 and 3
    x += dpp_u32<0x143, 0xC, false>(x);  // row_bcast:31 into rows 2 and 3
    return x;
}

// Workgroup per chunk; the last workgroup reduces the previous finalise's
// (residual, dangling) partials.  Each wave takes whole tasks; per round a
// lane holds kLaneEdges consecutive edge codes (16-byte loads, the next
// round's already in flight), finds the pair of each edge from the
// new-pair flags (DPP wave prefix over the per-lane flag counts), reads
// the pairs' values and adds each to its row.
template<int kRowsLog, int kThreads, int kLaneEdges, int kDepth>
__global__ __launch_bounds__(kThreads, kRowsLog == 13 ? 8 : 4) void pr_bins_accumulate_kernel(BinArgs a)
{
    constexpr uint32_t kRows = 1u << kRowsLog;
    constexpr int kWaves = kThreads / kWave;
    constexpr int kVec = kLaneEdges / 8;                 // 16-byte loads per lane and round
    constexpr uint32_t kRound = kWave * kLaneEdges;      // edges per wave and round
    static_assert(kWaves == static_cast<int>(bins::kChunkWaves), "the plan makes one task per wave");
    __shared__ unsigned long long acc[kRows];
    __shared__ uint4 stask[bins::kChunkWaves];  // the chunk's task records
    __shared__ double red[kWaves][2];
    if (blockIdx.x == gridDim.x - 1) {
        double r, g;
        reduce_partials<kThreads>(a.prev, a.nprev, r, g, red);
        if (threadIdx.x == 0) {
            *a.resid_out = r;
            *a.dangl_out = g;
        }
        return;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    // The lane's codes of the round starting at edge e0 of a task ending at
    // `end`.  The loads are unconditional (no branch, so the compiler can
    // keep several rounds in flight): a group past the task's end reads the
    // task's last group again (a line already fetched) and is masked by the
    // edge count later.
    auto load_codes = [&](uint32_t e0, uint32_t end, u32x4 (&raw)[kVec]) {
#pragma unroll
        for (int q = 0; q < kVec; q++) {
            const uint32_t e = min(e0 + lane * kLaneEdges + 8 * q, (end - 1u) & ~7u);
            const uint32_t off = e << 1;  // bytes; codes < 2^31
            raw[q] = __builtin_nontemporal_load(
                reinterpret_cast<const u32x4 *>(reinterpret_cast<const char *>(a.codes) + off));
        }
    };
    for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x - 1) {
        const uint4 ch = a.chunks[c];
        for (uint32_t i = threadIdx.x; i < kRows; i += kThreads) acc[i] = 0ull;
        // Task records come from LDS in the loop: a global load there would
        // make every wait for it a wait for all the loads in flight.
        for (uint32_t i = threadIdx.x; i < ch.y - ch.x; i += kThreads) stask[i] = a.tasks[ch.x + i];
        __syncthreads();
        // The wave's rounds, over its tasks of the chunk (t = ch.x + wid,
        // t + kWaves, ...: one task per wave with the plan's even splits),
        // are software-pipelined as set out above the loop below.
        uint32_t t = __builtin_amdgcn_readfirstlane(ch.x + wid);
        if (t < ch.y) {
            uint4 tk = stask[t - ch.x];   // task of the round being staged
            uint32_t base = tk.z;
            uint32_t le = tk.x;           // edge of the round whose codes are loaded next
            uint32_t lt = t;              // and its task
            uint4 ltk = tk;
            // Loads the codes of the round at (lt, le) into raw, then steps
            // (lt, le) to the following round; false when the wave had no
            // round left.  The load is unconditional (a clamped, unused one
            // at the end), so `raw` never merges an old value with a load in
            // flight.
            auto load_next = [&](u32x4 (&raw)[kVec]) {
                load_codes(le, ltk.y, raw);
                if (lt >= ch.y) return false;
                if (le + kRound < ltk.y) {
                    le += kRound;
                } else {
                    lt = __builtin_amdgcn_readfirstlane(lt + kWaves);
                    if (lt < ch.y) {
                        ltk = stask[lt - ch.x];
                        le = ltk.x;
                    }
                    // else le stays: the loads past the wave's last round
                    // repeat its codes, which the caches still hold
                }
                return true;
            };
            uint32_t se = tk.x;           // edge of the round staged next
            // Takes `raw` as the round at se of task tk (codes kept packed,
            // two per word), computes its pairs and issues its value loads
            // into (cd, vv, nn).
            auto code_of = [](const uint32_t (&cd)[kLaneEdges / 2], int j) {
                return (cd[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
            };
            // With 8 edges per lane, a lane's pair values come from one window
            // of 8 consecutive pairs: 2 vector loads instead of 8 gathers,
            // which the vector memory pipeline and the L2 serve at a far
            // lower request rate (at R-MAT scale 22 the iteration went from
            // 125 to 108 us with it).
            constexpr bool kWindow = kLaneEdges == 8;
            auto stage = [&](const u32x4 (&raw)[kVec], uint32_t (&cd)[kLaneEdges / 2], float (&vv)[kLaneEdges],
                             uint32_t &nn, uint32_t &fl, bool &full) {
                const uint32_t e = se + lane * kLaneEdges;
                full = se + kRound <= tk.y;  // wave-uniform: every lane has kLaneEdges edges
                nn = e < tk.y ? min(static_cast<uint32_t>(kLaneEdges), tk.y - e) : 0u;
#pragma unroll
                for (int q = 0; q < kVec; q++)
#pragma unroll
                    for (int h = 0; h < 4; h++) cd[4 * q + h] = raw[q][h];
                // The lane's new-pair flags as a bit mask (edge j = bit j), cut
                // to its edge count: flags past it are not counted, so a
                // masked load repeats the lane's last pair (no new line).
                uint32_t f = 0;
#pragma unroll
                for (int j = 0; j < kLaneEdges; j++) f |= ((code_of(cd, j) >> 15) & 1u) << j;
                f &= (1u << nn) - 1u;  // nn <= 16
                const uint32_t cnt = static_cast<uint32_t>(__popc(f));
                // The lane's first pair: the flags of the lanes before it
                // (R-MAT scale 22: 99.5 -> 95.5 us per iteration against a
                // prefix by ballots of the count's bits).
                const uint32_t incl = wave_inclusive_scan(cnt);
                const uint32_t pre = incl - cnt;
                const uint32_t tot = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), kWave - 1));
                const uint32_t p0 = base + pre;
                fl = f;
                if constexpr (kWindow) {
                    // The lane's 8 edges use at most 8 consecutive pairs from
                    // that of its first edge: two 16-byte loads (the bins are
                    // padded by 8 values) instead of 8 gathers.
                    const uint32_t q0 = min(p0 + (f & 1u), a.npairs - 1);
                    const f32x4u *w = reinterpret_cast<const f32x4u *>(reinterpret_cast<const char *>(a.bins) + (q0 << 2));
                    const f32x4u w0 = w[0], w1 = w[1];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        vv[i] = w0[i];
                        vv[4 + i] = w1[i];
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < kLaneEdges; j++) {
                        const uint32_t p = p0 + static_cast<uint32_t>(__popc(f & ((2u << j) - 1u)));
                        // 32-bit byte offset from the base (pairs < 2^30): the
                        // load takes the base in SGPRs, no 64-bit address per lane
                        const uint32_t off = min(p, a.npairs - 1) << 2;
                        vv[j] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(a.bins) + off);
                    }
                }
                base += tot;
                se += kRound;
                if (se >= tk.y) {
                    t = __builtin_amdgcn_readfirstlane(t + kWaves);
                    if (t < ch.y) {
                        tk = stask[t - ch.x];
                        base = tk.z;
                        se = tk.x;
                    }
                }
            };
            // Adds a staged round.  Masked edges add nothing: their codes
            // repeat one group (the loads past a task's end are clamped to
            // its last group), so adding zeros there would pile every masked
            // lane onto the same few LDS words.  The mask is applied here, a
            // round after the loads, so that waiting for them never waits for
            // the next round's.
            auto add_round = [&](const uint32_t (&cd)[kLaneEdges / 2], const float (&vv)[kLaneEdges], uint32_t nn,
                                 uint32_t fl, auto masked) {
                constexpr bool kMasked = decltype(masked)::value;
                // Window: edge j's pair is window entry o_j, the number of
                // new-pair flags of edges 1..j (0 <= o_j <= j).  With 2^14 rows
                // (4 waves per SIMD, registers to spare), w is the window
                // shifted left by o_j: each flag shifts it by one, entry 0 is
                // the edge's value, and edge j shifts entries 0 .. 7 - j only,
                // since edge j + m reads no entry past m.  That takes half the
                // VALU instructions of comparing o_j with each candidate, but
                // more registers than the 2^13-row kernel's 64 (8 waves per
                // SIMD), which keeps the comparisons.
                constexpr bool kShift = kWindow && kRowsLog == 14;
                float w[kLaneEdges];
#pragma unroll
                for (int i = 0; i < kLaneEdges; i++) w[i] = vv[i];
#pragma unroll
                for (int j = 0; j < kLaneEdges; j++) {
                    float x = vv[j];
                    if constexpr (kShift) {
                        if (j > 0) {
                            const bool s = (fl >> j) & 1u;
#pragma unroll
                            for (int i = 0; i < kLaneEdges - j; i++) w[i] = s ? w[i + 1] : w[i];
                        }
                        x = w[0];
                    } else if constexpr (kWindow) {
                        const uint32_t o = static_cast<uint32_t>(__popc((fl >> 1) & ((1u << j) - 1u)));
                        x = vv[0];
#pragma unroll
                        for (int i = 1; i <= j; i++) x = o == static_cast<uint32_t>(i) ? vv[i] : x;
                    }
                    if (!kMasked || j < static_cast<int>(nn))
                        atomicAdd(&acc[code_of(cd, j) & (kRows - 1)], bin_fixed(x));
                }
            };
            // With 2^14 rows a full round (every lane kLaneEdges edges: all
            // rounds but a task's last) adds without the per-edge test of the
            // edge count and its exec-mask branches (95.5 -> 94.5 us per
            // iteration at R-MAT scale 22).  The 2^13-row kernel, at 8 waves
            // per SIMD, ran twice as slow with the second copy of the loop.
            auto add = [&](const uint32_t (&cd)[kLaneEdges / 2], const float (&vv)[kLaneEdges], uint32_t nn,
                           uint32_t fl, bool full) {
                if (kRowsLog == 14 && full) add_round(cd, vv, nn, fl, std::false_type{});
                else add_round(cd, vv, nn, fl, std::true_type{});
            };
            // kDepth rounds of pair values in flight.  Rounds R_0, R_1, ... of
            // the wave; per step r: the codes of R_{r+kDepth+1} are loaded,
            // R_r is added, and R_{r+kDepth} is staged into R_r's buffers from
            // codes loaded a step earlier (issued before the values of
            // R_{r+kDepth-1}, so waiting for them leaves those in flight).
            // Buffers are used in turn with compile-time indices (the loop is
            // unrolled over 2 x kDepth steps), so no register holding a load in
            // flight is ever copied, which would make the compiler wait for
            // it.  Stages and code loads run unconditionally: past the wave's
            // last round they see no edge (counts 0, loads clamped), so the
            // compiler's count of loads in flight is the same on every path.
            static_assert(kDepth >= 2, "the pipeline keeps at least two rounds in flight");
            u32x4 raw[2][kVec];
            uint32_t code[kDepth][kLaneEdges / 2], n[kDepth], fl[kDepth];
            bool full[kDepth];      // wave-uniform: the round of the buffer is full
            float v[kDepth][kLaneEdges];
            bool real[kDepth];      // wave-uniform: R_r of the buffer exists
            bool next_real = true;  // the round whose codes were loaded last exists
            load_next(raw[0]);
#pragma unroll
            for (int i = 0; i < kDepth; i++) {
                real[i] = next_real;
                next_real = load_next(raw[(i + 1) & 1]);
                stage(raw[i & 1], code[i], v[i], n[i], fl[i], full[i]);
            }
            for (;;) {
#pragma unroll
                for (int i = 0; i < 2 * kDepth; i++) {
                    const int cur = i % kDepth;
                    const bool have = load_next(raw[(i + kDepth + 1) & 1]);
                    add(code[cur], v[cur], n[cur], fl[cur], full[cur]);
                    if (!real[(i + 1) % kDepth]) goto wave_done;  // rounds are consecutive
                    stage(raw[(i + kDepth) & 1], code[cur], v[cur], n[cur], fl[cur], full[cur]);
                    real[cur] = next_real;
                    next_real = have;
                }
            }
        wave_done:;
        }
        __syncthreads();
        float *out = a.partials + static_cast<size_t>(ch.w) * kRows;
        for (uint32_t i = threadIdx.x; i < kRows; i += kThreads)
            out[i] = static_cast<float>(acc[i]) * a.partial_scale;
        __syncthreads();  // acc is cleared for the next chunk
    }
}

// Workgroup per scatter block: new ranks of the block's vertices (phase 1)
// or the initial ranks (phase 0), their contributions staged in LDS, then
// written to each of the block's pairs (runs in flat order, each wave a
// contiguous range of them).  Each vertex's partials are summed
// by a group of G lanes (G a power of two, as many as the block's vertex
// count leaves, at most one per partial), lane i adding partials i, i + G,
// ... in order and the group combining by a fixed xor tree.
#ifndef BW_BINS_SCATTER_UNROLL
#define BW_BINS_SCATTER_UNROLL 8  // pair stores per lane and step of the scatter
#endif
template<int kRowsLog, int kThreads>
__global__ __launch_bounds__(kThreads) void pr_bins_finalise_kernel(BinArgs a)
{
    constexpr int kScatterUnroll = BW_BINS_SCATTER_UNROLL;
    constexpr uint32_t kWaves = kThreads / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    __shared__ float contrib[bins::kBlockVerts];
    __shared__ uint32_t rflat[bins::kMaxRuns + 1];
    __shared__ uint32_t rbin[bins::kMaxRuns + 1];
    __shared__ double red[kThreads / kWave][2];
    const double D = (a.phase && a.mode == BW_PR_DANGLING_UNIFORM) ? *a.dangl : 0.0;
    double resid = 0.0, dangl = 0.0;
    for (uint32_t b = blockIdx.x; b < a.nblocks; b += gridDim.x) {
        const uint4 bl = a.blocks[b];
        const uint32_t k = bl.x >> kRowsLog, rbase = k << kRowsLog;
        const uint32_t c0 = a.chunk_base[k], c1 = a.chunk_base[k + 1];
        const uint32_t nv = bl.y - bl.x, ns = a.phase ? c1 - c0 : 0u;
        uint32_t G = 1;
        while (G < kWave && 2 * G * nv <= kThreads && 2 * G <= ns) G *= 2;
        const uint32_t sub = threadIdx.x & (G - 1);
        for (uint32_t v0 = 0; v0 < nv; v0 += kThreads / G) {
            const uint32_t vi = v0 + threadIdx.x / G;
            const bool live = vi < nv;
            double s = 0.0;
            if (live && a.phase) {
                // partial of slot c0 + q for vertex bl.x + vi: [c0 + q][bl.x + vi - rbase]
                const float *pv = a.partials + (static_cast<size_t>(c0) << kRowsLog) + (bl.x - rbase) + vi;
                uint32_t q = sub;
                for (; q + 3 * G < ns; q += 4 * G) {
                    float t[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) t[i] = pv[static_cast<size_t>(q + i * G) << kRowsLog];
#pragma unroll
                    for (int i = 0; i < 4; i++) s += static_cast<double>(t[i]);
                }
                for (; q < ns; q += G) s += static_cast<double>(pv[static_cast<size_t>(q) << kRowsLog]);
            }
            for (uint32_t o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, static_cast<int>(o), kWave);
            if (live && sub == 0) {
                const uint32_t v = bl.x + vi;
                const uint32_t dv = a.deg[v];
                float nr, cv;
                if (!a.phase) {
                    nr = 1.0f / static_cast<float>(a.V);
                    cv = (a.divide && dv != 0) ? nr / static_cast<float>(dv) : nr;
                } else {
                    nr = new_rank(a.mode, a.V, static_cast<float>(s), D);
                    resid += fabs(static_cast<double>(nr) - static_cast<double>(a.rank[v]));
                    cv = dv != 0 ? nr / static_cast<float>(dv) : 0.0f;
                }
                if (dv == 0) dangl += static_cast<double>(nr);
                a.rank[v] = nr;
                contrib[vi] = cv * a.pair_scale;
            }
        }
        if (a.scatter) {
            for (uint32_t i = threadIdx.x; i <= bl.w; i += kThreads) {
                rflat[i] = a.run_flat[bl.z + i];
                rbin[i] = a.run_bin[bl.z + i];
            }
            __syncthreads();
            // Wave w writes the flat range [fbeg, fend) of the block's pairs;
            // each lane finds the run of its first flat by binary search,
            // then follows the runs forward (its flats only increase).  The
            // pairs' sources are stored in flat order: sequential reads.
            const uint32_t f0 = rflat[0], total = rflat[bl.w] - f0;
            const uint32_t per = ((total + kWaves - 1) / kWaves + kWave - 1) & ~static_cast<uint32_t>(kWave - 1);
            const uint32_t fbeg = f0 + min(wid * per, total), fend = min(fbeg + per, f0 + total);
            uint32_t lo = 0;
            if (fbeg + lane < fend) {
                const uint32_t f = fbeg + lane;
                uint32_t hi = bl.w;  // rflat[lo] <= f < rflat[hi]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (rflat[mid] <= f) lo = mid;
                    else hi = mid;
                }
            }
            // The lane's current run in registers: its flat end and the bin
            // offset (first bin - first flat); LDS is read only on a crossing.
            // (Groups of 4 pairs per 16-byte store were measured 55% slower.)
            uint32_t runEnd = rflat[lo + 1], runShift = rbin[lo] - rflat[lo];
            for (uint32_t fr = fbeg; fr < fend; fr += kWave * kScatterUnroll) {
                uint32_t p[kScatterUnroll];
                uint16_t src[kScatterUnroll];
#pragma unroll
                for (int u = 0; u < kScatterUnroll; u++) {
                    const uint32_t f = fr + lane + u * kWave;
                    p[u] = 0xFFFFFFFFu;
                    src[u] = 0;
                    if (f < fend) {
                        src[u] = a.psrc[f];
                        while (runEnd <= f) {
                            lo++;
                            runEnd = rflat[lo + 1];
                            runShift = rbin[lo] - rflat[lo];
                        }
                        p[u] = runShift + f;
                    }
                }
#pragma unroll
                for (int u = 0; u < kScatterUnroll; u++)
                    if (p[u] != 0xFFFFFFFFu) a.bins[p[u]] = contrib[src[u]];
            }
        }
        __syncthreads();  // contrib and the run table are reused by the next block
    }
    write_partial<kThreads>(a.part, resid, dangl, red);
}

// Reduction of the last launch's partials (after the loop).
__global__ __launch_bounds__(kBlock) void pr_reduce_kernel(const double2 *p, uint32_t n, double *resid_out,
                                                           double *dangl_out)
{
    __shared__ double red[kBlock / kWave][2];
    double r, g;
    reduce_partials(p, n, r, g, red);
    if (threadIdx.x == 0) {
        *resid_out = r;
        *dangl_out = g;
    }
}

// Initial contributions: 1/V (pagerank/pagerank.cpp:88), divided by the
// out-degree unless the reference's undivided first iteration is asked for;
// per-block dangling partials go to part[blockIdx].
__global__ __launch_bounds__(kBlock) void pr_fused_init_kernel(uint32_t V, const uint32_t *deg, float *c,
                                                               int divide, double2 *part)
{
    __shared__ double red[kBlock / kWave];
    const float init = 1.0f / static_cast<float>(V);
    double dangl = 0.0;
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < V; v += gridDim.x * kBlock) {
        const uint32_t dv = deg[v];
        c[v] = (divide && dv != 0) ? init / static_cast<float>(dv) : init;
        if (dv == 0) dangl += static_cast<double>(init);
    }
    const double t = block_sum_f64(dangl, red);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(0.0, t);
}

}  // namespace
}  // namespace bw

using namespace bw;

struct bw_pr_plan {
    uint32_t V = 0, E = 0;
    uint32_t ntiles = 0, nlong = 0, npieces = 0;
    uint32_t hot = 0;        // LDS hot-table size of the rows kernel (0, 8192, 16384)
    uint32_t rows_grid = 0;  // rows-kernel workgroups (0: one per step)
    uint32_t sub = 1;        // row tiles per rows-kernel workgroup and step (1, 2 or 4)
    uint32_t sorted = 0;     // rows phase by source-sorted windows of this many tiles (0: off)
    uint32_t sorted_width = 0;  // tiles per window of the built windows
    uint32_t win_batch = 16;    // gathers in flight per thread in the window kernel (8 or 16)
    const uint32_t *rcols = nullptr;  // the graph's reverse columns (for building the windows)
    std::vector<uint4> tiles_host;
    uint4 *windows = nullptr;
    uint32_t nwin = 0;
    uint32_t *scols = nullptr;
    uint16_t *sperm = nullptr;
    uint4 *pieces = nullptr;
    LongRow *lrows = nullptr;
    uint4 *tiles = nullptr;
    float2 *partials = nullptr;
    uint32_t *deg = nullptr;
    double2 *part = nullptr;  // max(nlong + ntiles, kInitBlocks) per-workgroup partials
    const void *roff = nullptr;  // the graph's reverse offsets (for building the bins)
    // Binned mode (bw_pr_plan_bin): 2^bin_log2 rows per partition, 0 = off.
    struct Bins {
        uint32_t log2 = 0, chunk_tasks = 0, nchunks = 0, nblocks = 0, pairs = 0, slots = 0, K = 0;
        uint32_t ntasks = 0, nruns = 0;  // task records; run entries (terminators included)
        uint64_t ncodes = 0;             // edge codes, padding included
        uint32_t lane_edges = 8;  // edge codes per lane and round of the accumulate kernel
        uint16_t *codes = nullptr;
        uint4 *tasks = nullptr, *chunks = nullptr, *blocks = nullptr;
        uint32_t *chunk_base = nullptr, *run_flat = nullptr, *run_bin = nullptr;
        uint16_t *psrc = nullptr;
        float *vals = nullptr, *partials = nullptr;
        double2 *part = nullptr;
        void release()
        {
            for (void *p : {static_cast<void *>(codes), static_cast<void *>(tasks), static_cast<void *>(chunks),
                            static_cast<void *>(blocks), static_cast<void *>(chunk_base),
                            static_cast<void *>(run_flat), static_cast<void *>(run_bin), static_cast<void *>(psrc),
                            static_cast<void *>(vals), static_cast<void *>(partials), static_cast<void *>(part)})
                if (p) (void)hipFree(p);
            *this = Bins();
        }
    } bins;
    bool binned = false;
    ~bw_pr_plan()
    {
        for (void *p : {static_cast<void *>(pieces), static_cast<void *>(lrows), static_cast<void *>(tiles),
                        static_cast<void *>(partials), static_cast<void *>(deg), static_cast<void *>(part),
                        static_cast<void *>(windows), static_cast<void *>(scols), static_cast<void *>(sperm)})
            if (p) (void)hipFree(p);
        bins.release();
    }
};

extern "C" int bw_pr_plan_create(bw_pr_plan **out, uint64_t V, uint64_t E, const void *rev_offsets,
                                 int32_t offset_bytes, const uint32_t *rev_cols)
{
    if (!out || !rev_offsets || (E && !rev_cols)) return fail(BW_HIP_ERR_ARG, "bw_pr_plan_create: invalid argument");
    if (offset_bytes != 4) return fail(BW_HIP_ERR_ARG, "the fused kernel reads uint32 offsets");
    if (V == 0 || V >= 0x7FFFFFFFull || E >= 0xFFFFFFFFull)
        return fail(BW_HIP_ERR_ARG, "fused PageRank needs 0 < V < 2^31 and E < 2^32 - 1");
    *out = nullptr;
    std::vector<uint32_t> off(V + 1);
    BW_HIP_CHECK(hipMemcpy(off.data(), rev_offsets, sizeof(uint32_t) * (V + 1), hipMemcpyDeviceToHost));
    if (off[V] != E) return fail(BW_HIP_ERR_ARG, "reverse offsets do not end at the edge count");

    std::vector<uint4> pieces, tiles;
    std::vector<LongRow> lrows;
    uint64_t r0 = 0;
    auto close_tile = [&](uint64_t r1) {
        if (r1 > r0)
            tiles.push_back(make_uint4(static_cast<uint32_t>(r0), static_cast<uint32_t>(r1), off[r0], off[r1]));
    };
    for (uint64_t v = 0; v < V; v++) {
        const uint32_t n = off[v + 1] - off[v];
        if (n > kTileNnz) {
            close_tile(v);
            const uint32_t np = (n + kTileNnz - 1) / kTileNnz;
            const uint32_t base = static_cast<uint32_t>(pieces.size());
            lrows.push_back({static_cast<uint32_t>(v), np, base, 0});
            for (uint32_t p = 0; p < np; p++) {
                const uint32_t s = off[v] + p * kTileNnz;
                pieces.push_back(make_uint4(static_cast<uint32_t>(v), base + p, s, std::min(s + kTileNnz, off[v + 1])));
            }
            r0 = v + 1;
            continue;
        }
        if (v - r0 >= kRowsMax || off[v + 1] - off[r0] > kTileNnz) {
            close_tile(v);
            r0 = v;
        }
    }
    close_tile(V);
    if (tiles.size() + lrows.size() >= 0x7FFFFFFFull || pieces.size() >= 0x7FFFFFFFull)
        return fail(BW_HIP_ERR_ARG, "graph needs too many tiles");

    auto plan = new bw_pr_plan();
    plan->rcols = rev_cols;
    plan->roff = rev_offsets;
    plan->tiles_host = tiles;
    plan->V = static_cast<uint32_t>(V);
    plan->E = static_cast<uint32_t>(E);
    plan->ntiles = static_cast<uint32_t>(tiles.size());
    plan->nlong = static_cast<uint32_t>(lrows.size());
    plan->npieces = static_cast<uint32_t>(pieces.size());
    auto cleanup = [&](int st) {
        delete plan;
        return st;
    };
    const size_t npart = std::max<size_t>(tiles.size() + lrows.size(), kInitBlocks);
    auto n1 = [](size_t n) { return std::max<size_t>(n, 1); };
    if (hipMalloc(&plan->pieces, sizeof(uint4) * n1(pieces.size())) != hipSuccess ||
        hipMalloc(&plan->lrows, sizeof(LongRow) * n1(lrows.size())) != hipSuccess ||
        hipMalloc(&plan->tiles, sizeof(uint4) * n1(tiles.size())) != hipSuccess ||
        hipMalloc(&plan->partials, sizeof(float2) * n1(pieces.size())) != hipSuccess ||
        hipMalloc(&plan->deg, sizeof(uint32_t) * V) != hipSuccess ||
        hipMalloc(&plan->part, sizeof(double2) * npart) != hipSuccess)
        return cleanup(fail(BW_HIP_ERR_NOMEM, "device allocation of the PageRank plan failed"));
    auto up = [](void *dst, const void *src, size_t bytes) {
        return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(plan->pieces, pieces.data(), sizeof(uint4) * pieces.size()) ||
        !up(plan->lrows, lrows.data(), sizeof(LongRow) * lrows.size()) ||
        !up(plan->tiles, tiles.data(), sizeof(uint4) * tiles.size()))
        return cleanup(fail(BW_HIP_ERR_RUNTIME, "upload of the PageRank plan failed"));
    int st = bw_pr_zero_degrees(V, plan->deg, nullptr);
    if (st == BW_HIP_OK) st = bw_pr_degrees_edge_list(E, rev_cols, plan->deg, nullptr);
    if (st == BW_HIP_OK && hipDeviceSynchronize() != hipSuccess)
        st = fail(BW_HIP_ERR_RUNTIME, "PageRank plan construction failed");
    if (st != BW_HIP_OK) return cleanup(st);
    *out = plan;
    return BW_HIP_OK;
}

extern "C" int bw_pr_plan_destroy(bw_pr_plan *plan)
{
    delete plan;
    return BW_HIP_OK;
}

extern "C" int bw_pr_plan_info(const bw_pr_plan *plan, uint32_t *ntiles, uint32_t *nlong, uint32_t *npieces)
{
    if (!plan) return fail(BW_HIP_ERR_ARG, "plan is NULL");
    if (ntiles) *ntiles = plan->ntiles;
    if (nlong) *nlong = plan->nlong;
    if (npieces) *npieces = plan->npieces;
    return BW_HIP_OK;
}

extern "C" int bw_pr_plan_tune(bw_pr_plan *plan, uint32_t hot_vertices, uint32_t rows_grid,
                               uint32_t tiles_per_workgroup)
{
    if (!plan) return fail(BW_HIP_ERR_ARG, "plan is NULL");
    const uint32_t h = hot_vertices, t = tiles_per_workgroup;
    // The combinations whose LDS fits 160 KB per workgroup.
    const bool ok = (t == 1 && (h == 0 || h == 8192 || h == 16384)) ||
                    (t == 2 && (h == 0 || h == 8192 || h == 16384)) || (t == 4 && (h == 0 || h == 8192));
    if (!ok)
        return fail(BW_HIP_ERR_ARG, "unsupported launch shape: hot_vertices 0/8192/16384 with 1 or 2 tiles per "
                                    "workgroup, 0/8192 with 4");
    plan->hot = h;
    plan->rows_grid = rows_grid;
    plan->sub = t;
    return BW_HIP_OK;
}

extern "C" int bw_pr_plan_sort_windows(bw_pr_plan *plan, int32_t tiles_per_window, int32_t gathers_per_round)
{
    if (!plan) return fail(BW_HIP_ERR_ARG, "plan is NULL");
    if (gathers_per_round != 8 && gathers_per_round != 16)
        return fail(BW_HIP_ERR_ARG, "gathers_per_round must be 8 or 16");
    plan->win_batch = static_cast<uint32_t>(gathers_per_round);
    if (tiles_per_window == 0) {
        plan->sorted = 0;
        return BW_HIP_OK;
    }
    if (tiles_per_window != 1 && tiles_per_window != 2 && tiles_per_window != 4)
        return fail(BW_HIP_ERR_ARG, "tiles_per_window must be 0, 1, 2 or 4");
    const uint32_t width = static_cast<uint32_t>(tiles_per_window);
    if (plan->windows && plan->sorted_width != width) {
        (void)hipFree(plan->windows);
        (void)hipFree(plan->scols);
        (void)hipFree(plan->sperm);
        plan->windows = nullptr;
        plan->scols = nullptr;
        plan->sperm = nullptr;
    }
    if (!plan->windows) {
        // Windows: up to `width` consecutive tiles with contiguous edges.
        const std::vector<uint4> &t = plan->tiles_host;
        std::vector<uint4> win;
        for (size_t i = 0; i < t.size();) {
            size_t j = i + 1;
            while (j < t.size() && j - i < width && t[j].z == t[j - 1].w && t[j].w - t[i].z <= width * kTileNnz)
                j++;
            win.push_back(make_uint4(static_cast<uint32_t>(i), static_cast<uint32_t>(j - i), t[i].z, t[j - 1].w));
            i = j;
        }
        const uint64_t E = plan->E;
        std::vector<uint32_t> cols(E), sc(std::max<uint64_t>(E, 1), 0);
        std::vector<uint16_t> sp(std::max<uint64_t>(E, 1), 0);
        if (E) BW_HIP_CHECK(hipMemcpy(cols.data(), plan->rcols, sizeof(uint32_t) * E, hipMemcpyDeviceToHost));
        // Sort each window's edges by source (ties by position), in parallel.
        const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned th = 0; th < nth; th++)
            pool.emplace_back([&, th] {
                std::vector<uint64_t> key;
                for (size_t w = th; w < win.size(); w += nth) {
                    const uint32_t e0 = win[w].z, n = win[w].w - win[w].z;
                    key.resize(n);
                    for (uint32_t k = 0; k < n; k++) key[k] = (static_cast<uint64_t>(cols[e0 + k]) << 16) | k;
                    std::sort(key.begin(), key.end());
                    for (uint32_t k = 0; k < n; k++) {
                        sc[e0 + k] = static_cast<uint32_t>(key[k] >> 16);
                        sp[e0 + k] = static_cast<uint16_t>(key[k] & 0xFFFFu);
                    }
                }
            });
        for (auto &th : pool) th.join();
        uint4 *dw = nullptr;
        uint32_t *dsc = nullptr;
        uint16_t *dsp = nullptr;
        auto release = [&] {
            if (dw) (void)hipFree(dw);
            if (dsc) (void)hipFree(dsc);
            if (dsp) (void)hipFree(dsp);
        };
        if (hipMalloc(&dw, sizeof(uint4) * std::max<size_t>(win.size(), 1)) != hipSuccess ||
            hipMalloc(&dsc, sizeof(uint32_t) * sc.size()) != hipSuccess ||
            hipMalloc(&dsp, sizeof(uint16_t) * sp.size()) != hipSuccess) {
            release();
            return fail(BW_HIP_ERR_NOMEM, "device allocation of the sorted windows failed");
        }
        if ((!win.empty() && hipMemcpy(dw, win.data(), sizeof(uint4) * win.size(), hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(dsc, sc.data(), sizeof(uint32_t) * sc.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dsp, sp.data(), sizeof(uint16_t) * sp.size(), hipMemcpyHostToDevice) != hipSuccess) {
            release();
            return fail(BW_HIP_ERR_RUNTIME, "upload of the sorted windows failed");
        }
        plan->windows = dw;
        plan->scols = dsc;
        plan->sperm = dsp;
        plan->nwin = static_cast<uint32_t>(win.size());
        plan->sorted_width = width;
    }
    plan->sorted = width;
    return BW_HIP_OK;
}

namespace {

int check_bin_args(const bw_pr_plan *plan, int32_t rows_log2, int32_t chunk_tasks, int32_t lane_edges)
{
    if (!plan) return fail(BW_HIP_ERR_ARG, "plan is NULL");
    if (rows_log2 != 0 && rows_log2 != 13 && rows_log2 != 14)
        return fail(BW_HIP_ERR_ARG, "rows_log2 must be 0, 13 or 14");
    if (rows_log2 == 0) return BW_HIP_OK;
    if (chunk_tasks < 1 || chunk_tasks > 4096) return fail(BW_HIP_ERR_ARG, "chunk_tasks must be in [1, 4096]");
    if (lane_edges != 8 && lane_edges != 16) return fail(BW_HIP_ERR_ARG, "lane_edges must be 8 or 16");
    return BW_HIP_OK;
}

// The per-iteration buffers of a plan with `pairs` pairs and `slots` partials.
int alloc_bin_buffers(bw_pr_plan::Bins &nb, uint32_t rows_log2, uint32_t pairs, uint32_t slots, uint32_t nblocks)
{
    const size_t rows = size_t{1} << rows_log2;
    // 8 values of padding: the accumulate reads 8-pair windows from a lane's first pair
    if (hipMalloc(&nb.vals, sizeof(float) * (static_cast<size_t>(pairs) + 8)) != hipSuccess ||
        hipMalloc(&nb.partials, sizeof(float) * rows * std::max<size_t>(slots, 1)) != hipSuccess ||
        hipMalloc(&nb.part, sizeof(double2) * std::max<size_t>(nblocks, 1)) != hipSuccess)
        return BW_HIP_ERR_NOMEM;
    return BW_HIP_OK;
}

// Host build (bins::build_plan): the reverse CSR is copied to the host.
int bin_host(bw_pr_plan *plan, uint32_t rows_log2, uint32_t chunk_tasks, bw_pr_plan::Bins &nb)
{
    const uint32_t V = plan->V, E = plan->E;
    std::vector<uint32_t> off(static_cast<size_t>(V) + 1), cols(std::max<uint32_t>(E, 1));
    BW_HIP_CHECK(hipMemcpy(off.data(), plan->roff, sizeof(uint32_t) * (static_cast<size_t>(V) + 1),
                           hipMemcpyDeviceToHost));
    if (E) BW_HIP_CHECK(hipMemcpy(cols.data(), plan->rcols, sizeof(uint32_t) * E, hipMemcpyDeviceToHost));
    bins::Plan P;
    std::string err;
    if (!bins::build_plan(off.data(), cols.data(), V, E, rows_log2, chunk_tasks, P, err))
        return fail(BW_HIP_ERR_ARG, "binned PageRank plan: " + err);
    static_assert(sizeof(bins::Quad) == sizeof(uint4), "Quad and uint4 share a layout");
    bool ok = true;
    auto put = [&ok](auto *&dst, const auto &vec, size_t min_elems = 1) {
        using T = typename std::decay_t<decltype(vec)>::value_type;
        const size_t n = std::max(vec.size(), min_elems);
        if (!ok || hipMalloc(reinterpret_cast<void **>(&dst), sizeof(T) * n) != hipSuccess) {
            ok = false;
            return;
        }
        if (!vec.empty() && hipMemcpy(dst, vec.data(), sizeof(T) * vec.size(), hipMemcpyHostToDevice) != hipSuccess)
            ok = false;
    };
    put(nb.codes, P.codes);
    put(nb.tasks, P.tasks);
    put(nb.chunks, P.chunks);
    put(nb.blocks, P.blocks);
    put(nb.chunk_base, P.chunk_base);
    put(nb.run_flat, P.run_flat);
    put(nb.run_bin, P.run_bin);
    put(nb.psrc, P.psrc);
    if (!ok || alloc_bin_buffers(nb, rows_log2, P.pairs, P.slots, static_cast<uint32_t>(P.blocks.size())) != BW_HIP_OK)
        return fail(BW_HIP_ERR_NOMEM, "device allocation or upload of the binned PageRank plan failed");
    nb.nchunks = static_cast<uint32_t>(P.chunks.size());
    nb.nblocks = static_cast<uint32_t>(P.blocks.size());
    nb.pairs = P.pairs;
    nb.slots = P.slots;
    nb.K = P.K;
    nb.ntasks = static_cast<uint32_t>(P.tasks.size());
    nb.nruns = static_cast<uint32_t>(P.run_flat.size());
    nb.ncodes = P.codes.size();
    return BW_HIP_OK;
}

// Device build (bins::build_plan_device): same arrays, no copy to the host.
int bin_device(bw_pr_plan *plan, uint32_t rows_log2, uint32_t chunk_tasks, bw_pr_plan::Bins &nb)
{
    bins::DeviceBins d;
    std::string err;
    const int st = bins::build_plan_device(static_cast<const uint32_t *>(plan->roff), plan->rcols, plan->V, plan->E,
                                           rows_log2, chunk_tasks, nullptr, d, err);
    if (st != BW_HIP_OK) return fail(st, "binned PageRank plan: " + err);
    nb.codes = d.codes;
    nb.tasks = d.tasks;
    nb.chunks = d.chunks;
    nb.blocks = d.blocks;
    nb.chunk_base = d.chunk_base;
    nb.run_flat = d.run_flat;
    nb.run_bin = d.run_bin;
    nb.psrc = d.psrc;
    if (alloc_bin_buffers(nb, rows_log2, d.pairs, d.slots, d.nblocks) != BW_HIP_OK)
        return fail(BW_HIP_ERR_NOMEM, "device allocation of the binned PageRank buffers failed");
    nb.nchunks = d.nchunks;
    nb.nblocks = d.nblocks;
    nb.pairs = d.pairs;
    nb.slots = d.slots;
    nb.K = d.K;
    nb.ntasks = d.ntasks;
    nb.nruns = d.nruns;
    nb.ncodes = d.ncodes;
    return BW_HIP_OK;
}

int bin_impl(bw_pr_plan *plan, int32_t rows_log2, int32_t chunk_tasks, int32_t lane_edges, bool on_host)
{
    BW_TRY(check_bin_args(plan, rows_log2, chunk_tasks, lane_edges));
    if (rows_log2 == 0) {
        plan->binned = false;
        return BW_HIP_OK;
    }
    auto &B = plan->bins;
    if (!on_host && B.log2 == static_cast<uint32_t>(rows_log2) && B.chunk_tasks == static_cast<uint32_t>(chunk_tasks)) {
        B.lane_edges = static_cast<uint32_t>(lane_edges);
        plan->binned = true;
        return BW_HIP_OK;
    }
    // A new shape is built beside the current one and replaces it only once
    // it is complete: a failed call leaves the plan as it was.
    bw_pr_plan::Bins nb;
    const int st = on_host ? bin_host(plan, static_cast<uint32_t>(rows_log2), static_cast<uint32_t>(chunk_tasks), nb)
                           : bin_device(plan, static_cast<uint32_t>(rows_log2), static_cast<uint32_t>(chunk_tasks), nb);
    if (st != BW_HIP_OK) {
        nb.release();
        return st;
    }
    nb.log2 = static_cast<uint32_t>(rows_log2);
    nb.chunk_tasks = static_cast<uint32_t>(chunk_tasks);
    nb.lane_edges = static_cast<uint32_t>(lane_edges);
    B.release();
    B = nb;
    plan->binned = true;
    return BW_HIP_OK;
}

}  // namespace

extern "C" int bw_pr_plan_bin(bw_pr_plan *plan, int32_t rows_log2, int32_t chunk_tasks, int32_t lane_edges)
{
    return bin_impl(plan, rows_log2, chunk_tasks, lane_edges, false);
}

extern "C" int bw_pr_plan_bin_host(bw_pr_plan *plan, int32_t rows_log2, int32_t chunk_tasks, int32_t lane_edges)
{
    return bin_impl(plan, rows_log2, chunk_tasks, lane_edges, true);
}

extern "C" int bw_pr_plan_bin_array(const bw_pr_plan *plan, int32_t which, void *host_dst, uint64_t capacity_bytes,
                                    uint64_t *bytes)
{
    if (!plan || !bytes) return fail(BW_HIP_ERR_ARG, "invalid argument");
    if (!plan->binned) return fail(BW_HIP_ERR_ARG, "the plan is not binned");
    const auto &B = plan->bins;
    const void *src = nullptr;
    uint64_t n = 0;
    switch (which) {
    case BW_PR_BIN_CODES: src = B.codes; n = sizeof(uint16_t) * B.ncodes; break;
    case BW_PR_BIN_TASKS: src = B.tasks; n = sizeof(uint4) * B.ntasks; break;
    case BW_PR_BIN_CHUNKS: src = B.chunks; n = sizeof(uint4) * B.nchunks; break;
    case BW_PR_BIN_CHUNK_BASE: src = B.chunk_base; n = sizeof(uint32_t) * (static_cast<uint64_t>(B.K) + 1); break;
    case BW_PR_BIN_BLOCKS: src = B.blocks; n = sizeof(uint4) * B.nblocks; break;
    case BW_PR_BIN_RUN_FLAT: src = B.run_flat; n = sizeof(uint32_t) * B.nruns; break;
    case BW_PR_BIN_RUN_BIN: src = B.run_bin; n = sizeof(uint32_t) * B.nruns; break;
    case BW_PR_BIN_PSRC: src = B.psrc; n = sizeof(uint16_t) * B.pairs; break;
    default: return fail(BW_HIP_ERR_ARG, "unknown plan array " + std::to_string(which));
    }
    *bytes = n;
    if (!host_dst) return BW_HIP_OK;
    if (capacity_bytes < n) return fail(BW_HIP_ERR_ARG, "destination too small for the plan array");
    if (n) BW_HIP_CHECK(hipMemcpy(host_dst, src, n, hipMemcpyDeviceToHost));
    return BW_HIP_OK;
}

extern "C" int bw_pr_plan_bin_info(const bw_pr_plan *plan, uint32_t *pairs, uint32_t *chunks, uint32_t *blocks)
{
    if (!plan) return fail(BW_HIP_ERR_ARG, "plan is NULL");
    if (pairs) *pairs = plan->binned ? plan->bins.pairs : 0;
    if (chunks) *chunks = plan->binned ? plan->bins.nchunks : 0;
    if (blocks) *blocks = plan->binned ? plan->bins.nblocks : 0;
    return BW_HIP_OK;
}

extern "C" int bw_pr_plan_degrees(const bw_pr_plan *plan, const uint32_t **deg)
{
    if (!plan || !deg) return fail(BW_HIP_ERR_ARG, "invalid argument");
    *deg = plan->deg;
    return BW_HIP_OK;
}

// acc layout: dangling mass of r_k at [k] (k = 0 .. iterations), residual of
// iteration k+1 at [iterations + 1 + k] (k = 0 .. iterations - 1), one
// scratch slot at the end.
extern "C" uint64_t bw_pr_fused_acc_len(int32_t iterations)
{ return 2 * static_cast<uint64_t>(std::max(iterations, 0)) + 3; }

namespace {

constexpr int kBinAccThreads = 1024;
#ifndef BW_BINS_DEPTH14
#define BW_BINS_DEPTH14 2
#endif
#ifndef BW_BINS_FIN_THREADS
#define BW_BINS_FIN_THREADS 512
#endif
constexpr int kBinFinThreads = BW_BINS_FIN_THREADS;

// The binned iteration loop: initial ranks and bins, then per iteration one
// accumulate and one finalise launch; ranks are updated in place.
int run_binned(const bw_pr_plan *plan, int32_t iterations, int32_t mode, int32_t nodiv_quirk, float *ranks,
               double *dangl, double *resid, double *spare, double *resid_host, float *kernel_ms, hipStream_t s)
{
    const auto &B = plan->bins;
    using Kernel = void (*)(BinArgs);
    const bool l16 = B.lane_edges == 16;
    const Kernel accK = B.log2 == 13 ? (l16 ? pr_bins_accumulate_kernel<13, kBinAccThreads, 16, 2>
                                            : pr_bins_accumulate_kernel<13, kBinAccThreads, 8, 2>)
                                     : (l16 ? pr_bins_accumulate_kernel<14, kBinAccThreads, 16, 2>
                                            : pr_bins_accumulate_kernel<14, kBinAccThreads, 8, BW_BINS_DEPTH14>);
    const Kernel finK = B.log2 == 13 ? pr_bins_finalise_kernel<13, kBinFinThreads>
                                     : pr_bins_finalise_kernel<14, kBinFinThreads>;
    BinArgs a;
    a.V = plan->V;
    a.codes = B.codes;
    a.tasks = B.tasks;
    a.chunks = B.chunks;
    a.nchunks = B.nchunks;
    a.npairs = std::max<uint32_t>(B.pairs, 1);
    a.bins = B.vals;
    a.partials = B.partials;
    a.chunk_base = B.chunk_base;
    a.blocks = B.blocks;
    a.nblocks = B.nblocks;
    a.run_flat = B.run_flat;
    a.run_bin = B.run_bin;
    a.psrc = B.psrc;
    a.deg = plan->deg;
    a.rank = ranks;
    a.prev = B.part;
    a.nprev = B.nblocks;
    a.resid_out = spare;
    a.dangl_out = spare;
    a.dangl = dangl;
    a.part = B.part;
    a.mode = mode;
    a.phase = 0;
    a.scatter = 1;
    a.divide = !(nodiv_quirk && mode == BW_PR_DANGLING_DROP);
    const int S = bins_fixed_shift(plan->V, plan->E, !a.divide);
    a.pair_scale = std::ldexp(1.0f, S - 32);
    a.partial_scale = std::ldexp(1.0f, -S);
    const uint32_t finGrid = B.nblocks, accGrid = B.nchunks + 1;
    hipLaunchKernelGGL(finK, dim3(finGrid), dim3(kBinFinThreads), 0, s, a);
    BW_LAUNCH_CHECK();
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (kernel_ms) {
        BW_HIP_CHECK(hipEventCreate(&ev0));
        BW_HIP_CHECK(hipEventCreate(&ev1));
        BW_HIP_CHECK(hipEventRecord(ev0, s));
    }
    for (int32_t it = 1; it <= iterations; it++) {
        a.resid_out = it > 1 ? resid + (it - 2) : spare;
        a.dangl_out = dangl + (it - 1);
        a.dangl = dangl + (it - 1);
        hipLaunchKernelGGL(accK, dim3(accGrid), dim3(kBinAccThreads), 0, s, a);
        BW_LAUNCH_CHECK();
        a.phase = 1;
        a.scatter = it < iterations;
        hipLaunchKernelGGL(finK, dim3(finGrid), dim3(kBinFinThreads), 0, s, a);
        BW_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(pr_reduce_kernel, dim3(1), dim3(kBlock), 0, s, B.part, finGrid, resid + (iterations - 1),
                       dangl + iterations);
    BW_LAUNCH_CHECK();
    if (kernel_ms) {
        BW_HIP_CHECK(hipEventRecord(ev1, s));
        BW_HIP_CHECK(hipEventSynchronize(ev1));
        BW_HIP_CHECK(hipEventElapsedTime(kernel_ms, ev0, ev1));
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
    }
    if (resid_host) {
        BW_HIP_CHECK(hipMemcpyAsync(resid_host, resid, sizeof(double) * iterations, hipMemcpyDeviceToHost, s));
        BW_HIP_CHECK(hipStreamSynchronize(s));
    }
    return BW_HIP_OK;
}

}  // namespace

extern "C" int bw_pr_fused_run(const bw_graph_dev *g, const bw_pr_plan *plan, int32_t iterations, int32_t mode,
                               int32_t nodiv_quirk, float *ranks, float *scratch, double *acc,
                               double *resid_host, float *kernel_ms, const bw_launch *cfg)
{
    if (!g || !plan || !ranks || !scratch || !acc || iterations < 1)
        return fail(BW_HIP_ERR_ARG, "bw_pr_fused_run: invalid argument");
    if (mode != BW_PR_DANGLING_DROP && mode != BW_PR_DANGLING_UNIFORM)
        return fail(BW_HIP_ERR_ARG, "unknown PageRank mode");
    if (g->vertex_count != plan->V || g->edge_count != plan->E || g->offset_bytes != 4 || !g->rev_offsets ||
        (plan->E && !g->rev_cols))
        return fail(BW_HIP_ERR_ARG, "graph does not match the plan");
    hipStream_t s = stream_of(cfg);
    const uint32_t V = plan->V;
    double *dangl = acc, *resid = acc + iterations + 1, *spare = acc + 2 * iterations + 2;
    if (plan->binned) return run_binned(plan, iterations, mode, nodiv_quirk, ranks, dangl, resid, spare, resid_host,
                                        kernel_ms, s);

    // Ping-pong so the last iteration lands in `ranks`.
    float *buf[2];
    buf[iterations % 2] = ranks;
    buf[(iterations + 1) % 2] = scratch;
    const int divide = !(nodiv_quirk && mode == BW_PR_DANGLING_DROP);
    const uint32_t initGrid = std::min<uint32_t>((V + kBlock - 1) / kBlock, kInitBlocks);
    hipLaunchKernelGGL(pr_fused_init_kernel, dim3(initGrid), dim3(kBlock), 0, s, V, plan->deg, buf[0], divide,
                       plan->part);
    BW_LAUNCH_CHECK();

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (kernel_ms) {
        BW_HIP_CHECK(hipEventCreate(&ev0));
        BW_HIP_CHECK(hipEventCreate(&ev1));
        BW_HIP_CHECK(hipEventRecord(ev0, s));
    }
    const uint32_t steps = plan->sorted ? plan->nlong + plan->nwin
                                        : (plan->nlong + plan->ntiles + plan->sub - 1) / plan->sub;
    const uint32_t rowsGrid = plan->rows_grid ? std::min(plan->rows_grid, steps) : steps;
    auto rowsKernel = plan->sub == 4 ? (plan->hot == 8192 ? pr_rows_kernel<8192, 4> : pr_rows_kernel<0, 4>)
                    : plan->sub == 2 ? (plan->hot == 16384 ? pr_rows_kernel<16384, 2>
                                        : plan->hot == 8192 ? pr_rows_kernel<8192, 2> : pr_rows_kernel<0, 2>)
                    : plan->hot == 16384 ? pr_rows_kernel<16384, 1>
                    : plan->hot == 8192  ? pr_rows_kernel<8192, 1>
                                         : pr_rows_kernel<0, 1>;
    auto piecesKernel = plan->hot == 16384 ? pr_pieces_kernel<16384>
                      : plan->hot == 8192  ? pr_pieces_kernel<8192>
                                           : pr_pieces_kernel<0>;
    // Pieces: persistent like the rows when a grid is set, plus the
    // reducing workgroup.
    const uint32_t piecesGrid = (plan->rows_grid ? std::min(plan->rows_grid, plan->npieces) : plan->npieces) + 1;
    uint32_t nprev = initGrid;
    for (int32_t it = 1; it <= iterations; it++) {
        IterArgs a;
        a.V = V;
        a.roff = static_cast<const uint32_t *>(g->rev_offsets);
        a.rcols = g->rev_cols;
        a.deg = plan->deg;
        a.cin = buf[(it - 1) % 2];
        a.cout = buf[it % 2];
        a.pieces = plan->pieces;
        a.lrows = plan->lrows;
        a.tiles = plan->tiles;
        a.partials = plan->partials;
        a.npieces = plan->npieces;
        a.nlong = plan->nlong;
        a.ntiles = plan->ntiles;
        a.windows = plan->windows;
        a.nwin = plan->nwin;
        a.scols = plan->scols;
        a.sperm = plan->sperm;
        a.prev = plan->part;
        a.nprev = nprev;
        a.resid_out = it > 1 ? resid + (it - 2) : spare;
        a.dangl_out = dangl + (it - 1);
        a.dangl = dangl + (it - 1);
        a.part = plan->part;
        a.mode = mode;
        a.not_last = it < iterations;
        a.cin_divided = (it > 1) || divide;
        hipLaunchKernelGGL(piecesKernel, dim3(piecesGrid), dim3(kBlock), 0, s, a);
        BW_LAUNCH_CHECK();
        if (rowsGrid) {
            if (plan->sorted) {
                using WinKernel = void (*)(IterArgs);
                const bool b8 = plan->win_batch == 8;
                WinKernel k = plan->sorted == 4 ? (b8 ? pr_windows_kernel<4, 8> : pr_windows_kernel<4, 16>)
                            : plan->sorted == 2 ? (b8 ? pr_windows_kernel<2, 8> : pr_windows_kernel<2, 16>)
                                                : (b8 ? pr_windows_kernel<1, 8> : pr_windows_kernel<1, 16>);
                hipLaunchKernelGGL(k, dim3(rowsGrid), dim3(kBlock * plan->sorted), 0, s, a);
            } else
                hipLaunchKernelGGL(rowsKernel, dim3(rowsGrid), dim3(kBlock * plan->sub), 0, s, a);
            BW_LAUNCH_CHECK();
        }
        nprev = rowsGrid;
    }
    hipLaunchKernelGGL(pr_reduce_kernel, dim3(1), dim3(kBlock), 0, s, plan->part, nprev, resid + (iterations - 1),
                       dangl + iterations);
    BW_LAUNCH_CHECK();
    if (kernel_ms) {
        BW_HIP_CHECK(hipEventRecord(ev1, s));
        BW_HIP_CHECK(hipEventSynchronize(ev1));
        BW_HIP_CHECK(hipEventElapsedTime(kernel_ms, ev0, ev1));
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
    }
    if (resid_host) {
        BW_HIP_CHECK(hipMemcpyAsync(resid_host, resid, sizeof(double) * iterations, hipMemcpyDeviceToHost, s));
        BW_HIP_CHECK(hipStreamSynchronize(s));
    }
    return BW_HIP_OK;
}
