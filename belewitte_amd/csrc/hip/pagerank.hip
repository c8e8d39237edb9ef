// PageRank iteration, degree and consolidate kernels for gfx950 and their
// C ABI entry points.
//
// Each kernel computes one reference kernel of pagerank/*.cu (merijn/
// Belewitte) with the same FP32 arithmetic per contribution: IEEE division
// pr/deg, the double/float base terms of the two consolidate kernels.  The
// mapping is MI355X's own:
//   - pull forms: a workgroup stages its rows' in-edge contributions in LDS
//     in rounds of 2048, gathered with every load in flight; rows with more
//     than 256 in-edges are summed one wave each, longest first, in the
//     same launch (the pull plan).  vertex-pull and vertex-pull-nodiv keep
//     the reference's sequential FP32 sum in CSR order, bit for bit; the
//     _warp forms combine lane-strided partials by shuffles where the
//     reference adds them atomically.  The dividing forms gather
//     per-vertex quotients computed once per iteration (the same IEEE
//     values);
//   - reverse edge lists: a segmented reduction over destination runs, one
//     store per run, atomics only for runs cut by a workgroup's range;
//   - forward edge lists and push forms: propagation blocking
//     (pagerank_scatter.hip) in bw_pr_run and the plugin; the per-call
//     entry points keep one atomic per edge;
//   - FP64 reductions of the residual and dangling mass, one atomic per
//     workgroup on a grid of two workgroups per CU.

#include "common.hpp"

#include <algorithm>
#include <memory>
#include <vector>

namespace bw {
namespace {

constexpr float kDamp = 0.85f;  // pagerank/pagerank.hpp:9

__device__ __forceinline__ void add_f32(float *p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void add_f64(double *p, double v) { unsafeAtomicAdd(p, v); }

// ---------------------------------------------------------------- degrees

// Occurrence counts of values[i * STRIDE + off] (the degree kernels'
// atomicAdd(&deg[x], 1) per edge).  Each wave reads 64 consecutive entries
// and adds every run of equal values inside them with one atomic from the
// run's last lane: a source-sorted edge list, whose lanes would otherwise
// all add to one counter, takes an atomic per run instead of per edge.
// Values below kHotIds are counted in LDS and added to deg once per
// workgroup: a graph relabelled by degree has its hubs there, whose
// counters would otherwise take one global atomic per edge, one after the
// other.
constexpr uint32_t kHotIds = 16384;

template<int STRIDE>
__global__ __launch_bounds__(256) void pr_degrees_runs_kernel(uint64_t E, const uint32_t *__restrict__ values,
                                                              int off, uint32_t *deg)
{
    __shared__ uint32_t hot[kHotIds];
    for (uint32_t i = threadIdx.x; i < kHotIds; i += blockDim.x) hot[i] = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
    for (uint64_t w = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave; w * kWave < E;
         w += waves) {
        const uint64_t i = w * kWave + lane;
        const bool valid = i < E;
        const uint32_t x = valid ? values[i * STRIDE + off] : 0xFFFFFFFFu;
        const uint32_t prev = __shfl_up(x, 1, kWave), next = __shfl_down(x, 1, kWave);
        const bool head = lane == 0 || prev != x;
        const bool tail = lane == kWave - 1 || next != x;
        const uint64_t heads = __ballot(head);
        const uint64_t upto = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull);
        const uint32_t h = 63u - static_cast<uint32_t>(__builtin_clzll(heads & upto));
        if (valid && tail) {
            if (x < kHotIds) atomicAdd(&hot[x], lane - h + 1u);
            else atomicAdd(&deg[x], lane - h + 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHotIds; i += blockDim.x)
        if (hot[i]) atomicAdd(&deg[i], hot[i]);
}

// Workgroups of the degree count: two per CU (the LDS counters take 64 KB),
// fewer for a short list, so that each flushes its counters after a long
// stretch of entries.
inline uint32_t degree_grid(uint64_t E)
{
    const uint64_t need = (E + 255) / 256;
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(need, 2ull * device_cus())));
}

template<typename Off>
__global__ __launch_bounds__(1024) void pr_degrees_csr_kernel(uint64_t V, const Off *__restrict__ off, uint32_t *deg)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride)
        atomicAdd(&deg[v], static_cast<uint32_t>(off[v + 1] - off[v]));
}

// ------------------------------------------------------------- iteration

template<bool REV>
__device__ __forceinline__ void push_edge(uint32_t in, uint32_t out, const uint32_t *deg, const float *pr,
                                          float *np)
{
    // pagerank/edgelist.cu:15-21 (forward: origin = in) and
    // rev_edgelist.cu:15-21 (reverse list: origin = out).
    const uint32_t origin = REV ? out : in, dest = REV ? in : out;
    const uint32_t d = deg[origin];
    float r = 0.0f;
    if (d != 0) r = pr[origin] / static_cast<float>(d);
    add_f32(&np[dest], r);
}

template<bool REV>
__global__ __launch_bounds__(1024) void pr_edge_list_kernel(uint64_t E, const uint32_t *__restrict__ in,
                                                            const uint32_t *__restrict__ out,
                                                            const uint32_t *__restrict__ deg,
                                                            const float *__restrict__ pr, float *np, int vec)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t done = 0;
    if (vec) {
        const uint64_t E4 = E / 4;
        const uint4 *in4 = reinterpret_cast<const uint4 *>(in);
        const uint4 *out4 = reinterpret_cast<const uint4 *>(out);
        for (uint64_t i = gid; i < E4; i += stride) {
            const uint4 a = in4[i], b = out4[i];
            push_edge<REV>(a.x, b.x, deg, pr, np);
            push_edge<REV>(a.y, b.y, deg, pr, np);
            push_edge<REV>(a.z, b.z, deg, pr, np);
            push_edge<REV>(a.w, b.w, deg, pr, np);
        }
        done = E4 * 4;
    }
    for (uint64_t i = done + gid; i < E; i += stride) push_edge<REV>(in[i], out[i], deg, pr, np);
}

template<bool REV>
__global__ __launch_bounds__(1024) void pr_struct_edge_list_kernel(uint64_t E, const uint32_t *__restrict__ edges,
                                                                   const uint32_t *__restrict__ deg,
                                                                   const float *__restrict__ pr, float *np,
                                                                   int vec)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t done = 0;
    if (vec) {
        const uint64_t E2 = E / 2;
        const uint4 *e4 = reinterpret_cast<const uint4 *>(edges);
        for (uint64_t i = gid; i < E2; i += stride) {
            const uint4 p = e4[i];
            push_edge<REV>(p.x, p.y, deg, pr, np);
            push_edge<REV>(p.z, p.w, deg, pr, np);
        }
        done = E2 * 2;
    }
    for (uint64_t i = done + gid; i < E; i += stride) push_edge<REV>(edges[2 * i], edges[2 * i + 1], deg, pr, np);
}

// ------------------------------------------------- reverse edge lists

constexpr uint32_t kSegThreads = 256;
constexpr uint32_t kSegPerThread = 16;
constexpr uint32_t kSegBlockEdges = kSegThreads * kSegPerThread;
constexpr uint32_t kNoKey = 0xFFFFFFFFu;

// Destination and source of reverse edge e: the SoA lists hold the
// destination in in[] and the source in out[]; the AoS pairs (in, out).
template<bool AOS>
__device__ __forceinline__ void rev_edge(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b, uint64_t e,
                                         uint32_t &dst, uint32_t &src)
{
    if constexpr (AOS) {
        dst = a[2 * e];
        src = a[2 * e + 1];
    } else {
        dst = a[e];
        src = b[e];
    }
}

// revEdgeListPageRank / revStructEdgeListPageRank (pagerank/rev_edgelist.cu:
// 3-23, rev_struct_edgelist.cu:3-24) for wave64.  The reverse lists are
// grouped by destination, so the per-edge atomic adds of the reference
// become a segmented reduction: a workgroup takes 4096 consecutive edges, 16
// per thread; each thread sums its runs of equal destination, a block-wide
// segmented scan carries a run's partial sum across threads, and the thread
// holding a run's last edge writes the run's total once.  A run cut by the
// workgroup's edge range (its destination continues in the neighbouring
// workgroup) is added atomically; every other run is stored, since no other
// workgroup touches its destination.  The sum order within a destination
// differs from the reference's (whose atomics have no fixed order).
template<bool AOS, bool PRE>
__global__ __launch_bounds__(kSegThreads) void pr_rev_segmented_kernel(uint64_t E, const uint32_t *__restrict__ a,
                                                                       const uint32_t *__restrict__ b,
                                                                       const uint32_t *__restrict__ deg,
                                                                       const float *__restrict__ pr, float *np,
                                                                       int vec)
{
    __shared__ uint32_t firstKeys[kSegThreads], lastKeys[kSegThreads];
    __shared__ float sums[kSegThreads];
    __shared__ float waveTotal[kSegThreads / kWave];
    __shared__ int waveHead[kSegThreads / kWave];
    const uint32_t t = threadIdx.x, lane = t & (kWave - 1), wv = t / kWave;
    const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * kSegBlockEdges;
    const uint64_t bEnd = b0 + kSegBlockEdges < E ? b0 + kSegBlockEdges : E;
    const uint64_t base = b0 + static_cast<uint64_t>(t) * kSegPerThread;

    uint32_t key[kSegPerThread], src[kSegPerThread];
    if (vec && base + kSegPerThread <= E) {
        if constexpr (AOS) {
            const uint4 *p = reinterpret_cast<const uint4 *>(a + 2 * base);
#pragma unroll
            for (uint32_t q = 0; q < kSegPerThread / 2; q++) {
                const uint4 x = p[q];
                key[2 * q] = x.x;
                src[2 * q] = x.y;
                key[2 * q + 1] = x.z;
                src[2 * q + 1] = x.w;
            }
        } else {
            const uint4 *pk = reinterpret_cast<const uint4 *>(a + base);
            const uint4 *ps = reinterpret_cast<const uint4 *>(b + base);
#pragma unroll
            for (uint32_t q = 0; q < kSegPerThread / 4; q++) {
                const uint4 x = pk[q], y = ps[q];
                key[4 * q] = x.x, key[4 * q + 1] = x.y, key[4 * q + 2] = x.z, key[4 * q + 3] = x.w;
                src[4 * q] = y.x, src[4 * q + 1] = y.y, src[4 * q + 2] = y.z, src[4 * q + 3] = y.w;
            }
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kSegPerThread; k++) {
            key[k] = kNoKey;
            src[k] = 0;
            if (base + k < E) rev_edge<AOS>(a, b, base + k, key[k], src[k]);
        }
    }
    // Contributions (rev_edgelist.cu:15-21: origin = out[], zero for a
    // vertex without out-edges), then the running sum of each run inside the
    // thread.
    // The gathers are unconditional (a missing edge reads source 0 and is
    // masked afterwards), so all of them are in flight together.
    // PRE: pr already holds each source's contribution (pr_contrib_kernel).
    uint32_t dg[kSegPerThread];
    float pv[kSegPerThread], c[kSegPerThread];
    if constexpr (!PRE) {
#pragma unroll
        for (uint32_t k = 0; k < kSegPerThread; k++) dg[k] = deg[src[k]];
    }
#pragma unroll
    for (uint32_t k = 0; k < kSegPerThread; k++) pv[k] = pr[src[k]];
#pragma unroll
    for (uint32_t k = 0; k < kSegPerThread; k++) {
        if constexpr (PRE) c[k] = key[k] != kNoKey ? pv[k] : 0.0f;
        else c[k] = (key[k] != kNoKey && dg[k] != 0) ? pv[k] / static_cast<float>(dg[k]) : 0.0f;
    }
#pragma unroll
    for (uint32_t k = 1; k < kSegPerThread; k++)
        if (key[k] == key[k - 1]) c[k] += c[k - 1];
    const uint32_t firstKey = key[0], lastKey = key[kSegPerThread - 1];
    firstKeys[t] = firstKey;
    lastKeys[t] = lastKey;
    __syncthreads();

    // Segmented inclusive scan over the threads of (carry out, head): a
    // thread whose edges are one run continuing the previous thread's last
    // run adds the previous carry; any other thread starts a new segment.
    const bool chained = t > 0 && firstKey == lastKey && lastKeys[t - 1] == firstKey;
    float x = c[kSegPerThread - 1];
    int h = chained ? 0 : 1;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const float px = __shfl_up(x, off, kWave);
        const int ph = __shfl_up(h, off, kWave);
        if (static_cast<int>(lane) >= off) {
            if (!h) x += px;
            h |= ph;
        }
    }
    if (lane == kWave - 1) {
        waveTotal[wv] = x;
        waveHead[wv] = h;
    }
    __syncthreads();
    if (!h) {
        // No segment head from the wave's start to here: the segment comes in
        // from the previous waves.
        float inc = 0.0f;
        for (uint32_t w = 0; w < wv; w++) inc = waveTotal[w] + (waveHead[w] ? 0.0f : inc);
        x += inc;
    }
    sums[t] = x;
    __syncthreads();
    const float carryIn = (t > 0 && lastKeys[t - 1] == firstKey) ? sums[t - 1] : 0.0f;

    // The key after this thread's last edge, and the runs the workgroup's
    // edge range cuts.
    uint32_t nextKey = kNoKey;
    if (t + 1 < kSegThreads) nextKey = firstKeys[t + 1];
    uint32_t prevEdgeKey = kNoKey, nextEdgeKey = kNoKey, blockFirst = kNoKey, blockLast = kNoKey, unused = 0;
    if (b0 > 0) rev_edge<AOS>(a, b, b0 - 1, prevEdgeKey, unused);
    if (bEnd < E) rev_edge<AOS>(a, b, bEnd, nextEdgeKey, unused);
    rev_edge<AOS>(a, b, b0, blockFirst, unused);
    rev_edge<AOS>(a, b, bEnd - 1, blockLast, unused);
    const bool leftOpen = prevEdgeKey == blockFirst, rightOpen = nextEdgeKey == blockLast;
#pragma unroll
    for (uint32_t k = 0; k < kSegPerThread; k++) {
        if (key[k] == kNoKey) continue;
        const uint32_t after = k + 1 < kSegPerThread ? key[k + 1] : nextKey;
        if (after == key[k]) continue;  // not the last edge of its run here
        const float total = c[k] + (key[k] == firstKey ? carryIn : 0.0f);
        if ((leftOpen && key[k] == blockFirst) || (rightOpen && key[k] == blockLast)) add_f32(&np[key[k]], total);
        else np[key[k]] = total;
    }
}

// pagerank/vertex_push.cu:16-30
template<typename Off>
__global__ __launch_bounds__(1024) void pr_vertex_push_kernel(uint64_t V, const Off *__restrict__ off,
                                                              const uint32_t *__restrict__ cols,
                                                              const float *__restrict__ pr, float *np)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint64_t s = off[v], e = off[v + 1];
        const uint32_t d = static_cast<uint32_t>(e - s);
        if (d == 0) continue;
        const float r = pr[v] / static_cast<float>(d);
        for (uint64_t j = s; j < e; j++) add_f32(&np[cols[j]], r);
    }
}

// ------------------------------------------------------------ pull forms

// Rows per workgroup of the pull kernels (one thread per row), and the
// in-edges whose contributions are staged in LDS per round.
constexpr uint32_t kPullRows = 256;
constexpr uint32_t kPullChunk = 2048;
constexpr uint32_t kPullPerThread = kPullChunk / kPullRows;
constexpr uint32_t kSumBatch = 32;     // LDS reads in flight in a row's sequential sum
constexpr uint32_t kStrideBatch = 8;   // the same for a lane of a virtual warp

// The contribution one in-edge from u brings: pr[u] / deg[u] for the
// dividing variants (vertex_pull.cu:25, the warp form dividing by the
// degree read as int, vertex_pull_warp.cu:40), pr[u] itself for nodiv.
template<int DIV>
__device__ __forceinline__ float pull_contribution(const float *__restrict__ pr, const uint32_t *__restrict__ deg,
                                                  uint32_t u)
{
    if constexpr (DIV == 1) return pr[u] / static_cast<float>(deg[u]);
    else if constexpr (DIV == 2) return pr[u] / static_cast<float>(static_cast<int32_t>(deg[u]));
    else return pr[u];
}

// The pull iteration of pagerank/vertex_pull.cu:16-30, vertex_pull_nodiv.cu,
// vertex_pull_warp.cu:4-49 and vertex_pull_warp_nodiv.cu:4-47 for wave64:
// a workgroup owns kPullRows consecutive rows of the reverse CSR and streams
// their in-edges in rounds of kPullChunk.  All 256 threads gather a round's
// contributions (coalesced column reads, independent gathers) into LDS while
// the previous round is being summed, so a hub row no longer walks its
// in-list with one dependent load after another.
//
// W = 1 (vertex-pull, vertex-pull-nodiv): thread t sums row t alone, in CSR
// order, carrying its running sum from round to round -- the reference's
// sequential FP32 sum, bit for bit.  W > 1 (the _warp variants): the W lanes
// of a virtual warp take the row's in-edges in lane-strided order (lane l
// the edges at positions l, l + W, ... of the row, as the reference's lanes
// do) and their partials are combined by a shuffle tree where the reference
// adds them atomically, in no fixed order.  Rows go round-robin to the
// virtual warps; a round boundary cuts at most one row, whose partials the
// owning warp carries to the next round.
struct PullShortLds {
    float vals[2][kPullChunk];
    uint64_t rowOff[kPullRows + 1];
};

template<int DIV, int W, typename Off>
__device__ __forceinline__ void pull_short_rows(uint32_t block, PullShortLds &lds, uint64_t V,
                                                const Off *__restrict__ roff, const uint32_t *__restrict__ rcols,
                                                const uint32_t *__restrict__ deg, const float *__restrict__ pr,
                                                float *np, uint64_t long_degree)
{
    auto &vals = lds.vals;
    auto &rowOff = lds.rowOff;
    constexpr uint32_t kWarps = kPullRows / W;
    const uint32_t t = threadIdx.x;
    const uint64_t r0 = static_cast<uint64_t>(block) * kPullRows;
    const uint32_t nrows = static_cast<uint32_t>(V - r0 < kPullRows ? V - r0 : kPullRows);
    for (uint32_t i = t; i <= nrows; i += kPullRows) rowOff[i] = static_cast<uint64_t>(roff[r0 + i]);
    __syncthreads();
    const uint64_t e0 = rowOff[0], e1 = rowOff[nrows];
    const uint32_t warp = t / W, lane = t % W;
    auto is_long = [&](uint32_t i) { return rowOff[i + 1] - rowOff[i] > long_degree; };

    // Rows with more than long_degree in-edges are summed by
    // pr_pull_long_kernel; a round starting inside one jumps to its end.
    auto skip_long = [&](uint64_t c) {
        while (c < e1) {
            uint32_t lo = 0, hi = nrows - 1;  // the row holding edge c: last row starting at or before it
            while (lo < hi) {
                const uint32_t m = (lo + hi + 1) / 2;
                if (rowOff[m] <= c) lo = m;
                else hi = m - 1;
            }
            if (!is_long(lo)) break;
            c = rowOff[lo + 1];
        }
        return c;
    };

    // Loads of one round into registers; stored to LDS after the previous
    // round's sums, so those sums overlap the gathers' latency.  The loads
    // are unconditional (indices past the round repeat its last edge) so
    // that all column loads, then all gathers, are in flight together.
    float reg[kPullPerThread];
    auto gather = [&](uint64_t c0, uint64_t c1) {
        uint32_t u[kPullPerThread];
#pragma unroll
        for (uint32_t k = 0; k < kPullPerThread; k++) {
            const uint64_t e = c0 + k * kPullRows + t;
            u[k] = rcols[e < c1 ? e : c1 - 1];
        }
#pragma unroll
        for (uint32_t k = 0; k < kPullPerThread; k++) reg[k] = pull_contribution<DIV>(pr, deg, u[k]);
    };
    auto stage = [&](int buf) {
#pragma unroll
        for (uint32_t k = 0; k < kPullPerThread; k++) vals[buf][k * kPullRows + t] = reg[k];
    };

    // Rows without in-edges get an empty sum.
    for (uint32_t i = t; i < nrows; i += kPullRows)
        if (rowOff[i] == rowOff[i + 1]) np[r0 + i] = 0.0f;

    float carry = 0.0f;  // partial of this lane's row cut by the round boundary
    uint64_t c0 = skip_long(e0);
    uint64_t c1 = c0 + kPullChunk < e1 ? c0 + kPullChunk : e1;
    if (c0 < e1) {
        gather(c0, c1);
        stage(0);
    }
    __syncthreads();
    int buf = 0;
    while (c0 < e1) {
        const uint64_t n0 = skip_long(c1);
        const uint64_t n1 = n0 + kPullChunk < e1 ? n0 + kPullChunk : e1;
        if (n0 < e1) gather(n0, n1);
        const float *v = vals[buf];
        if constexpr (W == 1) {
            // Thread t owns row t.
            if (t < nrows && !is_long(t)) {
                const uint64_t s = rowOff[t], e = rowOff[t + 1];
                const uint64_t lo = s > c0 ? s : c0, hi = e < c1 ? e : c1;
                if (lo < hi) {
                    // In CSR order; the LDS reads of a batch are issued
                    // together, the adds stay sequential.
                    float acc = carry;
                    uint32_t j = static_cast<uint32_t>(lo - c0);
                    const uint32_t end = static_cast<uint32_t>(hi - c0);
                    for (; j + kSumBatch <= end; j += kSumBatch) {
                        float x[kSumBatch];
#pragma unroll
                        for (uint32_t k = 0; k < kSumBatch; k++) x[k] = v[j + k];
#pragma unroll
                        for (uint32_t k = 0; k < kSumBatch; k++) acc += x[k];
                    }
                    for (; j < end; j++) acc += v[j];
                    if (hi == e) np[r0 + t] = acc;
                    carry = hi == e ? 0.0f : acc;
                }
            }
        } else {
            // Rows overlapping the round: the first row whose end passes c0
            // up to the last row starting before c1 (binary searches in LDS).
            uint32_t lo = 0, hi = nrows;
            while (lo < hi) {
                const uint32_t m = (lo + hi) / 2;
                if (rowOff[m + 1] <= c0) lo = m + 1;
                else hi = m;
            }
            const uint32_t ra = lo;
            hi = nrows;
            while (lo < hi) {
                const uint32_t m = (lo + hi) / 2;
                if (rowOff[m] < c1) lo = m + 1;
                else hi = m;
            }
            const uint32_t rb = lo;  // rows [ra, rb) overlap the round
            const uint32_t first = ra + (warp + kWarps - ra % kWarps) % kWarps;
            for (uint32_t i = first; i < rb; i += kWarps) {
                const uint64_t s = rowOff[i], e = rowOff[i + 1];
                if (s == e || is_long(i)) continue;
                const uint64_t a = s > c0 ? s : c0, b = e < c1 ? e : c1;
                // Lane l takes the row positions congruent to l modulo W.
                const uint64_t p0 = a - s;
                uint32_t j = static_cast<uint32_t>(a - c0) + static_cast<uint32_t>((lane + W - p0 % W) % W);
                const uint32_t end = static_cast<uint32_t>(b - c0);
                float part = (a == s) ? 0.0f : carry;
                for (; j + (kStrideBatch - 1) * W < end; j += kStrideBatch * W) {
                    float x[kStrideBatch];
#pragma unroll
                    for (uint32_t k = 0; k < kStrideBatch; k++) x[k] = v[j + k * W];
#pragma unroll
                    for (uint32_t k = 0; k < kStrideBatch; k++) part += x[k];
                }
                for (; j < end; j += W) part += v[j];
                if (b == e) {
#pragma unroll
                    for (int o = W / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, W);
                    if (lane == 0) np[r0 + i] = part;
                    carry = 0.0f;
                } else {
                    carry = part;
                }
            }
        }
        if (n0 < e1) stage(buf ^ 1);
        __syncthreads();
        c0 = n0;
        c1 = n1;
        buf ^= 1;
    }
}

// Rows with more than long_degree in-edges, one wave per row (listed by
// the pull plan), so that the hubs of a skewed graph are summed side by
// side rather than one after another by the workgroup owning their rows.
// The wave gathers kLongRound contributions per round (8 per lane, all in
// flight together).  SEQ: lane 0 adds them in CSR order through LDS, the
// reference's sequential sum bit for bit.  Otherwise (the _warp variants)
// lane l keeps the partial of the positions congruent to l modulo 64 and
// the partials meet in a shuffle tree -- the reference's per-lane partials
// and atomic adds in another, equally valid order.
constexpr uint32_t kLongWaves = 4;
constexpr uint32_t kLongPerLane = 8;
constexpr uint32_t kLongRound = kWave * kLongPerLane;

struct PullLongLds {
    float vals[kLongWaves][2][kLongRound];
};

template<int DIV, bool SEQ, typename Off>
__device__ __forceinline__ void pull_long_rows(uint32_t block, PullLongLds &lds, const uint32_t *__restrict__ rows,
                                               uint32_t nrows, const Off *__restrict__ roff,
                                               const uint32_t *__restrict__ rcols, const uint32_t *__restrict__ deg,
                                               const float *__restrict__ pr, float *np)
{
    auto &vals = lds.vals;
    const uint32_t wv = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const uint32_t i = block * kLongWaves + wv;
    if (i >= nrows) return;
    const uint32_t row = rows[i];
    const uint64_t s = static_cast<uint64_t>(roff[row]), e = static_cast<uint64_t>(roff[row + 1]);
    float x[kLongPerLane];
    auto gather = [&](uint64_t c0) {
        uint32_t u[kLongPerLane];
#pragma unroll
        for (uint32_t k = 0; k < kLongPerLane; k++) {
            const uint64_t j = c0 + k * kWave + lane;
            u[k] = rcols[j < e ? j : e - 1];
        }
#pragma unroll
        for (uint32_t k = 0; k < kLongPerLane; k++) {
            const uint64_t j = c0 + k * kWave + lane;
            x[k] = j < e ? pull_contribution<DIV>(pr, deg, u[k]) : 0.0f;
        }
    };
    if constexpr (!SEQ) {
        float part = 0.0f;
        for (uint64_t c0 = s; c0 < e; c0 += kLongRound) {
            gather(c0);
#pragma unroll
            for (uint32_t k = 0; k < kLongPerLane; k++) part += x[k];
        }
        part = wave_sum_f32(part);
        if (lane == 0) np[row] = part;
    } else {
        float acc = 0.0f;
        int buf = 0;
        gather(s);
#pragma unroll
        for (uint32_t k = 0; k < kLongPerLane; k++) vals[wv][0][k * kWave + lane] = x[k];
        for (uint64_t c0 = s; c0 < e; c0 += kLongRound) {
            const uint64_t c1 = c0 + kLongRound < e ? c0 + kLongRound : e;
            if (c1 < e) gather(c1);
            // The staged round is visible to lane 0 once the wave's LDS
            // writes are ordered before its reads.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            if (lane == 0) {
                // In CSR order, four values per 16-byte LDS read.
                const float *v = vals[wv][buf];
                const float4 *v4 = reinterpret_cast<const float4 *>(v);
                const uint32_t n = static_cast<uint32_t>(c1 - c0);
                uint32_t q = 0;
                for (; 4 * (q + kSumBatch / 4) <= n; q += kSumBatch / 4) {
                    float4 y[kSumBatch / 4];
#pragma unroll
                    for (uint32_t k = 0; k < kSumBatch / 4; k++) y[k] = v4[q + k];
#pragma unroll
                    for (uint32_t k = 0; k < kSumBatch / 4; k++) {
                        acc += y[k].x;
                        acc += y[k].y;
                        acc += y[k].z;
                        acc += y[k].w;
                    }
                }
                for (uint32_t j = 4 * q; j < n; j++) acc += v[j];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            if (c1 < e) {
#pragma unroll
                for (uint32_t k = 0; k < kLongPerLane; k++) vals[wv][buf ^ 1][k * kWave + lane] = x[k];
            }
            buf ^= 1;
        }
        if (lane == 0) np[row] = acc;
    }
}

// c[v] = pr[v] / deg[v] (zero without out-edges): the value every
// out-edge of v carries in the dividing variants.  Computing it once per
// vertex instead of once per edge gives the same IEEE quotient, so the sums
// that gather it are the reference's bit for bit, with one gather per edge
// instead of two.  INT: the degree read as int (vertex_pull_warp.cu:40).
template<bool INT>
__global__ __launch_bounds__(256) void pr_contrib_kernel(uint64_t V, const uint32_t *__restrict__ deg,
                                                         const float *__restrict__ pr, float *c)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint32_t d = deg[v];
        const float q = INT ? pr[v] / static_cast<float>(static_cast<int32_t>(d)) : pr[v] / static_cast<float>(d);
        c[v] = d != 0 ? q : 0.0f;
    }
}

// One launch of a pull iteration: workgroups [0, long_blocks) sum the long
// rows (longest first, so the biggest hub starts at once), the others one
// block of kPullRows rows each.
template<int DIV, int W, typename Off>
__global__ __launch_bounds__(kPullRows) void pr_pull_kernel(uint64_t V, const Off *__restrict__ roff,
                                                             const uint32_t *__restrict__ rcols,
                                                             const uint32_t *__restrict__ deg,
                                                             const float *__restrict__ pr, float *np,
                                                             uint64_t long_degree, const uint32_t *__restrict__ rows,
                                                             uint32_t nlong, uint32_t long_blocks)
{
    static_assert(kLongWaves * kWave == kPullRows, "both parts use the same workgroup");
    __shared__ union alignas(16) {
        PullShortLds s;
        PullLongLds l;
    } lds;
    if (blockIdx.x < long_blocks)
        pull_long_rows<DIV, W == 1, Off>(blockIdx.x, lds.l, rows, nlong, roff, rcols, deg, pr, np);
    else
        pull_short_rows<DIV, W, Off>(blockIdx.x - long_blocks, lds.s, V, roff, rcols, deg, pr, np, long_degree);
}

// The long rows of a reverse CSR: rows whose in-degree exceeds `threshold`
// (appended in any order).
template<typename Off>
__global__ void pr_long_rows_kernel(uint64_t V, const Off *__restrict__ roff, uint64_t threshold, uint2 *rows,
                                    uint32_t *count)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint64_t d = static_cast<uint64_t>(roff[v + 1]) - static_cast<uint64_t>(roff[v]);
        if (d > threshold)
            rows[atomicAdd(count, 1u)] = make_uint2(static_cast<uint32_t>(v), d > 0xFFFFFFFFull ? 0xFFFFFFFFu
                                                                                             : static_cast<uint32_t>(d));
    }
}

template<typename Off>
__host__ __device__ inline uint32_t pr_warp_vals_bytes(uint32_t warps, uint32_t chunk)
{ return (warps * chunk * 4u + 7u) & ~7u; }

template<typename Off>
__host__ __device__ inline uint32_t pr_warp_lds_bytes(uint32_t warps, uint32_t chunk)
{ return pr_warp_vals_bytes<Off>(warps, chunk) + warps * (chunk + 1) * static_cast<uint32_t>(sizeof(Off)); }

// pagerank/vertex_push_warp.cu:4-49: chunk ranks and offsets staged in LDS,
// lanes stride over each vertex's out-edges with float atomics.
template<typename Off>
__global__ __launch_bounds__(1024) void pr_vertex_push_warp_kernel(uint32_t W, uint32_t C, uint64_t V,
                                                                   const Off *__restrict__ off,
                                                                   const uint32_t *__restrict__ cols,
                                                                   const float *__restrict__ pr, float *np)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t warps = blockDim.x / W;
    const uint32_t bw = threadIdx.x / W, lane = threadIdx.x % W;
    float *myRanks = reinterpret_cast<float *>(lds_raw) + bw * C;
    Off *myOff = reinterpret_cast<Off *>(lds_raw + pr_warp_vals_bytes<Off>(warps, C)) + bw * (C + 1);
    const uint64_t warpId = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / W;
    const uint64_t totalWarps = static_cast<uint64_t>(warps) * gridDim.x;
    for (uint64_t chunk = warpId; chunk * C < V; chunk += totalWarps) {
        const uint64_t base = chunk * C;
        const uint32_t end = static_cast<uint32_t>(std::min<uint64_t>(C, V - base));
        for (uint32_t i = lane; i < end; i += W) myRanks[i] = pr[base + i];
        for (uint32_t i = lane; i <= end; i += W) myOff[i] = off[base + i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint32_t v = 0; v < end; v++) {
            const uint64_t s = myOff[v];
            const uint32_t n = static_cast<uint32_t>(myOff[v + 1] - s);
            if (n == 0) continue;
            const float r = myRanks[v] / static_cast<float>(n);
            for (uint32_t i = lane; i < n; i += W) add_f32(&np[cols[s + i]], r);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------ consolidate

__global__ __launch_bounds__(1024) void pr_consolidate_kernel(uint64_t V, const uint32_t *__restrict__ deg,
                                                              float *pr, float *np, int notLast, int nodiv,
                                                              int mode, const double *__restrict__ dangling,
                                                              double *resid)
{
    __shared__ double lds[1024 / kWave];
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    const double d = static_cast<double>(kDamp);
    const double D = (mode == BW_PR_DANGLING_UNIFORM && dangling) ? *dangling : 0.0;
    // pagerank/pagerank.cu:154: (1 - dampening) / size in float.
    const float baseF = (1.0f - kDamp) / static_cast<float>(V);
    // pagerank/pagerank.cu:132: (1.0 - dampening) / size in double.
    const double baseD = (1.0 - d) / static_cast<double>(V);
    double diff = 0.0;
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
        const float acc = np[v];
        float nr;
        if (mode == BW_PR_DANGLING_DROP) {
            if (nodiv) nr = __builtin_fmaf(kDamp, acc, baseF);
            else nr = static_cast<float>(baseD + static_cast<double>(kDamp * acc));
        } else {
            nr = static_cast<float>(baseD + d * (static_cast<double>(acc) + D / static_cast<double>(V)));
        }
        const float old = pr[v];
        const uint32_t dv = (nodiv || mode == BW_PR_DANGLING_UNIFORM) && deg ? deg[v] : 0u;
        double oldRank = static_cast<double>(old);
        if (nodiv && mode == BW_PR_DANGLING_UNIFORM && dv != 0) oldRank *= static_cast<double>(dv);
        diff += fabs(static_cast<double>(nr) - oldRank);
        pr[v] = (nodiv && dv != 0 && notLast) ? nr / static_cast<float>(dv) : nr;
        np[v] = 0.0f;
    }
    if (resid) {
        const double t = block_sum_f64(diff, lds);
        if (threadIdx.x == 0) add_f64(resid, t);
    }
}

__global__ __launch_bounds__(1024) void pr_dangling_kernel(uint64_t V, const uint32_t *__restrict__ deg,
                                                           const float *__restrict__ pr, double *out)
{
    __shared__ double lds[1024 / kWave];
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    double acc = 0.0;
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride)
        if (deg[v] == 0) acc += static_cast<double>(pr[v]);
    const double t = block_sum_f64(acc, lds);
    if (threadIdx.x == 0 && t != 0.0) add_f64(out, t);
}

__global__ void pr_init_kernel(uint64_t V, float *pr, float *np, const uint32_t *deg, int predivide)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    // pagerank/pagerank.cpp:88: 1.0f / vertex_count.
    const float init = 1.0f / static_cast<float>(V);
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
        float x = init;
        if (predivide && deg[v] != 0) x = init / static_cast<float>(deg[v]);
        pr[v] = x;
        np[v] = 0.0f;
    }
}

__global__ void zero_f64_kernel(double *p, uint64_t n)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------------ host glue

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// Workgroups of the kernels that end in one FP64 atomic add per workgroup
// (residual, dangling mass): two per CU unless the caller fixed the grid.
// Those adds all hit one address and serialise; 16384 of them cost 200 us
// on MI355X, the grid-stride loop over more vertices per thread almost
// nothing.
uint32_t reduction_grid(const bw_launch *cfg, uint32_t grid)
{
    if (cfg && cfg->grid) return grid;
    return std::min<uint32_t>(grid, 2u * static_cast<uint32_t>(device_cus()));
}

template<typename F>
int with_offsets(const void *off, int32_t ob, F &&f)
{
    if (ob == 4) return f(static_cast<const uint32_t *>(off));
    if (ob == 8) return f(static_cast<const uint64_t *>(off));
    return fail(BW_HIP_ERR_ARG, "offset_bytes must be 4 or 8");
}

int warp_params(const bw_launch *cfg, uint32_t &W, uint32_t &C)
{
    W = (cfg && cfg->warp) ? cfg->warp : 32;
    C = (cfg && cfg->chunk) ? cfg->chunk : 32;
    if (W == 0 || W > 64 || (W & (W - 1)))
        return fail(BW_HIP_ERR_ARG, "Warp size should be a power of 2 no larger than 64!");
    if (C % W != 0) return fail(BW_HIP_ERR_ARG, "Chunk size should be a multiple of warp size!");
    return BW_HIP_OK;
}

}  // namespace
}  // namespace bw

using namespace bw;

extern "C" int bw_pr_zero_degrees(uint64_t V, uint32_t *deg, const bw_launch *cfg)
{
    if (V && !deg) return fail(BW_HIP_ERR_ARG, "degrees must be a device pointer");
    return bw_hip_fill_u32(deg, 0u, V, cfg ? cfg->stream : nullptr);
}

extern "C" int bw_pr_degrees_edge_list(uint64_t E, const uint32_t *origins, uint32_t *deg, const bw_launch *cfg)
{
    if (E == 0) return BW_HIP_OK;
    if (!origins || !deg) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    const uint32_t block = 256, grid = degree_grid(E);
    hipLaunchKernelGGL(pr_degrees_runs_kernel<1>, dim3(grid), dim3(block), 0, stream_of(cfg), E, origins, 0, deg);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_pr_degrees_struct_edge_list(uint64_t E, const uint32_t *edges, int32_t reverse, uint32_t *deg,
                                              const bw_launch *cfg)
{
    if (E == 0) return BW_HIP_OK;
    if (!edges || !deg) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    const uint32_t block = 256, grid = degree_grid(E);
    hipLaunchKernelGGL(pr_degrees_runs_kernel<2>, dim3(grid), dim3(block), 0, stream_of(cfg), E, edges,
                       reverse ? 1 : 0, deg);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_pr_degrees_csr(uint64_t V, const void *off, int32_t ob, uint32_t *deg, const bw_launch *cfg)
{
    if (V == 0) return BW_HIP_OK;
    if (!off || !deg) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, V, block, grid));
    return with_offsets(off, ob, [&](auto *o) {
        using Off = std::remove_const_t<std::remove_pointer_t<decltype(o)>>;
        hipLaunchKernelGGL((pr_degrees_csr_kernel<Off>), dim3(grid), dim3(block), 0, stream_of(cfg), V, o, deg);
        BW_LAUNCH_CHECK();
        return BW_HIP_OK;
    });
}

extern "C" int bw_pr_degrees_rev_csr(uint64_t V, const void *roff, int32_t ob, const uint32_t *rcols,
                                     uint32_t *deg, const bw_launch *cfg)
{
    // reverseCSRComputeDegrees counts each source once per reverse-CSR
    // entry; the entries are rcols[0 .. roff[V]), so the count runs over the
    // column array directly, edge-parallel.
    if (V == 0) return BW_HIP_OK;
    if (!roff || !rcols || !deg) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    uint64_t E = 0;
    if (ob == 4) {
        uint32_t e32 = 0;
        BW_HIP_CHECK(hipMemcpy(&e32, static_cast<const uint32_t *>(roff) + V, 4, hipMemcpyDeviceToHost));
        E = e32;
    } else if (ob == 8) {
        BW_HIP_CHECK(hipMemcpy(&E, static_cast<const uint64_t *>(roff) + V, 8, hipMemcpyDeviceToHost));
    } else {
        return fail(BW_HIP_ERR_ARG, "offset_bytes must be 4 or 8");
    }
    return bw_pr_degrees_edge_list(E, rcols, deg, cfg);
}

extern "C" int bw_pr_edge_list(uint64_t E, const uint32_t *in, const uint32_t *out, const uint32_t *deg,
                               const float *pr, float *np, const bw_launch *cfg)
{
    if (E == 0) return BW_HIP_OK;
    if (!in || !out || !deg || !pr || !np) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    const int vec = aligned16(in) && aligned16(out);
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, vec ? (E + 3) / 4 : E, block, grid));
    hipLaunchKernelGGL(pr_edge_list_kernel<false>, dim3(grid), dim3(block), 0, stream_of(cfg), E, in, out, deg, pr,
                       np, vec);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

// Launches the segmented reverse edge-list kernel: one 256-thread workgroup
// per 4096 edges (its fixed geometry; cfg's block and grid do not apply).
template<bool AOS, bool PRE = false>
static int launch_rev_segmented(uint64_t E, const uint32_t *a, const uint32_t *b, const uint32_t *deg,
                                const float *pr, float *np, hipStream_t s)
{
    const uint64_t blocks = (E + kSegBlockEdges - 1) / kSegBlockEdges;
    if (blocks > 0x7FFFFFFFull) return fail(BW_HIP_ERR_ARG, "too many edges for the reverse edge-list kernel");
    const int vec = aligned16(a) && (AOS || aligned16(b));
    hipLaunchKernelGGL((pr_rev_segmented_kernel<AOS, PRE>), dim3(static_cast<uint32_t>(blocks)), dim3(kSegThreads),
                       0, s, E, a, b, deg, pr, np, vec);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_pr_rev_edge_list(uint64_t E, const uint32_t *in, const uint32_t *out, const uint32_t *deg,
                                   const float *pr, float *np, const bw_launch *cfg)
{
    if (E == 0) return BW_HIP_OK;
    if (!in || !out || !deg || !pr || !np) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    return launch_rev_segmented<false>(E, in, out, deg, pr, np, stream_of(cfg));
}

static int struct_entry(bool rev, uint64_t E, const uint32_t *edges, const uint32_t *deg, const float *pr,
                        float *np, const bw_launch *cfg)
{
    if (E == 0) return BW_HIP_OK;
    if (!edges || !deg || !pr || !np) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    if (rev) return launch_rev_segmented<true>(E, edges, nullptr, deg, pr, np, stream_of(cfg));
    const int vec = aligned16(edges);
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, vec ? (E + 1) / 2 : E, block, grid));
    hipLaunchKernelGGL(pr_struct_edge_list_kernel<false>, dim3(grid), dim3(block), 0, stream_of(cfg), E, edges, deg,
                       pr, np, vec);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_pr_struct_edge_list(uint64_t E, const uint32_t *edges, const uint32_t *deg, const float *pr,
                                      float *np, const bw_launch *cfg)
{ return struct_entry(false, E, edges, deg, pr, np, cfg); }

extern "C" int bw_pr_rev_struct_edge_list(uint64_t E, const uint32_t *edges, const uint32_t *deg, const float *pr,
                                          float *np, const bw_launch *cfg)
{ return struct_entry(true, E, edges, deg, pr, np, cfg); }

extern "C" int bw_pr_vertex_push(uint64_t V, const void *off, int32_t ob, const uint32_t *cols, const float *pr,
                                 float *np, const bw_launch *cfg)
{
    if (V == 0) return BW_HIP_OK;
    if (!off || !cols || !pr || !np) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, V, block, grid));
    return with_offsets(off, ob, [&](auto *o) {
        using Off = std::remove_const_t<std::remove_pointer_t<decltype(o)>>;
        hipLaunchKernelGGL((pr_vertex_push_kernel<Off>), dim3(grid), dim3(block), 0, stream_of(cfg), V, o, cols, pr,
                           np);
        BW_LAUNCH_CHECK();
        return BW_HIP_OK;
    });
}

// Launches the chunked pull kernel: one 256-thread workgroup per 256 rows
// (the kernel's fixed geometry; cfg's block and grid do not apply), W lanes
// per row.
template<int DIV, int W>
static int launch_pull(uint64_t V, const void *roff, int32_t ob, const uint32_t *rcols, const uint32_t *deg,
                       const float *pr, float *np, hipStream_t s, uint64_t long_degree = ~0ull,
                       const uint32_t *long_rows = nullptr, uint32_t nlong = 0)
{
    const uint64_t blocks = (V + kPullRows - 1) / kPullRows;
    if (blocks > 0x7FFFFFFFull) return fail(BW_HIP_ERR_ARG, "too many vertices for the pull kernel");
    return with_offsets(roff, ob, [&](auto *o) {
        using Off = std::remove_const_t<std::remove_pointer_t<decltype(o)>>;
        const uint32_t long_blocks = (nlong + kLongWaves - 1) / kLongWaves;
        if (blocks + long_blocks > 0x7FFFFFFFull) return fail(BW_HIP_ERR_ARG, "too many vertices for the pull kernel");
        hipLaunchKernelGGL((pr_pull_kernel<DIV, W, Off>), dim3(static_cast<uint32_t>(blocks + long_blocks)),
                           dim3(kPullRows), 0, s, V, o, rcols, deg, pr, np, nlong ? long_degree : ~0ull, long_rows,
                           nlong, long_blocks);
        BW_LAUNCH_CHECK();
        return BW_HIP_OK;
    });
}

template<bool DIV>
static int pull_entry(uint64_t V, const void *roff, int32_t ob, const uint32_t *rcols, const uint32_t *deg,
                      const float *pr, float *np, const bw_launch *cfg)
{
    if (V == 0) return BW_HIP_OK;
    if (!roff || !rcols || !pr || !np || (DIV && !deg)) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    return launch_pull<DIV ? 1 : 0, 1>(V, roff, ob, rcols, deg, pr, np, stream_of(cfg));
}

// Rows with more in-edges than this are summed one wave each
// (pull_long_rows): beyond it, one thread's sequential sum of a row would
// hold up the workgroup owning it.
constexpr uint64_t kLongDegree = 256;

struct bw_pr_pull_plan {
    uint64_t V = 0;
    const void *roff = nullptr;
    int32_t ob = 4;
    uint32_t *rows = nullptr;
    uint32_t nlong = 0;
    float *contrib = nullptr;  // per-vertex contributions of the dividing variants
    ~bw_pr_pull_plan()
    {
        if (rows) (void)hipFree(rows);
        if (contrib) (void)hipFree(contrib);
    }
};

extern "C" int bw_pr_pull_plan_create(bw_pr_pull_plan **out, uint64_t V, const void *roff, int32_t ob, void *stream)
{
    if (!out) return fail(BW_HIP_ERR_ARG, "bw_pr_pull_plan_create: out is NULL");
    *out = nullptr;
    if (V && !roff) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    if (ob != 4 && ob != 8) return fail(BW_HIP_ERR_ARG, "offset_bytes must be 4 or 8");
    if (V > 0xFFFFFFFFull) return fail(BW_HIP_ERR_ARG, "the pull plan holds 32-bit row ids");
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto plan = std::make_unique<bw_pr_pull_plan>();
    plan->V = V, plan->roff = roff, plan->ob = ob;
    if (V) {
        uint64_t E = 0;
        if (ob == 4) {
            uint32_t e32 = 0;
            BW_HIP_CHECK(hipMemcpyAsync(&e32, static_cast<const uint32_t *>(roff) + V, 4, hipMemcpyDeviceToHost, s));
            BW_HIP_CHECK(hipStreamSynchronize(s));
            E = e32;
        } else {
            BW_HIP_CHECK(hipMemcpyAsync(&E, static_cast<const uint64_t *>(roff) + V, 8, hipMemcpyDeviceToHost, s));
            BW_HIP_CHECK(hipStreamSynchronize(s));
        }
        // At most E / (kLongDegree + 1) rows can exceed the threshold; the
        // count lives in the slot after them.
        const uint64_t cap = E / (kLongDegree + 1);
        if (cap) {
            // (row, degree) pairs appended on the device, then ordered by
            // decreasing degree on the host (once per plan), so the longest
            // rows are dispatched first.
            uint2 *pairs = nullptr;
            BW_HIP_CHECK(hipMalloc(&pairs, sizeof(uint2) * (cap + 1)));
            struct Free {
                void *p;
                ~Free() { (void)hipFree(p); }
            } free_pairs{pairs};
            uint32_t *count = reinterpret_cast<uint32_t *>(pairs + cap);
            BW_TRY(bw_hip_fill_u32(count, 0u, 1, s));
            uint32_t block = 256, grid = 0;
            BW_TRY(resolve_geometry(nullptr, V, block, grid));
            BW_TRY(with_offsets(roff, ob, [&](auto *o) {
                hipLaunchKernelGGL(pr_long_rows_kernel, dim3(grid), dim3(block), 0, s, V, o, kLongDegree, pairs,
                                   count);
                BW_LAUNCH_CHECK();
                return BW_HIP_OK;
            }));
            BW_HIP_CHECK(hipMemcpyAsync(&plan->nlong, count, 4, hipMemcpyDeviceToHost, s));
            BW_HIP_CHECK(hipStreamSynchronize(s));
            if (plan->nlong) {
                std::vector<uint2> h(plan->nlong);
                BW_HIP_CHECK(hipMemcpy(h.data(), pairs, sizeof(uint2) * h.size(), hipMemcpyDeviceToHost));
                std::sort(h.begin(), h.end(), [](const uint2 &a, const uint2 &b) {
                    return a.y > b.y || (a.y == b.y && a.x < b.x);
                });
                std::vector<uint32_t> rows(h.size());
                for (size_t k = 0; k < h.size(); k++) rows[k] = h[k].x;
                BW_HIP_CHECK(hipMalloc(&plan->rows, sizeof(uint32_t) * rows.size()));
                BW_HIP_CHECK(hipMemcpy(plan->rows, rows.data(), sizeof(uint32_t) * rows.size(), hipMemcpyHostToDevice));
            }
        }
    }
    if (V) BW_HIP_CHECK(hipMalloc(&plan->contrib, sizeof(float) * V));
    *out = plan.release();
    return BW_HIP_OK;
}

extern "C" int bw_pr_pull_plan_destroy(bw_pr_pull_plan *plan)
{
    delete plan;
    return BW_HIP_OK;
}

extern "C" int bw_pr_pull_plan_info(const bw_pr_pull_plan *plan, uint32_t *long_rows)
{
    if (!plan) return fail(BW_HIP_ERR_ARG, "plan is NULL");
    if (long_rows) *long_rows = plan->nlong;
    return BW_HIP_OK;
}

extern "C" int bw_pr_pull_iteration(const bw_pr_pull_plan *plan, int32_t variant, const uint32_t *rcols,
                                    const uint32_t *deg, const float *pr, float *np, const bw_launch *cfg)
{
    if (!plan) return fail(BW_HIP_ERR_ARG, "plan is NULL");
    const uint64_t V = plan->V;
    if (V == 0) return BW_HIP_OK;
    const bool div = variant == BW_PR_VERTEX_PULL || variant == BW_PR_VERTEX_PULL_WARP;
    if (!rcols || !pr || !np || (div && !deg)) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    hipStream_t s = stream_of(cfg);
    const void *ro = plan->roff;
    const int32_t ob = plan->ob;
    const uint32_t *lr = plan->rows;
    const uint32_t nl = plan->nlong;
    if (variant != BW_PR_VERTEX_PULL && variant != BW_PR_VERTEX_PULL_NODIV && variant != BW_PR_VERTEX_PULL_WARP &&
        variant != BW_PR_VERTEX_PULL_WARP_NODIV)
        return fail(BW_HIP_ERR_ARG, "bw_pr_pull_iteration takes the four pull variants");
    uint32_t W = 1, C = 1;
    if (variant == BW_PR_VERTEX_PULL_WARP || variant == BW_PR_VERTEX_PULL_WARP_NODIV) BW_TRY(warp_params(cfg, W, C));
    if (div) {
        // The dividing variants gather per-vertex contributions computed here.
        const uint64_t blocks = std::min<uint64_t>((V + 255) / 256, 8ull * device_cus());
        if (variant == BW_PR_VERTEX_PULL)
            hipLaunchKernelGGL(pr_contrib_kernel<false>, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s, V, deg,
                               pr, plan->contrib);
        else
            hipLaunchKernelGGL(pr_contrib_kernel<true>, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s, V, deg,
                               pr, plan->contrib);
        BW_LAUNCH_CHECK();
        pr = plan->contrib;
    }
    switch (W) {
    case 1: return launch_pull<0, 1>(V, ro, ob, rcols, deg, pr, np, s, kLongDegree, lr, nl);
    case 2: return launch_pull<0, 2>(V, ro, ob, rcols, deg, pr, np, s, kLongDegree, lr, nl);
    case 4: return launch_pull<0, 4>(V, ro, ob, rcols, deg, pr, np, s, kLongDegree, lr, nl);
    case 8: return launch_pull<0, 8>(V, ro, ob, rcols, deg, pr, np, s, kLongDegree, lr, nl);
    case 16: return launch_pull<0, 16>(V, ro, ob, rcols, deg, pr, np, s, kLongDegree, lr, nl);
    case 32: return launch_pull<0, 32>(V, ro, ob, rcols, deg, pr, np, s, kLongDegree, lr, nl);
    default: return launch_pull<0, 64>(V, ro, ob, rcols, deg, pr, np, s, kLongDegree, lr, nl);
    }
}

extern "C" int bw_pr_vertex_pull(uint64_t V, const void *roff, int32_t ob, const uint32_t *rcols,
                                 const uint32_t *deg, const float *pr, float *np, const bw_launch *cfg)
{ return pull_entry<true>(V, roff, ob, rcols, deg, pr, np, cfg); }

extern "C" int bw_pr_vertex_pull_nodiv(uint64_t V, const void *roff, int32_t ob, const uint32_t *rcols,
                                       const float *pr, float *np, const bw_launch *cfg)
{ return pull_entry<false>(V, roff, ob, rcols, nullptr, pr, np, cfg); }

extern "C" int bw_pr_vertex_push_warp(uint64_t V, const void *off, int32_t ob, const uint32_t *cols,
                                      const float *pr, float *np, const bw_launch *cfg)
{
    if (V == 0) return BW_HIP_OK;
    if (!off || !cols || !pr || !np) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    uint32_t W, C, block, grid;
    BW_TRY(warp_params(cfg, W, C));
    BW_TRY(resolve_geometry(cfg, (V + C - 1) / C * W, block, grid));
    return with_offsets(off, ob, [&](auto *o) {
        using Off = std::remove_const_t<std::remove_pointer_t<decltype(o)>>;
        const uint32_t lds = pr_warp_lds_bytes<Off>(block / W, C);
        if (lds > 64 * 1024) return fail(BW_HIP_ERR_ARG, "Insufficient shared memory for warp/chunk sizes");
        hipLaunchKernelGGL((pr_vertex_push_warp_kernel<Off>), dim3(grid), dim3(block), lds, stream_of(cfg), W, C, V,
                           o, cols, pr, np);
        BW_LAUNCH_CHECK();
        return BW_HIP_OK;
    });
}

template<bool DIV>
static int pull_warp_entry(uint64_t V, const void *roff, int32_t ob, const uint32_t *rcols, const uint32_t *deg,
                           const float *pr, float *np, const bw_launch *cfg)
{
    if (V == 0) return BW_HIP_OK;
    if (!roff || !rcols || !pr || !np || (DIV && !deg)) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    uint32_t W, C;
    BW_TRY(warp_params(cfg, W, C));
    constexpr int D = DIV ? 2 : 0;
    hipStream_t s = stream_of(cfg);
    switch (W) {
    case 1: return launch_pull<D, 1>(V, roff, ob, rcols, deg, pr, np, s);
    case 2: return launch_pull<D, 2>(V, roff, ob, rcols, deg, pr, np, s);
    case 4: return launch_pull<D, 4>(V, roff, ob, rcols, deg, pr, np, s);
    case 8: return launch_pull<D, 8>(V, roff, ob, rcols, deg, pr, np, s);
    case 16: return launch_pull<D, 16>(V, roff, ob, rcols, deg, pr, np, s);
    case 32: return launch_pull<D, 32>(V, roff, ob, rcols, deg, pr, np, s);
    default: return launch_pull<D, 64>(V, roff, ob, rcols, deg, pr, np, s);
    }
}

extern "C" int bw_pr_vertex_pull_warp(uint64_t V, const void *roff, int32_t ob, const uint32_t *rcols,
                                      const uint32_t *deg, const float *pr, float *np, const bw_launch *cfg)
{ return pull_warp_entry<true>(V, roff, ob, rcols, deg, pr, np, cfg); }

extern "C" int bw_pr_vertex_pull_warp_nodiv(uint64_t V, const void *roff, int32_t ob, const uint32_t *rcols,
                                            const float *pr, float *np, const bw_launch *cfg)
{ return pull_warp_entry<false>(V, roff, ob, rcols, nullptr, pr, np, cfg); }

extern "C" int bw_pr_consolidate(uint64_t V, const uint32_t *deg, float *pr, float *np, int32_t not_last,
                                 int32_t nodiv, int32_t mode, const double *dangling, double *resid,
                                 const bw_launch *cfg)
{
    if (V == 0) return BW_HIP_OK;
    if (!pr || !np) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    if (mode != BW_PR_DANGLING_DROP && mode != BW_PR_DANGLING_UNIFORM)
        return fail(BW_HIP_ERR_ARG, "unknown PageRank mode");
    if ((nodiv || mode == BW_PR_DANGLING_UNIFORM) && !deg)
        return fail(BW_HIP_ERR_ARG, "consolidate needs out-degrees in this mode");
    if (mode == BW_PR_DANGLING_UNIFORM && !dangling)
        return fail(BW_HIP_ERR_ARG, "UNIFORM mode needs the dangling mass");
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, V, block, grid, 1024));
    grid = reduction_grid(cfg, grid);
    hipLaunchKernelGGL(pr_consolidate_kernel, dim3(grid), dim3(block), 0, stream_of(cfg), V, deg, pr, np,
                       not_last ? 1 : 0, nodiv ? 1 : 0, mode, dangling, resid);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

namespace bw {
namespace {
__global__ void pr_predivide_kernel(uint64_t V, const uint32_t *__restrict__ deg, float *pr)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride)
        if (deg[v] != 0) pr[v] = pr[v] / static_cast<float>(deg[v]);
}
}  // namespace
}  // namespace bw

extern "C" int bw_pr_predivide(uint64_t V, const uint32_t *deg, float *pr, const bw_launch *cfg)
{
    if (V == 0) return BW_HIP_OK;
    if (!deg || !pr) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, V, block, grid));
    hipLaunchKernelGGL(pr_predivide_kernel, dim3(grid), dim3(block), 0, stream_of(cfg), V, deg, pr);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_pr_dangling_mass(uint64_t V, const uint32_t *deg, const float *pr, double *out,
                                   const bw_launch *cfg)
{
    if (V == 0) return BW_HIP_OK;
    if (!deg || !pr || !out) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, V, block, grid, 1024));
    grid = reduction_grid(cfg, grid);
    hipLaunchKernelGGL(pr_dangling_kernel, dim3(grid), dim3(block), 0, stream_of(cfg), V, deg, pr, out);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

namespace {

// Device scratch of one bw_pr_run call.
struct RunBuffers {
    uint32_t *deg = nullptr;
    float *np = nullptr;
    float *contrib = nullptr;  // per-vertex contributions (reverse edge lists)
    double *acc = nullptr;  // [iterations] residuals, then [iterations] dangling masses
    ~RunBuffers()
    {
        if (deg) (void)hipFree(deg);
        if (np) (void)hipFree(np);
        if (contrib) (void)hipFree(contrib);
        if (acc) (void)hipFree(acc);
    }
};

}  // namespace

extern "C" int bw_pr_run(const bw_graph_dev *g, int32_t variant, int32_t iterations, int32_t mode,
                         float *ranks, double *resid_host, const bw_launch *cfg_in)
{
    if (!g || !ranks || iterations < 0) return fail(BW_HIP_ERR_ARG, "bw_pr_run: invalid argument");
    if (variant < 0 || variant >= BW_PR_NUM_REFERENCE_VARIANTS)
        return fail(BW_HIP_ERR_ARG, "unknown PageRank variant " + std::to_string(variant));
    if (mode != BW_PR_DANGLING_DROP && mode != BW_PR_DANGLING_UNIFORM)
        return fail(BW_HIP_ERR_ARG, "unknown PageRank mode");
    const uint64_t V = g->vertex_count, E = g->edge_count;
    if (V == 0) return BW_HIP_OK;
    bw_launch cfg{};
    if (cfg_in) cfg = *cfg_in;
    cfg.grid = 0;  // every kernel sizes its own grid
    hipStream_t s = static_cast<hipStream_t>(cfg.stream);
    const bool nodiv = variant == BW_PR_VERTEX_PULL_NODIV || variant == BW_PR_VERTEX_PULL_WARP_NODIV;

    auto need = [&](const void *p, const char *what) -> int {
        if (!p && E) return fail(BW_HIP_ERR_ARG, std::string("variant needs the ") + what + " representation");
        return BW_HIP_OK;
    };
    switch (variant) {
    case BW_PR_EDGE_LIST: BW_TRY(need(g->el_in, "EdgeList")); BW_TRY(need(g->el_out, "EdgeList")); break;
    case BW_PR_REV_EDGE_LIST: BW_TRY(need(g->rel_in, "reverse EdgeList")); BW_TRY(need(g->rel_out, "reverse EdgeList")); break;
    case BW_PR_STRUCT_EDGE_LIST: BW_TRY(need(g->sel, "StructEdgeList")); break;
    case BW_PR_REV_STRUCT_EDGE_LIST: BW_TRY(need(g->rsel, "reverse StructEdgeList")); break;
    case BW_PR_VERTEX_PUSH:
    case BW_PR_VERTEX_PUSH_WARP:
        if (!g->offsets || (E && !g->cols)) return fail(BW_HIP_ERR_ARG, "variant needs the CSR representation");
        break;
    default:
        if (!g->rev_offsets || (E && !g->rev_cols))
            return fail(BW_HIP_ERR_ARG, "variant needs the reverse CSR representation");
    }

    RunBuffers buf;
    BW_HIP_CHECK(hipMalloc(&buf.deg, sizeof(uint32_t) * V));
    BW_HIP_CHECK(hipMalloc(&buf.np, sizeof(float) * V));
    const uint64_t nacc = 2 * static_cast<uint64_t>(std::max(iterations, 1));
    BW_HIP_CHECK(hipMalloc(&buf.acc, sizeof(double) * nacc));
    hipLaunchKernelGGL(zero_f64_kernel, dim3((nacc + 255) / 256), dim3(256), 0, s, buf.acc, nacc);
    BW_LAUNCH_CHECK();
    double *resid = buf.acc, *dang = buf.acc + std::max(iterations, 1);

    // Degrees once per run (pagerank/pagerank.cpp:102-112); the push
    // variants have none in the reference and get them from the CSR only
    // when UNIFORM mode needs the dangling vertices.
    BW_TRY(bw_pr_zero_degrees(V, buf.deg, &cfg));
    // With the forward CSR resident, every variant's out-degrees come from
    // its offsets: the same integers the variant's own degree kernel counts
    // (pagerank/pagerank.cu:40-123), without an atomic per edge.
    const bool needDeg = !(variant == BW_PR_VERTEX_PUSH || variant == BW_PR_VERTEX_PUSH_WARP) ||
                         mode == BW_PR_DANGLING_UNIFORM;
    if (needDeg && g->offsets) {
        BW_TRY(bw_pr_degrees_csr(V, g->offsets, g->offset_bytes, buf.deg, &cfg));
    } else switch (variant) {
    case BW_PR_EDGE_LIST: BW_TRY(bw_pr_degrees_edge_list(E, g->el_in, buf.deg, &cfg)); break;
    case BW_PR_REV_EDGE_LIST: BW_TRY(bw_pr_degrees_edge_list(E, g->rel_out, buf.deg, &cfg)); break;
    case BW_PR_STRUCT_EDGE_LIST: BW_TRY(bw_pr_degrees_struct_edge_list(E, g->sel, 0, buf.deg, &cfg)); break;
    case BW_PR_REV_STRUCT_EDGE_LIST: BW_TRY(bw_pr_degrees_struct_edge_list(E, g->rsel, 1, buf.deg, &cfg)); break;
    case BW_PR_VERTEX_PUSH:
    case BW_PR_VERTEX_PUSH_WARP:
        if (mode == BW_PR_DANGLING_UNIFORM) BW_TRY(bw_pr_degrees_csr(V, g->offsets, g->offset_bytes, buf.deg, &cfg));
        break;
    default:
        BW_TRY(bw_pr_degrees_rev_csr(V, g->rev_offsets, g->offset_bytes, g->rev_cols, buf.deg, &cfg));
    }

    {
        uint32_t block, grid;
        BW_TRY(resolve_geometry(nullptr, V, block, grid));
        hipLaunchKernelGGL(pr_init_kernel, dim3(grid), dim3(block), 0, s, V, ranks, buf.np, buf.deg,
                           (nodiv && mode == BW_PR_DANGLING_UNIFORM) ? 1 : 0);
        BW_LAUNCH_CHECK();
    }

    // The push forms scatter through a binned layout (propagation blocking)
    // instead of one global atomic per edge; a graph the layout refuses
    // (beyond 2^25 vertices or 2^32 - 1 edges) keeps the atomic kernels.
    bw_pr_scatter *scatter = nullptr;
    struct ScatterGuard {
        bw_pr_scatter *&p;
        ~ScatterGuard() { bw_pr_scatter_destroy(p); }
    } scatter_guard{scatter};
    if (E > 0 && (variant == BW_PR_EDGE_LIST || variant == BW_PR_STRUCT_EDGE_LIST || variant == BW_PR_VERTEX_PUSH ||
                  variant == BW_PR_VERTEX_PUSH_WARP)) {
        int st = BW_HIP_OK;
        if (variant == BW_PR_EDGE_LIST)
            st = bw_pr_scatter_create(&scatter, BW_PR_SCATTER_EDGE_LIST, V, E, g->el_in, g->el_out, 4, s);
        else if (variant == BW_PR_STRUCT_EDGE_LIST)
            st = bw_pr_scatter_create(&scatter, BW_PR_SCATTER_STRUCT, V, E, g->sel, nullptr, 4, s);
        else
            st = bw_pr_scatter_create(&scatter, BW_PR_SCATTER_CSR, V, E, g->offsets, g->cols, g->offset_bytes, s);
        if (st != BW_HIP_OK && st != BW_HIP_ERR_ARG) return st;
    }

    // The pull forms split off their long rows (one wave each) once per run.
    bw_pr_pull_plan *pull = nullptr;
    struct PullGuard {
        bw_pr_pull_plan *&p;
        ~PullGuard() { bw_pr_pull_plan_destroy(p); }
    } pull_guard{pull};
    if (variant == BW_PR_VERTEX_PULL || variant == BW_PR_VERTEX_PULL_NODIV || variant == BW_PR_VERTEX_PULL_WARP ||
        variant == BW_PR_VERTEX_PULL_WARP_NODIV)
        BW_TRY(bw_pr_pull_plan_create(&pull, V, g->rev_offsets, g->offset_bytes, s));

    for (int32_t it = 1; it <= iterations; it++) {
        if (mode == BW_PR_DANGLING_UNIFORM) BW_TRY(bw_pr_dangling_mass(V, buf.deg, ranks, &dang[it - 1], &cfg));
        if (pull) {
            BW_TRY(bw_pr_pull_iteration(pull, variant, g->rev_cols, buf.deg, ranks, buf.np, &cfg));
            BW_TRY(bw_pr_consolidate(V, buf.deg, ranks, buf.np, iterations > it, nodiv, mode,
                                     mode == BW_PR_DANGLING_UNIFORM ? &dang[it - 1] : nullptr, &resid[it - 1], &cfg));
            continue;
        }
        if (variant == BW_PR_REV_EDGE_LIST || variant == BW_PR_REV_STRUCT_EDGE_LIST) {
            // One gather per edge: each source's contribution computed once.
            if (!buf.contrib) BW_HIP_CHECK(hipMalloc(&buf.contrib, sizeof(float) * V));
            const uint64_t blocks = std::min<uint64_t>((V + 255) / 256, 8ull * device_cus());
            hipLaunchKernelGGL(pr_contrib_kernel<false>, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s, V,
                               buf.deg, ranks, buf.contrib);
            BW_LAUNCH_CHECK();
            if (E) {
                if (variant == BW_PR_REV_EDGE_LIST)
                    BW_TRY((launch_rev_segmented<false, true>(E, g->rel_in, g->rel_out, nullptr, buf.contrib, buf.np,
                                                              s)));
                else
                    BW_TRY((launch_rev_segmented<true, true>(E, g->rsel, nullptr, nullptr, buf.contrib, buf.np, s)));
            }
            const bool useDeg = nodiv || mode == BW_PR_DANGLING_UNIFORM;
            BW_TRY(bw_pr_consolidate(V, useDeg ? buf.deg : nullptr, ranks, buf.np, iterations > it, nodiv, mode,
                                     mode == BW_PR_DANGLING_UNIFORM ? &dang[it - 1] : nullptr, &resid[it - 1], &cfg));
            continue;
        }
        if (scatter) {
            BW_TRY(bw_pr_scatter_iteration(scatter, buf.deg, ranks, buf.np, s));
            const bool useDeg = nodiv || mode == BW_PR_DANGLING_UNIFORM;
            BW_TRY(bw_pr_consolidate(V, useDeg ? buf.deg : nullptr, ranks, buf.np, iterations > it, nodiv, mode,
                                     mode == BW_PR_DANGLING_UNIFORM ? &dang[it - 1] : nullptr, &resid[it - 1], &cfg));
            continue;
        }
        switch (variant) {
        case BW_PR_EDGE_LIST: BW_TRY(bw_pr_edge_list(E, g->el_in, g->el_out, buf.deg, ranks, buf.np, &cfg)); break;
        case BW_PR_REV_EDGE_LIST:
            BW_TRY(bw_pr_rev_edge_list(E, g->rel_in, g->rel_out, buf.deg, ranks, buf.np, &cfg));
            break;
        case BW_PR_STRUCT_EDGE_LIST: BW_TRY(bw_pr_struct_edge_list(E, g->sel, buf.deg, ranks, buf.np, &cfg)); break;
        case BW_PR_REV_STRUCT_EDGE_LIST:
            BW_TRY(bw_pr_rev_struct_edge_list(E, g->rsel, buf.deg, ranks, buf.np, &cfg));
            break;
        case BW_PR_VERTEX_PUSH:
            BW_TRY(bw_pr_vertex_push(V, g->offsets, g->offset_bytes, g->cols, ranks, buf.np, &cfg));
            break;
        case BW_PR_VERTEX_PUSH_WARP:
            BW_TRY(bw_pr_vertex_push_warp(V, g->offsets, g->offset_bytes, g->cols, ranks, buf.np, &cfg));
            break;
        case BW_PR_VERTEX_PULL:
            BW_TRY(bw_pr_vertex_pull(V, g->rev_offsets, g->offset_bytes, g->rev_cols, buf.deg, ranks, buf.np, &cfg));
            break;
        case BW_PR_VERTEX_PULL_NODIV:
            BW_TRY(bw_pr_vertex_pull_nodiv(V, g->rev_offsets, g->offset_bytes, g->rev_cols, ranks, buf.np, &cfg));
            break;
        case BW_PR_VERTEX_PULL_WARP:
            BW_TRY(bw_pr_vertex_pull_warp(V, g->rev_offsets, g->offset_bytes, g->rev_cols, buf.deg, ranks, buf.np,
                                          &cfg));
            break;
        case BW_PR_VERTEX_PULL_WARP_NODIV:
            BW_TRY(bw_pr_vertex_pull_warp_nodiv(V, g->rev_offsets, g->offset_bytes, g->rev_cols, ranks, buf.np,
                                                &cfg));
            break;
        }
        const bool useDeg = nodiv || mode == BW_PR_DANGLING_UNIFORM;
        BW_TRY(bw_pr_consolidate(V, useDeg ? buf.deg : nullptr, ranks, buf.np, iterations > it, nodiv, mode,
                                 mode == BW_PR_DANGLING_UNIFORM ? &dang[it - 1] : nullptr, &resid[it - 1], &cfg));
    }
    if (resid_host && iterations > 0) {
        BW_HIP_CHECK(hipMemcpyAsync(resid_host, resid, sizeof(double) * iterations, hipMemcpyDeviceToHost, s));
    }
    BW_HIP_CHECK(hipStreamSynchronize(s));
    return BW_HIP_OK;
}
