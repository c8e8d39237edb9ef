// Fused pull PageRank for gfx950 ("vertex-pull-nodiv-fused").
//
// Same recurrence as pagerank/vertex_pull_nodiv.cu + consolidateRankNoDiv
// (merijn/Belewitte): contributions c[u] = rank[u]/outdeg(u) are gathered
// along the reverse CSR and the consolidate step is fused into the same
// launch, so one iteration is one kernel and reads
//   offsets 4(V+1) + columns 4E + gathered contributions 4E
//   + out-degree 4V + own old contribution 4V, and writes 4V
// = 8E + 16V + 4 bytes, the HBM roofline of the iteration.
//
// Mapping (CSR-Adaptive, one 256-thread workgroup per work item):
//   * stream blocks: <= 256 consecutive short rows holding <= kStreamNnz
//     edges; the block gathers the contributions of all its edges into LDS
//     with coalesced column reads, then a power-of-two group of lanes per
//     row reduces the row from LDS;
//   * long rows (> kStreamNnz in-edges) are split into pieces of kLongPiece
//     edges, one workgroup per piece; the last piece to finish (counted with
//     an atomic) combines the pieces' partials in piece order and finalises
//     the row.  Long pieces come first in the item list so hub rows start
//     early.
// Every per-vertex sum is Kahan-compensated and combined in a fixed order,
// so ranks are reproducible run to run.  The FP64 dangling mass and L1
// residual are reduced per workgroup and added atomically into 64 slots
// per iteration (the next iteration sums the slots).

#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace bw {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kStreamNnz = 1024;
constexpr uint32_t kLongPiece = 8192;
constexpr int kSlots = 64;
constexpr float kDamp = 0.85f;
constexpr uint32_t kLongFlag = 0x80000000u;

struct LongRow {
    uint32_t row, pieces, base, pad;
};

struct Kahan {
    float s = 0.0f, c = 0.0f;  // c holds minus the lost low-order part
    __device__ __forceinline__ void add(float x)
    {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s - c; }
};

// (s1 - c1) + (s2 - c2) with the rounding error of the head sum kept.
__device__ __forceinline__ void kahan_merge(float &s, float &c, float s2, float c2)
{
    const float t = s + s2;
    const float bp = t - s;
    const float err = (s - (t - bp)) + (s2 - bp);  // TwoSum error
    s = t;
    c = (c + c2) - err;
}

__device__ __forceinline__ void add_slot(double *slots, double v)
{
    if (v != 0.0) unsafeAtomicAdd(&slots[blockIdx.x & (kSlots - 1)], v);
}

struct IterArgs {
    uint32_t V;
    const uint32_t *roff;
    const uint32_t *rcols;
    const uint32_t *deg;
    const float *cin;
    float *cout;
    const uint2 *items;
    uint32_t nitems;
    const LongRow *lrows;
    float2 *partials;
    uint32_t *counters;
    const double *dangl_in;  // kSlots, this iteration's dangling mass
    double *dangl_out;       // kSlots, next iteration's
    double *resid_out;       // kSlots
    int mode;
    int not_last;
    int cin_divided;
};

__device__ __forceinline__ void finalise(const IterArgs &a, uint32_t v, float sum, double D, double &resid,
                                         double &dangl)
{
    const double d = static_cast<double>(kDamp);
    float nr;
    if (a.mode == BW_PR_DANGLING_DROP) {
        // consolidateRankNoDiv, pagerank/pagerank.cu:154 (float base, fma).
        const float baseF = (1.0f - kDamp) / static_cast<float>(a.V);
        nr = __builtin_fmaf(kDamp, sum, baseF);
    } else {
        nr = static_cast<float>((1.0 - d) / static_cast<double>(a.V) +
                                d * (static_cast<double>(sum) + D / static_cast<double>(a.V)));
    }
    const uint32_t dv = a.deg[v];
    const float old = a.cin[v];
    double oldRank = static_cast<double>(old);
    if (a.cin_divided && dv != 0) oldRank *= static_cast<double>(dv);
    resid += fabs(static_cast<double>(nr) - oldRank);
    if (dv == 0) dangl += static_cast<double>(nr);
    a.cout[v] = (dv != 0 && a.not_last) ? nr / static_cast<float>(dv) : nr;
}

__global__ __launch_bounds__(kBlock) void pr_fused_kernel(IterArgs a)
{
    __shared__ float vals[kStreamNnz];
    __shared__ float wsum[kBlock / kWave][2];
    __shared__ double red[kBlock / kWave][2];
    __shared__ double Dsh;
    __shared__ int lastPiece;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wid = tid / kWave;

    if (tid < kWave) {
        double x = (a.mode == BW_PR_DANGLING_UNIFORM) ? a.dangl_in[tid] : 0.0;
        x = wave_sum_f64(x);
        if (tid == 0) Dsh = x;
    }
    __syncthreads();
    const double D = Dsh;
    double resid = 0.0, dangl = 0.0;

    for (uint32_t item = blockIdx.x; item < a.nitems; item += gridDim.x) {
        const uint2 it = a.items[item];
        if (it.y & kLongFlag) {
            const LongRow lr = a.lrows[it.x];
            const uint32_t p = it.y & ~kLongFlag;
            const uint32_t v = lr.row;
            const uint32_t rs = a.roff[v], re = a.roff[v + 1];
            const uint32_t s = rs + p * kLongPiece;
            const uint32_t e = min(s + kLongPiece, re);
            Kahan k;
            for (uint32_t j = s + tid; j < e; j += kBlock) k.add(a.cin[a.rcols[j]]);
            float ps = k.s, pc = k.c;
            for (int o = kWave / 2; o > 0; o >>= 1) {
                const float s2 = __shfl_xor(ps, o, kWave), c2 = __shfl_xor(pc, o, kWave);
                kahan_merge(ps, pc, s2, c2);
            }
            if (lane == 0) {
                wsum[wid][0] = ps;
                wsum[wid][1] = pc;
            }
            __syncthreads();
            if (tid == 0) {
                float bs = wsum[0][0], bc = wsum[0][1];
                for (int w = 1; w < kBlock / kWave; w++) kahan_merge(bs, bc, wsum[w][0], wsum[w][1]);
                a.partials[lr.base + p] = make_float2(bs, bc);
                __threadfence();
                const uint32_t done = atomicAdd(&a.counters[it.x], 1u);
                lastPiece = (done == lr.pieces - 1);
                if (lastPiece) {
                    __threadfence();
                    // Partials of the other pieces are read past the
                    // non-coherent vector cache.
                    float ts = 0.0f, tc = 0.0f;
                    for (uint32_t q = 0; q < lr.pieces; q++) {
                        const uint64_t bits = __hip_atomic_load(
                            reinterpret_cast<uint64_t *>(&a.partials[lr.base + q]), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
                        float2 pq;
                        std::memcpy(&pq, &bits, sizeof pq);
                        kahan_merge(ts, tc, pq.x, pq.y);
                    }
                    finalise(a, v, ts - tc, D, resid, dangl);
                    __hip_atomic_store(&a.counters[it.x], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            __syncthreads();
        } else {
            const uint32_t r0 = it.x, r1 = it.y;
            const uint32_t nrows = r1 - r0;
            const uint32_t e0 = a.roff[r0];
            const uint32_t nnz = a.roff[r1] - e0;
            for (uint32_t j = tid; j < nnz; j += kBlock) vals[j] = a.cin[a.rcols[e0 + j]];
            __syncthreads();
            // Lanes per row: the largest power of two with nrows groups
            // fitting the block, at most one wave.
            uint32_t tpr = kWave;
            while (tpr > 1 && tpr * nrows > kBlock) tpr >>= 1;
            const uint32_t row = tid / tpr, sub = tid % tpr;
            if (row < nrows) {
                const uint32_t v = r0 + row;
                const uint32_t rs = a.roff[v] - e0, re = a.roff[v + 1] - e0;
                Kahan k;
                for (uint32_t j = rs + sub; j < re; j += tpr) k.add(vals[j]);
                float ps = k.s, pc = k.c;
                for (uint32_t o = tpr / 2; o > 0; o >>= 1) {
                    const float s2 = __shfl_xor(ps, static_cast<int>(o), kWave);
                    const float c2 = __shfl_xor(pc, static_cast<int>(o), kWave);
                    kahan_merge(ps, pc, s2, c2);
                }
                if (sub == 0) finalise(a, v, ps - pc, D, resid, dangl);
            }
            __syncthreads();
        }
    }

    // Workgroup totals of the FP64 reductions.
    resid = wave_sum_f64(resid);
    dangl = wave_sum_f64(dangl);
    if (lane == 0) {
        red[wid][0] = resid;
        red[wid][1] = dangl;
    }
    __syncthreads();
    if (tid == 0) {
        double r = 0.0, g = 0.0;
        for (int w = 0; w < kBlock / kWave; w++) {
            r += red[w][0];
            g += red[w][1];
        }
        add_slot(a.resid_out, r);
        add_slot(a.dangl_out, g);
    }
}

// Initial contributions: 1/V (pagerank/pagerank.cpp:88), divided by the
// out-degree unless the reference's undivided first iteration is asked for;
// the initial dangling mass goes to slots[0..63].
__global__ __launch_bounds__(kBlock) void pr_fused_init_kernel(uint32_t V, const uint32_t *deg, float *c,
                                                               int divide, double *slots)
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
    if (threadIdx.x == 0) add_slot(slots, t);
}

__global__ void zero_f64(double *p, uint64_t n)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(blockDim.x) * gridDim.x)
        p[i] = 0.0;
}

}  // namespace
}  // namespace bw

using namespace bw;

struct bw_pr_plan {
    uint32_t V = 0, E = 0;
    uint32_t nitems = 0, nlong = 0, npieces = 0;
    uint2 *items = nullptr;
    LongRow *lrows = nullptr;
    float2 *partials = nullptr;
    uint32_t *counters = nullptr;
    uint32_t *deg = nullptr;
    ~bw_pr_plan()
    {
        for (void *p : {static_cast<void *>(items), static_cast<void *>(lrows), static_cast<void *>(partials),
                        static_cast<void *>(counters), static_cast<void *>(deg)})
            if (p) (void)hipFree(p);
    }
};

extern "C" int bw_pr_plan_create(bw_pr_plan **out, uint64_t V, uint64_t E, const void *rev_offsets,
                                 int32_t offset_bytes, const uint32_t *rev_cols)
{
    if (!out || !rev_offsets || (E && !rev_cols)) return fail(BW_HIP_ERR_ARG, "bw_pr_plan_create: invalid argument");
    if (offset_bytes != 4) return fail(BW_HIP_ERR_ARG, "the fused kernel reads uint32 offsets");
    if (V == 0 || V >= 0x7FFFFFFFull || E > 0xFFFFFFFFull)
        return fail(BW_HIP_ERR_ARG, "fused PageRank needs 0 < V < 2^31 and E < 2^32");
    *out = nullptr;
    std::vector<uint32_t> off(V + 1);
    BW_HIP_CHECK(hipMemcpy(off.data(), rev_offsets, sizeof(uint32_t) * (V + 1), hipMemcpyDeviceToHost));
    if (off[V] != E) return fail(BW_HIP_ERR_ARG, "reverse offsets do not end at the edge count");

    std::vector<uint2> items, stream;
    std::vector<LongRow> lrows;
    uint32_t pieces = 0;
    uint64_t r0 = 0;
    auto close_block = [&](uint64_t r1) {
        if (r1 > r0) stream.push_back(make_uint2(static_cast<uint32_t>(r0), static_cast<uint32_t>(r1)));
    };
    for (uint64_t v = 0; v < V; v++) {
        const uint32_t n = off[v + 1] - off[v];
        if (n > kStreamNnz) {
            close_block(v);
            const uint32_t k = static_cast<uint32_t>(lrows.size());
            const uint32_t np = (n + kLongPiece - 1) / kLongPiece;
            lrows.push_back({static_cast<uint32_t>(v), np, pieces, 0});
            for (uint32_t p = 0; p < np; p++) items.push_back(make_uint2(k, kLongFlag | p));
            pieces += np;
            r0 = v + 1;
            continue;
        }
        if (v - r0 >= static_cast<uint64_t>(kBlock) || off[v + 1] - off[r0] > kStreamNnz) {
            close_block(v);
            r0 = v;
        }
    }
    close_block(V);
    items.insert(items.end(), stream.begin(), stream.end());

    auto plan = new bw_pr_plan();
    plan->V = static_cast<uint32_t>(V);
    plan->E = static_cast<uint32_t>(E);
    plan->nitems = static_cast<uint32_t>(items.size());
    plan->nlong = static_cast<uint32_t>(lrows.size());
    plan->npieces = pieces;
    auto cleanup = [&](int st) { delete plan; return st; };
    if (hipMalloc(&plan->items, sizeof(uint2) * std::max<size_t>(items.size(), 1)) != hipSuccess ||
        hipMalloc(&plan->lrows, sizeof(LongRow) * std::max<size_t>(lrows.size(), 1)) != hipSuccess ||
        hipMalloc(&plan->partials, sizeof(float2) * std::max<uint32_t>(pieces, 1)) != hipSuccess ||
        hipMalloc(&plan->counters, sizeof(uint32_t) * std::max<size_t>(lrows.size(), 1)) != hipSuccess ||
        hipMalloc(&plan->deg, sizeof(uint32_t) * V) != hipSuccess)
        return cleanup(fail(BW_HIP_ERR_NOMEM, "device allocation of the PageRank plan failed"));
    if (hipMemcpy(plan->items, items.data(), sizeof(uint2) * items.size(), hipMemcpyHostToDevice) != hipSuccess ||
        (lrows.size() &&
         hipMemcpy(plan->lrows, lrows.data(), sizeof(LongRow) * lrows.size(), hipMemcpyHostToDevice) != hipSuccess))
        return cleanup(fail(BW_HIP_ERR_RUNTIME, "upload of the PageRank plan failed"));
    int st = bw_hip_fill_u32(plan->counters, 0u, std::max<size_t>(lrows.size(), 1), nullptr);
    if (st == BW_HIP_OK) st = bw_pr_zero_degrees(V, plan->deg, nullptr);
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

extern "C" int bw_pr_plan_info(const bw_pr_plan *plan, uint32_t *nitems, uint32_t *nlong, uint32_t *npieces)
{
    if (!plan) return fail(BW_HIP_ERR_ARG, "plan is NULL");
    if (nitems) *nitems = plan->nitems;
    if (nlong) *nlong = plan->nlong;
    if (npieces) *npieces = plan->npieces;
    return BW_HIP_OK;
}

extern "C" int bw_pr_plan_degrees(const bw_pr_plan *plan, const uint32_t **deg)
{
    if (!plan || !deg) return fail(BW_HIP_ERR_ARG, "invalid argument");
    *deg = plan->deg;
    return BW_HIP_OK;
}

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
    // acc layout: slots [iterations+1][2][kSlots]: dangling(k), residual(k).
    const uint64_t nacc = static_cast<uint64_t>(iterations + 1) * 2 * kSlots;
    hipLaunchKernelGGL(zero_f64, dim3(64), dim3(256), 0, s, acc, nacc);
    BW_LAUNCH_CHECK();
    auto dangl = [&](int k) { return acc + static_cast<uint64_t>(k) * 2 * kSlots; };
    auto resid = [&](int k) { return acc + static_cast<uint64_t>(k) * 2 * kSlots + kSlots; };

    // Ping-pong so the last iteration lands in `ranks`.
    float *buf[2];
    buf[iterations % 2] = ranks;
    buf[(iterations + 1) % 2] = scratch;
    const int divide = !(nodiv_quirk && mode == BW_PR_DANGLING_DROP);
    {
        const uint32_t grid = std::min<uint32_t>((V + kBlock - 1) / kBlock, 4096);
        hipLaunchKernelGGL(pr_fused_init_kernel, dim3(grid), dim3(kBlock), 0, s, V, plan->deg, buf[0], divide,
                           dangl(0));
        BW_LAUNCH_CHECK();
    }
    uint32_t grid = (cfg && cfg->grid) ? cfg->grid : std::min<uint32_t>(plan->nitems, static_cast<uint32_t>(device_cus()) * 16);
    grid = std::max<uint32_t>(grid, 1);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (kernel_ms) {
        BW_HIP_CHECK(hipEventCreate(&ev0));
        BW_HIP_CHECK(hipEventCreate(&ev1));
        BW_HIP_CHECK(hipEventRecord(ev0, s));
    }
    for (int32_t it = 1; it <= iterations; it++) {
        IterArgs a;
        a.V = V;
        a.roff = static_cast<const uint32_t *>(g->rev_offsets);
        a.rcols = g->rev_cols;
        a.deg = plan->deg;
        a.cin = buf[(it - 1) % 2];
        a.cout = buf[it % 2];
        a.items = plan->items;
        a.nitems = plan->nitems;
        a.lrows = plan->lrows;
        a.partials = plan->partials;
        a.counters = plan->counters;
        a.dangl_in = dangl(it - 1);
        a.dangl_out = dangl(it);
        a.resid_out = resid(it - 1);
        a.mode = mode;
        a.not_last = it < iterations;
        a.cin_divided = (it > 1) || divide;
        hipLaunchKernelGGL(pr_fused_kernel, dim3(grid), dim3(kBlock), 0, s, a);
        BW_LAUNCH_CHECK();
    }
    if (kernel_ms) {
        BW_HIP_CHECK(hipEventRecord(ev1, s));
        BW_HIP_CHECK(hipEventSynchronize(ev1));
        BW_HIP_CHECK(hipEventElapsedTime(kernel_ms, ev0, ev1));
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
    }
    if (resid_host) {
        std::vector<double> h(nacc);
        BW_HIP_CHECK(hipMemcpyAsync(h.data(), acc, sizeof(double) * nacc, hipMemcpyDeviceToHost, s));
        BW_HIP_CHECK(hipStreamSynchronize(s));
        for (int k = 0; k < iterations; k++) {
            double r = 0.0;
            for (int j = 0; j < kSlots; j++) r += h[static_cast<uint64_t>(k) * 2 * kSlots + kSlots + j];
            resid_host[k] = r;
        }
    }
    return BW_HIP_OK;
}

extern "C" uint64_t bw_pr_fused_acc_len(int32_t iterations)
{ return static_cast<uint64_t>(iterations + 1) * 2 * kSlots; }
