// PageRank iteration, degree and consolidate kernels for gfx950 and their
// C ABI entry points.
//
// Each kernel restates one reference kernel of pagerank/*.cu (merijn/
// Belewitte) with the same FP32 arithmetic per contribution: IEEE division
// pr/deg, float atomic adds where the reference pushes, the double/float
// base terms of the two consolidate kernels.  Mapping changes for wave64:
// vector loads on edge streams, per-vertex partial sums of the pull-warp
// kernels reduced inside the virtual warp and stored once (the vertex is
// owned by one virtual warp, so the reference's per-lane atomics are not
// needed), FP64 wave/block reductions for the residual and dangling mass.

#include "common.hpp"

#include <algorithm>
#include <vector>

namespace bw {
namespace {

constexpr float kDamp = 0.85f;  // pagerank/pagerank.hpp:9

__device__ __forceinline__ void add_f32(float *p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void add_f64(double *p, double v) { unsafeAtomicAdd(p, v); }

// ---------------------------------------------------------------- degrees

__global__ __launch_bounds__(1024) void pr_degrees_origins_kernel(uint64_t E, const uint32_t *__restrict__ origins,
                                                                  uint32_t *deg)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < E; i += stride)
        atomicAdd(&deg[origins[i]], 1u);
}

__global__ __launch_bounds__(1024) void pr_degrees_struct_kernel(uint64_t E, const uint32_t *__restrict__ edges,
                                                                 int rev, uint32_t *deg)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < E; i += stride)
        atomicAdd(&deg[edges[2 * i + rev]], 1u);
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

// pagerank/vertex_pull.cu:16-30 (DIV) and vertex_pull_nodiv.cu:16-28:
// sequential FP32 sum in CSR order, one thread per vertex, so the result is
// the reference's bit for bit.
template<bool DIV, typename Off>
__global__ __launch_bounds__(1024) void pr_vertex_pull_kernel(uint64_t V, const Off *__restrict__ roff,
                                                              const uint32_t *__restrict__ rcols,
                                                              const uint32_t *__restrict__ deg,
                                                              const float *__restrict__ pr, float *np)
{
    const uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
        float acc = 0.0f;
        const uint64_t s = roff[v], e = roff[v + 1];
        for (uint64_t j = s; j < e; j++) {
            const uint32_t u = rcols[j];
            if constexpr (DIV) acc += pr[u] / static_cast<float>(deg[u]);
            else acc += pr[u];
        }
        np[v] = acc;
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

// pagerank/vertex_pull_warp.cu:4-49 (DIV) and vertex_pull_warp_nodiv.cu:4-47.
template<bool DIV, typename Off>
__global__ __launch_bounds__(1024) void pr_vertex_pull_warp_kernel(uint32_t W, uint32_t C, uint64_t V,
                                                                   const Off *__restrict__ roff,
                                                                   const uint32_t *__restrict__ rcols,
                                                                   const uint32_t *__restrict__ deg,
                                                                   const float *__restrict__ pr, float *np)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t warps = blockDim.x / W;
    const uint32_t bw = threadIdx.x / W, lane = threadIdx.x % W;
    Off *myOff = reinterpret_cast<Off *>(lds_raw + pr_warp_vals_bytes<Off>(warps, C)) + bw * (C + 1);
    const uint64_t warpId = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / W;
    const uint64_t totalWarps = static_cast<uint64_t>(warps) * gridDim.x;
    for (uint64_t chunk = warpId; chunk * C < V; chunk += totalWarps) {
        const uint64_t base = chunk * C;
        const uint32_t end = static_cast<uint32_t>(std::min<uint64_t>(C, V - base));
        for (uint32_t i = lane; i <= end; i += W) myOff[i] = roff[base + i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint32_t v = 0; v < end; v++) {
            const uint64_t s = myOff[v];
            const uint32_t n = static_cast<uint32_t>(myOff[v + 1] - s);
            float part = 0.0f;
            for (uint32_t i = lane; i < n; i += W) {
                const uint32_t u = rcols[s + i];
                if constexpr (DIV) part += pr[u] / static_cast<float>(static_cast<int32_t>(deg[u]));
                else part += pr[u];
            }
            // Sum of the lanes' partials inside the virtual warp.
            for (uint32_t o = W / 2; o > 0; o >>= 1) part += __shfl_xor(part, static_cast<int>(o), kWave);
            if (lane == 0) np[base + v] += part;
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
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, E, block, grid));
    hipLaunchKernelGGL(pr_degrees_origins_kernel, dim3(grid), dim3(block), 0, stream_of(cfg), E, origins, deg);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_pr_degrees_struct_edge_list(uint64_t E, const uint32_t *edges, int32_t reverse, uint32_t *deg,
                                              const bw_launch *cfg)
{
    if (E == 0) return BW_HIP_OK;
    if (!edges || !deg) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, E, block, grid));
    hipLaunchKernelGGL(pr_degrees_struct_kernel, dim3(grid), dim3(block), 0, stream_of(cfg), E, edges,
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

extern "C" int bw_pr_rev_edge_list(uint64_t E, const uint32_t *in, const uint32_t *out, const uint32_t *deg,
                                   const float *pr, float *np, const bw_launch *cfg)
{
    if (E == 0) return BW_HIP_OK;
    if (!in || !out || !deg || !pr || !np) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    const int vec = aligned16(in) && aligned16(out);
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, vec ? (E + 3) / 4 : E, block, grid));
    hipLaunchKernelGGL(pr_edge_list_kernel<true>, dim3(grid), dim3(block), 0, stream_of(cfg), E, in, out, deg, pr,
                       np, vec);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

static int struct_entry(bool rev, uint64_t E, const uint32_t *edges, const uint32_t *deg, const float *pr,
                        float *np, const bw_launch *cfg)
{
    if (E == 0) return BW_HIP_OK;
    if (!edges || !deg || !pr || !np) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    const int vec = aligned16(edges);
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, vec ? (E + 1) / 2 : E, block, grid));
    if (rev)
        hipLaunchKernelGGL(pr_struct_edge_list_kernel<true>, dim3(grid), dim3(block), 0, stream_of(cfg), E, edges,
                           deg, pr, np, vec);
    else
        hipLaunchKernelGGL(pr_struct_edge_list_kernel<false>, dim3(grid), dim3(block), 0, stream_of(cfg), E, edges,
                           deg, pr, np, vec);
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

template<bool DIV>
static int pull_entry(uint64_t V, const void *roff, int32_t ob, const uint32_t *rcols, const uint32_t *deg,
                      const float *pr, float *np, const bw_launch *cfg)
{
    if (V == 0) return BW_HIP_OK;
    if (!roff || !rcols || !pr || !np || (DIV && !deg)) return fail(BW_HIP_ERR_ARG, "NULL device pointer");
    uint32_t block, grid;
    BW_TRY(resolve_geometry(cfg, V, block, grid));
    return with_offsets(roff, ob, [&](auto *o) {
        using Off = std::remove_const_t<std::remove_pointer_t<decltype(o)>>;
        hipLaunchKernelGGL((pr_vertex_pull_kernel<DIV, Off>), dim3(grid), dim3(block), 0, stream_of(cfg), V, o, rcols,
                           deg, pr, np);
        BW_LAUNCH_CHECK();
        return BW_HIP_OK;
    });
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
    uint32_t W, C, block, grid;
    BW_TRY(warp_params(cfg, W, C));
    BW_TRY(resolve_geometry(cfg, (V + C - 1) / C * W, block, grid));
    return with_offsets(roff, ob, [&](auto *o) {
        using Off = std::remove_const_t<std::remove_pointer_t<decltype(o)>>;
        const uint32_t lds = pr_warp_lds_bytes<Off>(block / W, C);
        if (lds > 64 * 1024) return fail(BW_HIP_ERR_ARG, "Insufficient shared memory for warp/chunk sizes");
        hipLaunchKernelGGL((pr_vertex_pull_warp_kernel<DIV, Off>), dim3(grid), dim3(block), lds, stream_of(cfg), W,
                           C, V, o, rcols, deg, pr, np);
        BW_LAUNCH_CHECK();
        return BW_HIP_OK;
    });
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
    BW_TRY(resolve_geometry(cfg, V, block, grid));
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
    BW_TRY(resolve_geometry(cfg, V, block, grid));
    hipLaunchKernelGGL(pr_dangling_kernel, dim3(grid), dim3(block), 0, stream_of(cfg), V, deg, pr, out);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

namespace {

// Device scratch of one bw_pr_run call.
struct RunBuffers {
    uint32_t *deg = nullptr;
    float *np = nullptr;
    double *acc = nullptr;  // [iterations] residuals, then [iterations] dangling masses
    ~RunBuffers()
    {
        if (deg) (void)hipFree(deg);
        if (np) (void)hipFree(np);
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
    switch (variant) {
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

    for (int32_t it = 1; it <= iterations; it++) {
        if (mode == BW_PR_DANGLING_UNIFORM) BW_TRY(bw_pr_dangling_mass(V, buf.deg, ranks, &dang[it - 1], &cfg));
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
