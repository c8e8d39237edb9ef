// Measurement probes (libbw_probes.so): a streaming copy that calibrates the
// HBM byte counters and a gather-rate probe that bounds random 4-byte
// gathers.  Kept out of the kernel library so that its ABI holds only the
// path's own entry points.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "bw_probes.h"

namespace {

thread_local std::string g_error;

int fail(const std::string &msg)
{
    g_error = msg;
    return 1;
}

#define PROBE_CHECK(call)                                                                            \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return (g_error = std::string(#call) + ": " + hipGetErrorString(e_), 2); \
    } while (0)

typedef uint32_t copy_u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void copy_u32x4_kernel(copy_u32x4 *__restrict__ dst,
                                                         const copy_u32x4 *__restrict__ src, uint64_t n4)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

__global__ void copy_u32_tail_kernel(uint32_t *dst, const uint32_t *src, uint64_t begin, uint64_t n)
{
    const uint64_t i = begin + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Thread t sums table[idx[r * threads + t]] over its rounds, kBatch loads in
// flight at a time, so consecutive lanes of a wave read consecutive index
// entries (the access shape of a row tile's gathers).
constexpr int kBatch = 16;

__global__ __launch_bounds__(256) void gather_kernel(const uint32_t *__restrict__ idx,
                                                     const float *__restrict__ table, uint32_t len,
                                                     float *__restrict__ out, uint32_t rounds)
{
    const uint64_t threads = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float acc = 0.0f;
    for (uint32_t r = 0; r < rounds; r++) {
        uint32_t c[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++)
            c[k] = __builtin_nontemporal_load(&idx[(static_cast<uint64_t>(r) * kBatch + k) * threads + t]);
        float v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) v[k] = c[k] < len ? table[c[k]] : 0.0f;
#pragma unroll
        for (int k = 0; k < kBatch; k++) acc += v[k];
    }
    out[t] = acc;
}

}  // namespace

extern "C" const char *bw_probe_last_error(void) { return g_error.c_str(); }

extern "C" int bw_probe_copy(uint32_t *dst, const uint32_t *src, uint64_t n, void *stream)
{
    if (n == 0) return 0;
    if (!dst || !src) return fail("NULL pointer in bw_probe_copy");
    if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15)
        return fail("bw_probe_copy needs 16-byte aligned buffers");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t n4 = n / 4;
    if (n4) {
        const uint64_t grid = (n4 + 255) / 256;
        if (grid > 0x7FFFFFFFull) return fail("bw_probe_copy: more than 2^31 workgroups");
        hipLaunchKernelGGL(copy_u32x4_kernel, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s,
                           reinterpret_cast<copy_u32x4 *>(dst), reinterpret_cast<const copy_u32x4 *>(src), n4);
        PROBE_CHECK(hipGetLastError());
    }
    if (n4 * 4 < n) {
        hipLaunchKernelGGL(copy_u32_tail_kernel, dim3(1), dim3(64), 0, s, dst, src, n4 * 4, n);
        PROBE_CHECK(hipGetLastError());
    }
    return 0;
}

extern "C" int bw_probe_gather(const uint32_t *idx, uint64_t n, const float *table, uint32_t table_len, float *out,
                               uint32_t threads, float *kernel_ms, void *stream)
{
    if (!idx || !table || !out || table_len == 0) return fail("bw_probe_gather: invalid argument");
    if (threads == 0 || threads % 256 != 0) return fail("bw_probe_gather: threads must be a multiple of 256");
    const uint64_t per_round = static_cast<uint64_t>(threads) * kBatch;
    if (n == 0 || n % per_round != 0 || n / per_round > 0xFFFFFFFFull)
        return fail("bw_probe_gather: n must be a positive multiple of 16 * threads");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (kernel_ms) {
        PROBE_CHECK(hipEventCreate(&e0));
        PROBE_CHECK(hipEventCreate(&e1));
        PROBE_CHECK(hipEventRecord(e0, s));
    }
    hipLaunchKernelGGL(gather_kernel, dim3(threads / 256), dim3(256), 0, s, idx, table, table_len, out,
                       static_cast<uint32_t>(n / per_round));
    PROBE_CHECK(hipGetLastError());
    if (kernel_ms) {
        PROBE_CHECK(hipEventRecord(e1, s));
        PROBE_CHECK(hipEventSynchronize(e1));
        PROBE_CHECK(hipEventElapsedTime(kernel_ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return 0;
}
