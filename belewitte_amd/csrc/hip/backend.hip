// HIP backend of the C ABI: device query, allocations, copies, streams,
// events and launch geometry.  Replaces CUDA.hpp/CUDA.cpp and
// utils/cuda_utils.hpp of the reference (paths relative to the reference
// repository); see include/belewitte_hip.h for the contract.

#include "common.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#ifndef BW_VERSION
#define BW_VERSION "belewitte-amd-dev"
#endif

namespace bw {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    std::string msg = std::string("HIP error #") + std::to_string(static_cast<int>(e)) + " ("
                    + file + ":" + std::to_string(line) + "): " + hipGetErrorString(e)
                    + " in " + what;
    set_error(msg);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return BW_HIP_ERR_NOMEM;
    if (e == hipErrorNoDevice) return BW_HIP_ERR_NODEVICE;
    return BW_HIP_ERR_RUNTIME;
}

int device_cus()
{
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lock(mu);
    if (static_cast<size_t>(dev) >= cache.size()) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cache[dev] = cus;
    }
    return cache[dev];
}

int resolve_geometry(const bw_launch *cfg, uint64_t work, uint32_t &block,
                     uint32_t &grid, uint32_t default_block)
{
    block = (cfg && cfg->block) ? cfg->block : default_block;
    if (block % kWave != 0 || block > 1024)
        return fail(BW_HIP_ERR_ARG, "block size must be a multiple of 64 and at most 1024, got "
                                    + std::to_string(block));
    if (cfg && cfg->grid) {
        grid = cfg->grid;
    } else {
        // Enough blocks to cover the work once, capped so a grid-stride
        // loop keeps every CU busy with several waves.
        uint64_t need = (work + block - 1) / block;
        uint64_t cap = static_cast<uint64_t>(device_cus()) * 2048 / block * 8;
        grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(need, cap)));
    }
    if (grid == 0) return fail(BW_HIP_ERR_ARG, "grid size must be positive");
    return BW_HIP_OK;
}

__global__ void fill_u32_kernel(uint32_t *dst, uint32_t value, uint64_t n)
{
    uint64_t stride = static_cast<uint64_t>(blockDim.x) * gridDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = value;
}

}  // namespace bw

using namespace bw;

extern "C" const char *bw_hip_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *bw_hip_version(void) { return BW_VERSION; }

extern "C" int bw_hip_device_count(int32_t *count)
{
    if (!count) return fail(BW_HIP_ERR_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice || e == hipErrorInsufficientDriver) {
        *count = 0;
        return BW_HIP_OK;
    }
    BW_HIP_CHECK(e);
    *count = n;
    return BW_HIP_OK;
}

extern "C" int bw_hip_set_device(int32_t device)
{
    int32_t n = 0;
    BW_TRY(bw_hip_device_count(&n));
    if (device < 0 || device >= n)
        return fail(BW_HIP_ERR_NODEVICE, "Non-existent device #" + std::to_string(device)
                                         + ", device count is " + std::to_string(n));
    BW_HIP_CHECK(hipSetDevice(device));
    return BW_HIP_OK;
}

extern "C" int bw_hip_get_device_props(int32_t device, bw_hip_props *p)
{
    if (!p) return fail(BW_HIP_ERR_ARG, "props is NULL");
    hipDeviceProp_t prop;
    BW_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    std::memset(p, 0, sizeof *p);
    std::snprintf(p->name, sizeof p->name, "%s", prop.name);
    std::snprintf(p->arch, sizeof p->arch, "%s", prop.gcnArchName);
    p->compute_units = prop.multiProcessorCount;
    p->wave_size = prop.warpSize;
    p->max_threads_per_block = prop.maxThreadsPerBlock;
    for (int i = 0; i < 3; i++) {
        p->max_block_dims[i] = prop.maxThreadsDim[i];
        p->max_grid_dims[i] = prop.maxGridSize[i];
    }
    p->shared_mem_per_block = prop.sharedMemPerBlock;
    p->total_global_mem = prop.totalGlobalMem;
    p->l2_cache_bytes = static_cast<uint64_t>(prop.l2CacheSize);
    p->clock_khz = prop.clockRate;
    p->memory_clock_khz = prop.memoryClockRate;
    p->memory_bus_width_bits = prop.memoryBusWidth;
    return BW_HIP_OK;
}

extern "C" int bw_hip_dev_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(BW_HIP_ERR_ARG, "ptr is NULL");
    *ptr = nullptr;
    BW_HIP_CHECK(hipMalloc(ptr, bytes ? bytes : 4));
    return BW_HIP_OK;
}

extern "C" int bw_hip_dev_free(void *ptr)
{
    if (ptr) BW_HIP_CHECK(hipFree(ptr));
    return BW_HIP_OK;
}

extern "C" int bw_hip_host_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(BW_HIP_ERR_ARG, "ptr is NULL");
    *ptr = nullptr;
    BW_HIP_CHECK(hipHostMalloc(ptr, bytes ? bytes : 4, hipHostMallocDefault));
    return BW_HIP_OK;
}

extern "C" int bw_hip_host_free(void *ptr)
{
    if (ptr) BW_HIP_CHECK(hipHostFree(ptr));
    return BW_HIP_OK;
}

extern "C" int bw_hip_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return BW_HIP_OK;
    if (!dst || !src) return fail(BW_HIP_ERR_ARG, "NULL pointer in memcpy_h2d");
    BW_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BW_HIP_OK;
}

extern "C" int bw_hip_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return BW_HIP_OK;
    if (!dst || !src) return fail(BW_HIP_ERR_ARG, "NULL pointer in memcpy_d2h");
    BW_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BW_HIP_OK;
}

extern "C" int bw_hip_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return BW_HIP_OK;
    if (!dst || !src) return fail(BW_HIP_ERR_ARG, "NULL pointer in memcpy_d2d");
    BW_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return BW_HIP_OK;
}

extern "C" int bw_hip_memcpy_async(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return BW_HIP_OK;
    if (!dst || !src) return fail(BW_HIP_ERR_ARG, "NULL pointer in memcpy_async");
    BW_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, static_cast<hipStream_t>(stream)));
    return BW_HIP_OK;
}

extern "C" int bw_hip_fill_u32(void *dst, uint32_t value, uint64_t count, void *stream)
{
    if (count == 0) return BW_HIP_OK;
    if (!dst) return fail(BW_HIP_ERR_ARG, "NULL pointer in fill_u32");
    uint32_t block = 256, grid = 0;
    BW_TRY(resolve_geometry(nullptr, count, block, grid));
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(block), 0,
                       static_cast<hipStream_t>(stream), static_cast<uint32_t *>(dst), value, count);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_hip_device_sync(void)
{
    BW_HIP_CHECK(hipDeviceSynchronize());
    return BW_HIP_OK;
}

extern "C" int bw_hip_stream_create(void **stream)
{
    if (!stream) return fail(BW_HIP_ERR_ARG, "stream is NULL");
    hipStream_t s;
    BW_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return BW_HIP_OK;
}

extern "C" int bw_hip_stream_destroy(void *stream)
{
    if (stream) BW_HIP_CHECK(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return BW_HIP_OK;
}

extern "C" int bw_hip_stream_sync(void *stream)
{
    BW_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return BW_HIP_OK;
}

extern "C" int bw_hip_event_create(void **event)
{
    if (!event) return fail(BW_HIP_ERR_ARG, "event is NULL");
    hipEvent_t e;
    BW_HIP_CHECK(hipEventCreate(&e));
    *event = e;
    return BW_HIP_OK;
}

extern "C" int bw_hip_event_destroy(void *event)
{
    if (event) BW_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return BW_HIP_OK;
}

extern "C" int bw_hip_event_record(void *event, void *stream)
{
    BW_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    return BW_HIP_OK;
}

extern "C" int bw_hip_event_sync(void *event)
{
    BW_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return BW_HIP_OK;
}

extern "C" int bw_hip_event_elapsed_ms(float *ms, void *start, void *end)
{
    if (!ms) return fail(BW_HIP_ERR_ARG, "ms is NULL");
    BW_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)));
    return BW_HIP_OK;
}

extern "C" int bw_hip_compute_division(uint64_t count, uint64_t *threads, uint64_t *blocks)
{
    if (!threads || !blocks) return fail(BW_HIP_ERR_ARG, "NULL output");
    int dev = 0;
    BW_HIP_CHECK(hipGetDevice(&dev));
    int cus = 0, maxThreads = 0;
    BW_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BW_HIP_CHECK(hipDeviceGetAttribute(&maxThreads, hipDeviceAttributeMaxThreadsPerBlock, dev));
    // Backend.cpp:52-70 with the rounding unit 64 (wave64) instead of 32.
    const uint64_t n = static_cast<uint64_t>(cus), m = static_cast<uint64_t>(maxThreads);
    auto ceil_div = [](uint64_t x, uint64_t y) { return (x + y - 1) / y; };
    uint64_t i = 1;
    uint64_t t = ceil_div(count, n);
    for (i = 1; t > m; i++) t = ceil_div(count, i * n);
    uint64_t adjust = t % kWave;
    t += adjust ? kWave - adjust : 0;
    if (t == 0) t = kWave;
    *threads = t;
    *blocks = std::max<uint64_t>(1, std::min(i * n, ceil_div(count, t)));
    return BW_HIP_OK;
}
