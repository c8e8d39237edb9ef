// Shared host/device helpers of libbelewitte_hip (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <string>

#include "belewitte_hip.h"

namespace bw {

// ------------------------------------------------------------ errors

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define BW_HIP_CHECK(call)                                                   \
    do {                                                                     \
        hipError_t bw_e_ = (call);                                           \
        if (bw_e_ != hipSuccess)                                             \
            return ::bw::hip_fail(bw_e_, #call, __FILE__, __LINE__);         \
    } while (0)

// Checks the launch that was just enqueued.
#define BW_LAUNCH_CHECK() BW_HIP_CHECK(hipGetLastError())

#define BW_TRY(call)                                                         \
    do {                                                                     \
        int bw_s_ = (call);                                                  \
        if (bw_s_ != BW_HIP_OK) return bw_s_;                                \
    } while (0)

constexpr int kWave = 64;
constexpr int32_t kUnvisited = INT_MAX;

inline hipStream_t stream_of(const bw_launch *cfg)
{ return cfg ? static_cast<hipStream_t>(cfg->stream) : nullptr; }

// Resolves block/grid for `work` items of a grid-stride kernel; the block
// must be a multiple of the wave size so every wave is full (the wave
// reductions rely on it).
int resolve_geometry(const bw_launch *cfg, uint64_t work, uint32_t &block,
                     uint32_t &grid, uint32_t default_block = 256);

// Number of compute units of the current device (cached per device).
int device_cus();

// -------------------------------------------------------- device side

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
    // xor butterfly over the 64 lanes; every lane ends with the total.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ float wave_sum_f32(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block-wide sum of a double; result valid in thread 0.  `lds` holds one
// slot per wave (16 for a 1024-thread block).
__device__ __forceinline__ double block_sum_f64(double v, double *lds)
{
    v = wave_sum_f64(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (wid == 0) {
        const int nw = blockDim.x / kWave;
        t = lane < nw ? lds[lane] : 0.0;
        t = wave_sum_f64(t);
    }
    return t;
}

// Frontier counters: the four Reduction strategies of bfs/bfs.hpp:32-116,
// with wave64 in place of the 32-wide warp.
template<int R>
struct FrontierCount;

template<>
struct FrontierCount<BW_RED_NORMAL> {
    __device__ __forceinline__ void update(uint32_t *frontier)
    { atomicAdd(frontier, 1u); }
    __device__ __forceinline__ void finalise(uint32_t *) {}
};

template<>
struct FrontierCount<BW_RED_BULK> {
    uint32_t count = 0;
    __device__ __forceinline__ void update(uint32_t *) { count++; }
    __device__ __forceinline__ void finalise(uint32_t *frontier)
    { if (count) atomicAdd(frontier, count); }
};

template<>
struct FrontierCount<BW_RED_WAVE> {
    uint32_t count = 0;
    __device__ __forceinline__ void update(uint32_t *) { count++; }
    __device__ __forceinline__ void finalise(uint32_t *frontier)
    {
        uint32_t total = wave_sum_u32(count);
        if ((threadIdx.x & (kWave - 1)) == 0 && total) atomicAdd(frontier, total);
    }
};

template<>
struct FrontierCount<BW_RED_BLOCK> {
    uint32_t count = 0;
    __device__ __forceinline__ void update(uint32_t *) { count++; }
    __device__ __forceinline__ void finalise(uint32_t *frontier)
    {
        __shared__ uint32_t partial[1024 / kWave];
        uint32_t total = wave_sum_u32(count);
        const int lane = threadIdx.x & (kWave - 1);
        const int wid = threadIdx.x / kWave;
        if (lane == 0) partial[wid] = total;
        __syncthreads();
        if (wid == 0) {
            const int nw = blockDim.x / kWave;
            uint32_t t = lane < nw ? partial[lane] : 0u;
            t = wave_sum_u32(t);
            if (lane == 0 && t) atomicAdd(frontier, t);
        }
    }
};

// atomicMin on a depth that only ever decreases within a level: a plain
// load that already shows <= nd proves the atomic would not change the
// value nor count it, so the atomic is skipped.  A stale cached copy can
// only be larger than the current value, which merely costs the atomic.
__device__ __forceinline__ bool relax_depth(int32_t *levels, uint32_t v, int32_t nd)
{
    if (levels[v] <= nd) return false;
    return atomicMin(&levels[v], nd) > nd;
}

template<typename Off>
__device__ __forceinline__ uint64_t load_off(const Off *o, uint64_t i)
{ return static_cast<uint64_t>(o[i]); }

}  // namespace bw
