// The binned PageRank plan built on the device (pagerank_plan.hip): the same
// arrays, bit for bit, as the host builder of pagerank_bins_plan.hpp, from a
// reverse CSR already in device memory, with sorts and scans in place of
// the host's sequential passes.  Only the per-partition tables (K <= 2048
// entries) and the chunk/task records are assembled on the host.
#pragma once

#include <cstdint>
#include <string>

#include <hip/hip_runtime.h>

namespace bw {
namespace bins {

// Device arrays of a plan (owned by the caller after a successful build).
struct DeviceBins {
    uint16_t *codes = nullptr;
    uint4 *tasks = nullptr, *chunks = nullptr, *blocks = nullptr;
    uint32_t *chunk_base = nullptr, *run_flat = nullptr, *run_bin = nullptr;
    uint16_t *psrc = nullptr;
    uint32_t K = 0, pairs = 0, slots = 0, nchunks = 0, nblocks = 0, ntasks = 0, nruns = 0;
    uint64_t ncodes = 0;
};

// Builds the plan of the graph whose reverse CSR is (roff, rcols) on the
// device (uint32 offsets).  Returns a BW_HIP_* status; on an error every
// device allocation it made is released and `err` says why.
int build_plan_device(const uint32_t *roff, const uint32_t *rcols, uint32_t V, uint32_t E, uint32_t rows_log2,
                      uint32_t chunk_tasks, hipStream_t stream, DeviceBins &out, std::string &err);

// Frees what build_plan_device allocated.
void release(DeviceBins &b);

}  // namespace bins
}  // namespace bw
