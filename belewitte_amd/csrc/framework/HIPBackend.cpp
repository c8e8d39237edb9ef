#include "HIPBackend.hpp"

#include <algorithm>

void hipAssert(int status, const char *file, int line)
{
    if (status == BW_HIP_OK) return;
    std::cout.flush();
    std::cerr << "HIP error #" << status << " (" << file << ":" << line << "):" << std::endl
              << bw_hip_last_error() << std::endl
              << std::endl
              << "Callstack:" << std::endl;
    dump_stack_trace(status);
}

// ------------------------------------------------------------------ Backend

Backend::~Backend() {}

size_t Backend::platformCount() { return devicesPerPlatform.size(); }

int Backend::deviceCount(size_t platform) { return devicesPerPlatform.at(platform); }

void Backend::listPlatforms(bool verbose)
{
    for (size_t i = 0; i < devicesPerPlatform.size(); i++) queryPlatform(i, verbose);
}

void Backend::listDevices(size_t platform, bool verbose)
{
    for (int i = 0; i < devicesPerPlatform.at(platform); i++) queryDevice(platform, i, verbose);
}

std::pair<size_t, size_t> Backend::computeDivision(size_t count)
{
    // Smallest multiple i of the CU count whose per-block share fits the
    // block limit, block size rounded up to whole 64-wide waves.
    const size_t cus = std::max<size_t>(numComputeUnits, 1);
    const size_t maxT = std::max<size_t>(maxThreadsPerBlock, 64);
    auto ceil_div = [](size_t x, size_t y) { return (x + y - 1) / y; };
    size_t i = 1;
    size_t threads = ceil_div(count, cus);
    for (i = 1; threads > maxT; i++) threads = ceil_div(count, i * cus);
    if (threads % 64) threads += 64 - threads % 64;
    if (threads == 0) threads = 64;
    const size_t blocks = std::max<size_t>(1, std::min(i * cus, ceil_div(count, threads)));
    return {threads, blocks};
}

// --------------------------------------------------------------- allocations

HIPBackend::hip_alloc_t::hip_alloc_t(size_t bytes, bool readonly) : byteSize(bytes), read_only(readonly)
{
    void *h = nullptr, *d = nullptr;
    BW_CHK(bw_hip_host_alloc(&h, bytes));
    hostPtr = std::shared_ptr<void>(h, [](void *p) { bw_hip_host_free(p); });
    std::memset(h, 0, bytes);
    BW_CHK(bw_hip_dev_alloc(&d, bytes));
    devPtr = std::shared_ptr<void>(d, [](void *p) { bw_hip_dev_free(p); });
}

void HIPBackend::hip_alloc_t::registerLocalAlloc(void *field, const hip_alloc_t &sub)
{
    localAllocs.push_back(sub);
    localAllocs.back().associatedPtr = static_cast<void **>(field);
    *static_cast<void **>(field) = sub.host();
}

void HIPBackend::hip_alloc_t::copyHostToDev()
{
    if (!hostPtr) return;
    for (auto &sub : localAllocs) sub.copyHostToDev();
    BW_CHK(bw_hip_memcpy_h2d(devPtr.get(), hostPtr.get(), byteSize));
    if (associatedPtr) *associatedPtr = devPtr.get();
}

void HIPBackend::hip_alloc_t::copyDevToHost()
{
    if (!hostPtr) return;
    // The host copy of a struct keeps its device pointers for the
    // launchers; only plain arrays are read back.
    if (localAllocs.empty()) BW_CHK(bw_hip_memcpy_d2h(hostPtr.get(), devPtr.get(), byteSize));
    for (auto &sub : localAllocs) sub.copyDevToHost();
}

void HIPBackend::hip_alloc_t::free()
{
    localAllocs.clear();
    hostPtr.reset();
    devPtr.reset();
    byteSize = 0;
    associatedPtr = nullptr;
}

// ------------------------------------------------------------------ platform

HIPBackend &HIPBackend::get()
{
    static HIPBackend backend;
    return backend;
}

HIPBackend::HIPBackend()
{
    int32_t n = 0;
    devicesPerPlatform_.push_back(0);
    if (bw_hip_device_count(&n) != BW_HIP_OK || n <= 0) return;
    devicesPerPlatform_[0] = n;
    for (int32_t i = 0; i < n; i++) {
        bw_hip_props p;
        BW_CHK(bw_hip_get_device_props(i, &p));
        props_.push_back(p);
    }
    maxDims_ = 3;
    initialised_ = true;
    setDevice(0, 0);
}

HIPBackend::~HIPBackend() {}

void HIPBackend::queryPlatform(size_t platform, bool verbose)
{
    if (platform >= devicesPerPlatform.size()) {
        std::cerr << "Non-existent platform #" << platform << ", platform count is " << devicesPerPlatform.size()
                  << std::endl;
        std::exit(EXIT_FAILURE);
    }
    std::cout << "Platform #" << platform << ":" << std::endl;
    listDevices(platform, verbose);
}

void HIPBackend::queryDevice(size_t platform, int dev, bool verbose)
{
    if (platform >= devicesPerPlatform.size() || dev < 0 || dev >= devicesPerPlatform[platform]) {
        std::cerr << "Non-existent device #" << dev << std::endl;
        std::exit(EXIT_FAILURE);
    }
    const bw_hip_props &p = props_[static_cast<size_t>(dev)];
    std::cout << "    Device Number: " << dev << std::endl;
    std::cout << "\tDevice Name: " << p.name << std::endl;
    std::cout << "\tArchitecture: " << p.arch << std::endl << std::endl;
    if (!verbose) return;
    std::cout << "\tWave Size: " << p.wave_size << std::endl;
    std::cout << "\tCompute Units: " << p.compute_units << std::endl;
    std::cout << "\tThreads Per Block: " << p.max_threads_per_block << std::endl << std::endl;
    std::cout << "\tMax Block Dimensions: " << p.max_block_dims[0] << " x " << p.max_block_dims[1] << " x "
              << p.max_block_dims[2] << std::endl;
    std::cout << "\tMax Grid Dimensions: " << p.max_grid_dims[0] << " x " << p.max_grid_dims[1] << " x "
              << p.max_grid_dims[2] << std::endl << std::endl;
    std::cout << "\tGlobal Mem: " << p.total_global_mem << std::endl;
    std::cout << "\tL2 Cache: " << p.l2_cache_bytes << std::endl;
    std::cout << "\tLDS Per Block: " << p.shared_mem_per_block << std::endl << std::endl;
}

void HIPBackend::setDevice(size_t platform, int device)
{
    if (platform >= devicesPerPlatform.size() || device < 0 || device >= devicesPerPlatform[platform]) {
        std::cerr << "Non-existent device #" << device << ", device count for platform " << platform << " is "
                  << (platform < devicesPerPlatform.size() ? devicesPerPlatform[platform] : 0) << std::endl;
        std::exit(EXIT_FAILURE);
    }
    BW_CHK(bw_hip_set_device(device));
    const bw_hip_props &p = props_[static_cast<size_t>(device)];
    numComputeUnits_ = static_cast<size_t>(p.compute_units);
    maxThreadsPerBlock_ = static_cast<size_t>(p.max_threads_per_block);
    maxSharedMem_ = static_cast<size_t>(p.shared_mem_per_block);
    maxBlockSizes_ = {static_cast<size_t>(p.max_block_dims[0]), static_cast<size_t>(p.max_block_dims[1]),
                      static_cast<size_t>(p.max_block_dims[2])};
    maxGridSizes_ = {static_cast<size_t>(p.max_grid_dims[0]), static_cast<size_t>(p.max_grid_dims[1]),
                     static_cast<size_t>(p.max_grid_dims[2])};
}

void HIPBackend::setWorkSizes(size_t dims, std::vector<size_t> blockSizes, std::vector<size_t> gridSizes,
                              size_t sharedMem)
{
    if (dims < 1 || dims > 3) {
        std::cerr << "Invalid number of dimensions: " << dims << std::endl;
        std::exit(EXIT_FAILURE);
    }
    if (sharedMem > maxSharedMem) {
        std::cerr << "Insufficient shared memory, " << sharedMem << " request, " << maxSharedMem << " available."
                  << std::endl;
        std::exit(EXIT_FAILURE);
    }
    if (blockSizes.size() != dims || gridSizes.size() != dims) {
        std::cerr << "Number of block/grid sizes doesn't match the number of dimensions (" << dims << ")"
                  << std::endl;
        std::exit(EXIT_FAILURE);
    }
    size_t block = 1, grid = 1;
    for (size_t d = 0; d < dims; d++) {
        block *= blockSizes[d];
        grid *= gridSizes[d];
    }
    launch_.block = static_cast<uint32_t>(block);
    launch_.grid = static_cast<uint32_t>(grid);
}
