// The HIP platform of the plugin framework, on top of the C ABI of
// libbelewitte_hip (include/belewitte_hip.h).
//
// Same Platform concept as CUDABackend (CUDA.hpp of merijn/Belewitte):
// alloc_t<T> pairs a pinned host buffer with a device buffer, nested arrays
// of a struct are registered with registerLocalAlloc and their device
// addresses patched into the host copy on copyHostToDev; setWorkSizes
// stores the launch geometry; runKernel(kernel, args...) launches.
//
// What differs, by design for MI355X: a "kernel" is a host launcher with the
// device kernel's signature.  runKernel hands it, for every alloc_t argument,
// the device pointer -- except for allocations that own nested arrays (the
// graph representation structs), which are passed as their host copy whose
// array members already hold device pointers.  The launcher forwards the
// arrays as plain kernel arguments through the C ABI, so the kernels never
// load a device-side descriptor first.
//
// Errors follow the reference convention: a failing call prints the C ABI
// message and a stack trace and exits with status | 128 (CUDA.cpp:7-26).
#ifndef BW_FRAMEWORK_HIPBACKEND_HPP
#define BW_FRAMEWORK_HIPBACKEND_HPP

#include <cstddef>
#include <cstring>
#include <iostream>
#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

#include "Util.hpp"
#include "belewitte_hip.h"

void hipAssert(int status, const char *file, int line);
#define BW_CHK(call) hipAssert((call), __FILE__, __LINE__)

class Backend
{
  public:
    virtual ~Backend();
    size_t platformCount();
    int deviceCount(size_t platform);
    void listPlatforms(bool verbose);
    void listDevices(size_t platform, bool verbose);
    // Backend.cpp:52-70 geometry with the wave64 rounding.
    std::pair<size_t, size_t> computeDivision(size_t count);

    virtual void queryPlatform(size_t platform, bool verbose) = 0;
    virtual void queryDevice(size_t platform, int device, bool verbose) = 0;
    virtual void setDevice(size_t platform, int device) = 0;
    virtual void setWorkSizes(size_t dim, std::vector<size_t> blockSizes, std::vector<size_t> gridSizes,
                              size_t sharedMem = 0) = 0;

  protected:
    Backend() = default;
    std::vector<int> devicesPerPlatform_;
    size_t numComputeUnits_ = 0, maxThreadsPerBlock_ = 0, maxDims_ = 0, maxSharedMem_ = 0;
    std::vector<size_t> maxBlockSizes_, maxGridSizes_;
    bool initialised_ = false;

  public:
    const std::vector<int> &devicesPerPlatform = devicesPerPlatform_;
    const size_t &numComputeUnits = numComputeUnits_;
    const size_t &maxThreadsPerBlock = maxThreadsPerBlock_;
    const size_t &maxDims = maxDims_;
    const size_t &maxSharedMem = maxSharedMem_;
    const std::vector<size_t> &maxBlockSizes = maxBlockSizes_;
    const std::vector<size_t> &maxGridSizes = maxGridSizes_;
    const bool &initialised = initialised_;
};

template<typename>
struct isBackendAllocTrait : public std::false_type {};

template<typename T>
constexpr bool isDeviceAlloc()
{ return isBackendAllocTrait<typename std::remove_reference<T>::type>::value; }

class HIPBackend : public Backend
{
  public:
    // Untyped pinned-host + device pair.
    class hip_alloc_t
    {
      public:
        hip_alloc_t() = default;
        hip_alloc_t(size_t bytes, bool readonly);

        explicit operator bool() const { return static_cast<bool>(hostPtr); }

        void copyHostToDev();
        void copyDevToHost();
        void free();
        void registerLocalAlloc(void *field, const hip_alloc_t &sub);

        void *host() const { return hostPtr.get(); }
        void *device() const { return devPtr.get(); }
        bool hasLocalAllocs() const { return !localAllocs.empty(); }
        size_t bytes() const { return byteSize; }

      protected:
        std::shared_ptr<void> hostPtr, devPtr;
        size_t byteSize = 0;
        bool read_only = false;
        void **associatedPtr = nullptr;
        std::vector<hip_alloc_t> localAllocs;
    };

    template<typename T>
    class alloc_t : public hip_alloc_t
    {
      public:
        alloc_t() = default;
        alloc_t(size_t n, bool readonly) : hip_alloc_t(sizeof(T) * n, readonly), count(n) {}
        alloc_t(alloc_t &&o) noexcept : hip_alloc_t(std::move(o)), count(o.count) { o.count = 0; }
        alloc_t &operator=(alloc_t &&o) noexcept
        {
            hip_alloc_t::operator=(std::move(o));
            count = o.count;
            o.count = 0;
            return *this;
        }
        alloc_t(const alloc_t &) = delete;
        alloc_t &operator=(const alloc_t &) = delete;

        T *operator->() const { return static_cast<T *>(host()); }
        T &operator*() const { return *static_cast<T *>(host()); }
        T &operator[](size_t i) const
        {
            checkError(i < count, "Attempting to index invalid offset! Index: ", i, " Max: ", count);
            return static_cast<T *>(host())[i];
        }
        T *begin() const { return static_cast<T *>(host()); }
        T *end() const { return static_cast<T *>(host()) + count; }
        T *devicePtr() const { return static_cast<T *>(device()); }

        // What a launcher receives for this allocation (see file comment).
        T *kernelArg() const
        { return hasLocalAllocs() ? static_cast<T *>(host()) : static_cast<T *>(device()); }

        void free()
        {
            hip_alloc_t::free();
            count = 0;
        }

        template<typename U>
        void allocLocal(U **field, size_t n)
        { registerLocalAlloc(static_cast<void *>(field), alloc_t<U>(n, false)); }

        size_t count = 0;
        const size_t &size = count;
    };

    typedef void *kernel_type;

    template<typename T> struct HostToDev { typedef T type; };
    template<typename T> struct HostToDev<alloc_t<T>> { typedef T *type; };
    template<typename T> struct HostToDev<alloc_t<T> &> { typedef T *type; };
    template<typename T> struct HostToDev<alloc_t<T> &&> { typedef T *type; };
    template<typename T> struct HostToDev<const alloc_t<T> &> { typedef T *type; };

    template<typename T> struct DevToHost { typedef T type; };
    template<typename T> struct DevToHost<T *> { typedef alloc_t<T> type; };

    template<typename... Args>
    struct kernel {
        using type = void (*)(typename HostToDev<Args>::type...);
    };

    static HIPBackend &get();

    void queryPlatform(size_t platform, bool verbose) override;
    void queryDevice(size_t platform, int device, bool verbose) override;
    void setDevice(size_t platform, int device) override;
    void setWorkSizes(size_t dims, std::vector<size_t> blockSizes, std::vector<size_t> gridSizes,
                      size_t sharedMem = 0) override;

    template<typename... Args>
    void runKernel(typename kernel<Args...>::type k, const Args &...args)
    { k(arg(args)...); }

    template<typename V>
    alloc_t<V> alloc(size_t count = 1)
    { return alloc_t<V>(count, false); }

    template<typename V>
    alloc_t<V> allocConstant(size_t count = 1)
    { return alloc_t<V>(count, true); }

    // Launch configuration set by the last setWorkSizes.
    const bw_launch &launchConfig() const { return launch_; }

  private:
    HIPBackend();
    ~HIPBackend() override;

    template<typename T>
    static T *arg(const alloc_t<T> &a) { return a.kernelArg(); }
    template<typename T, typename = std::enable_if_t<!isDeviceAlloc<T>()>>
    static const T &arg(const T &v) { return v; }

    bw_launch launch_{};
    std::vector<bw_hip_props> props_;
};

template<typename T>
struct isBackendAllocTrait<HIPBackend::alloc_t<T>> : public std::true_type {};

#endif
