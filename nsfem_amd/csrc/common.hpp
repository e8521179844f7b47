// Shared runtime pieces of the HIP library: error model, the library stream,
// RAII device arrays.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace nsfem
{
    // Error classes map 1:1 onto C-ABI status codes.  std::invalid_argument
    // and std::runtime_error keep the reference's exception split.
    struct DeviceError : std::runtime_error
    {
        using std::runtime_error::runtime_error;
    };
    struct NonFiniteError : std::runtime_error
    {
        using std::runtime_error::runtime_error;
    };

    inline void hipCheck(hipError_t rc, const char * what, const char * file, int line)
    {
        if (rc != hipSuccess)
        {
            throw DeviceError(std::string(what) + " failed: " + hipGetErrorName(rc) + " (" + hipGetErrorString(rc) +
                              ") at " + file + ":" + std::to_string(line));
        }
    }
#define NSFEM_HIP(call) ::nsfem::hipCheck((call), #call, __FILE__, __LINE__)

    // One non-blocking stream per process for everything the library
    // enqueues; entry points that the reference runs synchronously wait on it
    // before returning.
    hipStream_t libraryStream();
    void syncLibraryStream();

    template <typename T>
    class DeviceArray
    {
        T * ptr_ = nullptr;
        size_t n_ = 0;

    public:
        DeviceArray() = default;
        explicit DeviceArray(size_t n) { resize(n); }
        DeviceArray(const DeviceArray &) = delete;
        DeviceArray & operator=(const DeviceArray &) = delete;
        DeviceArray(DeviceArray && o) noexcept : ptr_(o.ptr_), n_(o.n_)
        {
            o.ptr_ = nullptr;
            o.n_ = 0;
        }
        DeviceArray & operator=(DeviceArray && o) noexcept
        {
            if (this != &o)
            {
                release();
                ptr_ = o.ptr_;
                n_ = o.n_;
                o.ptr_ = nullptr;
                o.n_ = 0;
            }
            return *this;
        }
        ~DeviceArray() { release(); }

        void release() noexcept
        {
            if (ptr_)
            {
                (void)hipFree(ptr_);
            }
            ptr_ = nullptr;
            n_ = 0;
        }
        void resize(size_t n)
        {
            release();
            if (n > 0)
            {
                NSFEM_HIP(hipMalloc(reinterpret_cast<void **>(&ptr_), n * sizeof(T)));
                n_ = n;
            }
        }
        void upload(const T * src, size_t n)
        {
            resize(n);
            if (n)
            {
                NSFEM_HIP(hipMemcpyAsync(ptr_, src, n * sizeof(T), hipMemcpyHostToDevice, libraryStream()));
                syncLibraryStream();
            }
        }
        void upload(const std::vector<T> & v) { upload(v.data(), v.size()); }
        std::vector<T> download() const
        {
            std::vector<T> h(n_);
            if (n_)
            {
                NSFEM_HIP(hipMemcpyAsync(h.data(), ptr_, n_ * sizeof(T), hipMemcpyDeviceToHost, libraryStream()));
                syncLibraryStream();
            }
            return h;
        }
        void zero()
        {
            if (n_)
            {
                NSFEM_HIP(hipMemsetAsync(ptr_, 0, n_ * sizeof(T), libraryStream()));
            }
        }
        T * get() { return ptr_; }
        const T * get() const { return ptr_; }
        size_t size() const { return n_; }
    };

    inline int ceilDiv(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }
} // namespace nsfem
