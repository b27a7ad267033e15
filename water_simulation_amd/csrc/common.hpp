// common.hpp -- shared host-side plumbing of libwsim: status codes, HIP error
// translation and a stream-tagged device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

namespace wsim {

using u32 = uint32_t;
using u64 = uint64_t;
using i32 = int32_t;
using f32 = float;
using f64 = double;

// Status codes returned across the C ABI.  They map one to one onto the
// exception types the reference throws (cuda_check.h:10-16,
// particle_store.cu:12-16, particle_store.h:112-114, loader.cpp:60-74).
enum Status : int {
    kOk = 0,
    kRuntimeError = 1,     // std::runtime_error
    kOutOfRange = 2,       // std::out_of_range
    kInvalidArgument = 3,  // std::invalid_argument
    kDeviceError = 4,      // HIP runtime failure (reference: CUDA error -> runtime_error)
};

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void fail(int status, const std::string& msg) { throw Error(status, msg); }

[[noreturn]] inline void hip_fail(hipError_t e, const char* expr, const char* file, int line) {
    std::string m = "HIP error at ";
    m += file;
    m += ":";
    m += std::to_string(line);
    m += "\n  call: ";
    m += expr;
    m += "\n  err:  ";
    m += hipGetErrorString(e);
    throw Error(kDeviceError, m);
}

#define WSIM_HIP(expr)                                                   \
    do {                                                                 \
        hipError_t wsim_e_ = (expr);                                     \
        if (wsim_e_ != hipSuccess) ::wsim::hip_fail(wsim_e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define WSIM_LAUNCH_CHECK() WSIM_HIP(hipGetLastError())

// Move-only device allocation.  Allocation happens only at construction or
// reserve() time, never per substep, so plain hipMalloc/hipFree are used.
template <class T> class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t count) : size_(count) {
        if (count) WSIM_HIP(hipMalloc(reinterpret_cast<void**>(&ptr_), count * sizeof(T)));
    }
    ~DeviceBuffer() {
        if (ptr_) (void)hipFree(ptr_);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), size_(o.size_) {
        o.ptr_ = nullptr;
        o.size_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            if (ptr_) (void)hipFree(ptr_);
            ptr_ = o.ptr_;
            size_ = o.size_;
            o.ptr_ = nullptr;
            o.size_ = 0;
        }
        return *this;
    }
    T* data() noexcept { return ptr_; }
    const T* data() const noexcept { return ptr_; }
    size_t size() const noexcept { return size_; }
    void zero(hipStream_t s) {
        if (ptr_ && size_) WSIM_HIP(hipMemsetAsync(ptr_, 0, size_ * sizeof(T), s));
    }
    // Grow (discarding contents) to hold at least `count` elements.
    void ensure(size_t count) {
        if (count > size_) *this = DeviceBuffer<T>(count);
    }

private:
    T* ptr_ = nullptr;
    size_t size_ = 0;
};

// Pinned host allocation for the few scalars read back each substep.
template <class T> class PinnedBuffer {
public:
    PinnedBuffer() = default;
    explicit PinnedBuffer(size_t count, unsigned flags = 0) : size_(count) {
        if (count)
            WSIM_HIP(hipHostMalloc(reinterpret_cast<void**>(&ptr_), count * sizeof(T), flags));
    }
    ~PinnedBuffer() {
        if (ptr_) (void)hipHostFree(ptr_);
    }
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    PinnedBuffer(PinnedBuffer&& o) noexcept : ptr_(o.ptr_), size_(o.size_) {
        o.ptr_ = nullptr;
        o.size_ = 0;
    }
    PinnedBuffer& operator=(PinnedBuffer&& o) noexcept {
        if (this != &o) {
            if (ptr_) (void)hipHostFree(ptr_);
            ptr_ = o.ptr_;
            size_ = o.size_;
            o.ptr_ = nullptr;
            o.size_ = 0;
        }
        return *this;
    }
    T* data() noexcept { return ptr_; }
    T& operator[](size_t i) { return ptr_[i]; }
    size_t size() const noexcept { return size_; }

private:
    T* ptr_ = nullptr;
    size_t size_ = 0;
};

inline u32 div_up(u64 a, u64 b) { return static_cast<u32>((a + b - 1) / b); }

}  // namespace wsim
