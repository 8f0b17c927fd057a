// HIP runtime helpers shared by the host driver and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cupf {

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

#define CUPF_HIP(call) ::cupf::hip_check((call), #call)

// Owning device allocation.  Moves, never copies.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { allocate(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) { o.ptr_ = nullptr; o.bytes_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = o.ptr_; bytes_ = o.bytes_;
            o.ptr_ = nullptr; o.bytes_ = 0;
        }
        return *this;
    }
    void allocate(size_t bytes) {
        release();
        if (bytes == 0) return;
        CUPF_HIP(hipMalloc(&ptr_, bytes));
        bytes_ = bytes;
    }
    // Reallocate only when the current buffer is too small.
    void reserve(size_t bytes) { if (bytes > bytes_) allocate(bytes); }
    void release() {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    void* get() const { return ptr_; }
    size_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

template <typename T>
DeviceBuffer upload_vector(const std::vector<T>& v, hipStream_t s) {
    DeviceBuffer d(v.size() * sizeof(T));
    if (!v.empty()) CUPF_HIP(hipMemcpyAsync(d.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
}

}  // namespace cupf
