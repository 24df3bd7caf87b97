// Owning, move-only device buffer with the reference's DeviceArray interface
// (ZKX/cusignal_cpp/include/cusignal/runtime/device_array.h:19-121) on the HIP
// runtime: from_host / to_host copy synchronously on the default stream,
// reset() frees, size() counts elements; blocks are counted by the device
// allocation tracker.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <vector>

#include "device_allocation_tracker.h"
#include "errors.h"

namespace cusignal {

namespace cuda_utils {
inline void check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    const ErrorCode code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? ErrorCode::OUT_OF_MEMORY
                                                                                        : ErrorCode::CUDA_RUNTIME_ERROR;
    throw_error(code, std::string(what) + ": " + hipGetErrorString(e), {std::to_string(static_cast<int>(e)),
                                                                       hipGetErrorName(e), hipGetErrorString(e)});
}
inline void synchronize_stream() { check_hip(hipStreamSynchronize(nullptr), "hipStreamSynchronize"); }
}  // namespace cuda_utils

template <typename T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(std::size_t count) { allocate(count); }
    ~DeviceArray() noexcept
    {
        if (data_) (void)cuda_utils::tracked_device_free(data_);
    }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : data_(o.data_), size_(o.size_)
    {
        o.data_ = nullptr;
        o.size_ = 0;
    }
    DeviceArray& operator=(DeviceArray&& o) noexcept
    {
        if (this != &o) {
            if (data_) (void)cuda_utils::tracked_device_free(data_);
            data_ = o.data_;
            size_ = o.size_;
            o.data_ = nullptr;
            o.size_ = 0;
        }
        return *this;
    }

    static DeviceArray from_host(const std::vector<T>& host)
    {
        DeviceArray out(host.size());
        if (!host.empty()) {
            cuda_utils::check_hip(hipMemcpy(out.data_, host.data(), sizeof(T) * host.size(), hipMemcpyHostToDevice),
                                  "hipMemcpy H2D");
            cuda_utils::synchronize_stream();
        }
        return out;
    }

    std::vector<T> to_host() const
    {
        std::vector<T> host(size_);
        if (size_) {
            cuda_utils::synchronize_stream();
            cuda_utils::check_hip(hipMemcpy(host.data(), data_, sizeof(T) * size_, hipMemcpyDeviceToHost),
                                  "hipMemcpy D2H");
        }
        return host;
    }

    void reset()
    {
        if (data_) {
            const hipError_t e = cuda_utils::tracked_device_free(data_);
            data_ = nullptr;
            size_ = 0;
            cuda_utils::check_hip(e, "hipFree");
        }
    }
    void reset(std::size_t count) { allocate(count); }

    [[nodiscard]] std::size_t size() const noexcept { return size_; }
    [[nodiscard]] bool empty() const noexcept { return size_ == 0; }
    [[nodiscard]] T* data() noexcept { return data_; }
    [[nodiscard]] const T* data() const noexcept { return data_; }

private:
    void allocate(std::size_t count)
    {
        reset();
        if (count) {
            cuda_utils::check_hip(cuda_utils::tracked_device_malloc(reinterpret_cast<void**>(&data_), sizeof(T) * count),
                                  "hipMalloc");
            size_ = count;
        }
    }
    T* data_ = nullptr;
    std::size_t size_ = 0;
};

}  // namespace cusignal
