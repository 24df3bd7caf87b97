// Process-wide accounting of the device memory the operator API's own
// objects hold (DeviceArray, FFTInterface buffers), with the reference's
// entry points (ZKX/cusignal_cpp/include/cusignal/runtime/
// device_allocation_tracker.h:14-101): a tracked allocate / free pair and the
// live and peak byte counts since the last peak reset.  Memory the library
// allocates internally (plans, workspaces) is not counted, as in the
// reference.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <mutex>
#include <unordered_map>

namespace cusignal::cuda_utils {

namespace detail {

class DeviceAllocationLedger {
public:
    void add(void* p, std::size_t bytes)
    {
        std::lock_guard<std::mutex> g(m_);
        live_[p] = bytes;
        current_ += bytes;
        if (current_ > peak_) peak_ = current_;
    }
    void remove(void* p)
    {
        std::lock_guard<std::mutex> g(m_);
        auto it = live_.find(p);
        if (it == live_.end()) return;
        current_ -= it->second;
        live_.erase(it);
    }
    void reset_peak()
    {
        std::lock_guard<std::mutex> g(m_);
        peak_ = current_;
    }
    std::size_t current()
    {
        std::lock_guard<std::mutex> g(m_);
        return current_;
    }
    std::size_t peak()
    {
        std::lock_guard<std::mutex> g(m_);
        return peak_;
    }

private:
    std::mutex m_;
    std::unordered_map<void*, std::size_t> live_;
    std::size_t current_ = 0;
    std::size_t peak_ = 0;
};

inline DeviceAllocationLedger& ledger()
{
    static DeviceAllocationLedger l;
    return l;
}

}  // namespace detail

// hipMalloc that records the block; a failed allocation records nothing.
inline hipError_t tracked_device_malloc(void** p, std::size_t bytes)
{
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess && p && *p) detail::ledger().add(*p, bytes);
    return e;
}

// hipFree that forgets the block once the free succeeded.
inline hipError_t tracked_device_free(void* p)
{
    const hipError_t e = hipFree(p);
    if (e == hipSuccess && p) detail::ledger().remove(p);
    return e;
}

// the reference's spellings of the same pair
template <typename T>
inline hipError_t tracked_cuda_malloc(T** p, std::size_t bytes)
{
    return tracked_device_malloc(reinterpret_cast<void**>(p), bytes);
}
inline hipError_t tracked_cuda_free(void* p) { return tracked_device_free(p); }

inline void reset_tracked_device_peak() { detail::ledger().reset_peak(); }
inline std::size_t tracked_device_current_bytes() { return detail::ledger().current(); }
inline std::size_t tracked_device_peak_bytes() { return detail::ledger().peak(); }

}  // namespace cusignal::cuda_utils
