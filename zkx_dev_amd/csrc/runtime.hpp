// Runtime services shared by the radar translation units.
#pragma once

#include "zkx_hip_common.hpp"

namespace zkx {

void clear_last_error();

// Device twiddle table W_n^k = exp(-2 pi i k/n), k < n, built once per
// (device, n) in FP64 and kept for the life of the process.
int twiddles(int n, const float2** out);

// A per-host-thread device buffer that only grows, the analogue of the
// reference's thread_local CFAR workspace (radartools_typed.cu:13-23).  The
// reference runs everything on the default stream; here calls take a stream,
// so every use is a lease: acquire() makes the new stream wait for the work
// of the previous lease (an event recorded by release()), whichever stream
// that ran on.  On a stream under graph capture no event is used and the
// caller orders the captured work.
class ThreadWorkspace {
public:
    int acquire(size_t bytes, hipStream_t st, void** out);
    void release(hipStream_t st);

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
    int device_ = -1;
    hipEvent_t ev_ = nullptr;
    bool pending_ = false;
};

// Holds a lease for the scope of the call that enqueues the work.
class WorkspaceLease {
public:
    WorkspaceLease(ThreadWorkspace& ws, hipStream_t st) : ws_(ws), st_(st) {}
    ~WorkspaceLease()
    {
        if (held_) ws_.release(st_);
    }
    int get(size_t bytes, void** out)
    {
        const int rc = ws_.acquire(bytes, st_, out);
        held_ = held_ || rc == 0;
        return rc;
    }
    WorkspaceLease(const WorkspaceLease&) = delete;
    WorkspaceLease& operator=(const WorkspaceLease&) = delete;

private:
    ThreadWorkspace& ws_;
    hipStream_t st_;
    bool held_ = false;
};

}  // namespace zkx
