// Runtime services shared by the radar translation units.
#pragma once

#include "zkx_hip_common.hpp"

namespace zkx {

void clear_last_error();

// Device twiddle table W_n^k = exp(-2 pi i k/n), k < n, built once per
// (device, n) in FP64 and kept for the life of the process.
int twiddles(int n, const float2** out);

// A per-host-thread device buffer that only grows, the analogue of the
// reference's thread_local CFAR workspace (radartools_typed.cu:13-23).
class ThreadWorkspace {
public:
    int reserve(size_t bytes, void** out);

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
    int device_ = -1;
};

}  // namespace zkx
