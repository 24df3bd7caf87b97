// Library-level runtime: last-error record, twiddle tables, per-thread
// workspaces, dtype widening.  Counterpart of the reference's runtime layer
// (ZKX/cusignal_cpp/include/cusignal/runtime/: errors.h, cuda_error.h,
// host_output_finalize.h) reduced to what the radar path needs.
#include "runtime.hpp"

#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace zkx {

namespace {
thread_local std::string g_last_error;
std::mutex g_tw_mutex;
std::map<std::pair<int, int>, float2*> g_twiddles;  // (device, n) -> table
}  // namespace

void set_last_error(int code, const char* msg)
{
    g_last_error = std::string("[") + std::to_string(code) + "] " + (msg ? msg : "");
}

void clear_last_error() { g_last_error.clear(); }

int twiddles(int n, const float2** out)
{
    int dev = 0;
    ZKX_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    auto it = g_twiddles.find({dev, n});
    if (it != g_twiddles.end()) {
        *out = it->second;
        return ZKX_SUCCESS;
    }
    // W_n^k = exp(-2 pi i k / n), computed in FP64 and rounded once.
    std::vector<float2> host(static_cast<size_t>(n));
    const double two_pi = 6.283185307179586476925286766559;
    for (int k = 0; k < n; ++k) {
        const double a = -two_pi * static_cast<double>(k) / static_cast<double>(n);
        host[k] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
    float2* dptr = nullptr;
    ZKX_HIP_CHECK(hipMalloc(&dptr, sizeof(float2) * static_cast<size_t>(n)));
    ZKX_HIP_CHECK(hipMemcpy(dptr, host.data(), sizeof(float2) * host.size(), hipMemcpyHostToDevice));
    g_twiddles[{dev, n}] = dptr;
    *out = dptr;
    return ZKX_SUCCESS;
}

namespace {
bool capturing(hipStream_t st)
{
    hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &status) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return status != hipStreamCaptureStatusNone;
}
}  // namespace

int ThreadWorkspace::acquire(size_t bytes, hipStream_t st, void** out)
{
    int dev = 0;
    ZKX_HIP_CHECK(hipGetDevice(&dev));
    if (dev != device_) {
        // another device: its event and buffer belong to the old one
        pending_ = false;
        if (ev_) (void)hipEventDestroy(ev_);
        ev_ = nullptr;
        ptr_ = nullptr;
        bytes_ = 0;
    }
    if (bytes > bytes_) {
        if (ptr_ != nullptr) {
            // the previous buffer may still be read by queued work on this device
            ZKX_HIP_CHECK(hipDeviceSynchronize());
            ZKX_HIP_CHECK(hipFree(ptr_));
            pending_ = false;
        }
        ptr_ = nullptr;
        bytes_ = 0;
        ZKX_HIP_CHECK(hipMalloc(&ptr_, bytes));
        bytes_ = bytes;
    }
    device_ = dev;
    if (pending_ && !capturing(st)) ZKX_HIP_CHECK(hipStreamWaitEvent(st, ev_, 0));
    *out = ptr_;
    return ZKX_SUCCESS;
}

void ThreadWorkspace::release(hipStream_t st)
{
    if (capturing(st)) return;
    if (!ev_ && hipEventCreateWithFlags(&ev_, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        ev_ = nullptr;
        return;
    }
    pending_ = hipEventRecord(ev_, st) == hipSuccess;
    if (!pending_) (void)hipGetLastError();
}

namespace {
__global__ void widen_c64_kernel(const float2* __restrict__ in, double2* __restrict__ out, size_t n)
{
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const float2 v = in[i];
        out[i] = make_double2(v.x, v.y);
    }
}
__global__ void widen_f32_kernel(const float* __restrict__ in, double* __restrict__ out, size_t n)
{
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[i] = in[i];
}
unsigned grid_for(size_t n)
{
    const size_t blocks = (n + 255) / 256;
    return static_cast<unsigned>(blocks < 8192 ? (blocks ? blocks : 1) : 8192);
}
}  // namespace

}  // namespace zkx

using namespace zkx;

extern "C" {

const char* zkx_last_error_message(void) { return g_last_error.c_str(); }

int zkx_version(void) { return 100; }

int zkx_widen_c64_to_c128(const zkx_complex64* in, zkx_complex128* out, size_t count, void* stream)
{
    clear_last_error();
    if (count == 0) return ZKX_SUCCESS;
    if (!in || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "widen: null pointer");
    hipLaunchKernelGGL(widen_c64_kernel, dim3(grid_for(count)), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const float2*>(in), reinterpret_cast<double2*>(out), count);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

int zkx_widen_f32_to_f64(const float* in, double* out, size_t count, void* stream)
{
    clear_last_error();
    if (count == 0) return ZKX_SUCCESS;
    if (!in || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "widen: null pointer");
    hipLaunchKernelGGL(widen_f32_kernel, dim3(grid_for(count)), dim3(256), 0, (hipStream_t)stream,
                       in, out, count);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // extern "C"
