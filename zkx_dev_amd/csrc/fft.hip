// The reference's FFT backend C ABI (fft_interface.h:189-447) on the gfx950
// LDS Stockham FFT, so this library can replace libzkx_fft_thrust.so at link
// time.  Shift, frequency and error-string helpers follow the reference
// backend's definitions (ZKX/cusignal_cpp/src/backends/fft/fft_thrust.cu:520-550,
// 1228-1361), including its fft_shift == fft_inverse_shift for 1-D arrays.
#include "../../include/zkx_fft.h"
#include "runtime.hpp"

#include <new>
#include <vector>

namespace zkx {
int fft_rows_dispatch(int n, const float2* x, long long rows, bool inverse, float scale, float2* y,
                      hipStream_t st);
}

namespace {

struct Plan {
    int nx;
    int batch;
    FFTDirection dir;
};

__global__ void shift1d_kernel(const float2* __restrict__ in, float2* __restrict__ out, int n, int half, int batch)
{
    const long long total = static_cast<long long>(n) * batch;
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long b = g / n;
        const int i = static_cast<int>(g % n);
        out[b * n + i] = in[b * n + (i + half) % n];
    }
}

__global__ void shift2d_kernel(const float2* __restrict__ in, float2* __restrict__ out, int nx, int ny, int sx, int sy)
{
    const long long total = static_cast<long long>(nx) * ny;
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int x = static_cast<int>(g % nx), y = static_cast<int>(g / nx);
        out[static_cast<long long>(y) * nx + x] =
            in[static_cast<long long>((y + sy) % ny) * nx + (x + sx) % nx];
    }
}

unsigned blocks_for(long long n)
{
    const long long b = (n + 255) / 256;
    return static_cast<unsigned>(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

FFTResult run_plan(FFTPlanHandle handle, void* out, const void* in, hipStream_t st)
{
    if (!handle || !out || !in) return FFT_INVALID_VALUE;
    const Plan* p = static_cast<const Plan*>(handle);
    const bool inv = p->dir == FFT_INVERSE;
    const float scale = inv ? 1.0f / static_cast<float>(p->nx) : 1.0f;
    const int rc = zkx::fft_rows_dispatch(p->nx, static_cast<const float2*>(in), p->batch, inv, scale,
                                          static_cast<float2*>(out), st);
    return rc == ZKX_SUCCESS ? FFT_SUCCESS : (rc == ZKX_OUT_OF_MEMORY ? FFT_ALLOC_FAILED : FFT_EXEC_FAILED);
}

// Frequency tables are written on the host; device destinations get a copy.
FFTResult store_freqs(float* dst, const std::vector<float>& v)
{
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, dst) == hipSuccess && attr.type == hipMemoryTypeDevice) {
        if (hipMemcpy(dst, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
            return FFT_EXEC_FAILED;
        return FFT_SUCCESS;
    }
    (void)hipGetLastError();
    for (size_t i = 0; i < v.size(); ++i) dst[i] = v[i];
    return FFT_SUCCESS;
}

}  // namespace

extern "C" {

FFTResult fft_plan_create_1d(FFTPlanHandle* plan, int nx, int batch, FFTType type, FFTDirection direction)
{
    if (!plan) return FFT_INVALID_VALUE;
    *plan = nullptr;
    if (type != FFT_TYPE_C2C || batch < 1 || nx < 2 || nx > 8192 || (nx & (nx - 1)) != 0)
        return FFT_INVALID_VALUE;
    if (direction != FFT_FORWARD && direction != FFT_INVERSE) return FFT_INVALID_VALUE;
    Plan* p = new (std::nothrow) Plan{nx, batch, direction};
    if (!p) return FFT_ALLOC_FAILED;
    // build the twiddle table now so execution never allocates
    const float2* tw = nullptr;
    if (zkx::twiddles(nx, &tw) != ZKX_SUCCESS) {
        delete p;
        return FFT_ALLOC_FAILED;
    }
    *plan = p;
    return FFT_SUCCESS;
}

FFTResult fft_plan_create_2d(FFTPlanHandle* plan, int, int, FFTType, FFTDirection)
{
    if (plan) *plan = nullptr;
    return FFT_INVALID_VALUE;
}

FFTResult fft_execute(FFTPlanHandle plan, void* output, const void* input)
{
    return run_plan(plan, output, input, nullptr);
}

FFTResult fft_execute_on_stream(FFTPlanHandle plan, void* output, const void* input, void* stream)
{
    return run_plan(plan, output, input, static_cast<hipStream_t>(stream));
}

FFTResult fft_execute_inplace(FFTPlanHandle plan, void* data) { return run_plan(plan, data, data, nullptr); }

FFTResult fft_plan_destroy(FFTPlanHandle plan)
{
    if (!plan) return FFT_INVALID_VALUE;
    delete static_cast<Plan*>(plan);
    return FFT_SUCCESS;
}

FFTResult fft_shift(void* output, const void* input, int nx, int batch)
{
    if (!output || !input || nx <= 0 || batch <= 0) return FFT_INVALID_VALUE;
    hipLaunchKernelGGL(shift1d_kernel, dim3(blocks_for(static_cast<long long>(nx) * batch)), dim3(256), 0, nullptr,
                       static_cast<const float2*>(input), static_cast<float2*>(output), nx, nx / 2, batch);
    return hipGetLastError() == hipSuccess ? FFT_SUCCESS : FFT_EXEC_FAILED;
}

FFTResult fft_inverse_shift(void* output, const void* input, int nx, int batch)
{
    return fft_shift(output, input, nx, batch);
}

FFTResult fft_shift_2d(void* output, const void* input, int nx, int ny)
{
    if (!output || !input || nx <= 0 || ny <= 0) return FFT_INVALID_VALUE;
    hipLaunchKernelGGL(shift2d_kernel, dim3(blocks_for(static_cast<long long>(nx) * ny)), dim3(256), 0, nullptr,
                       static_cast<const float2*>(input), static_cast<float2*>(output), nx, ny, nx / 2, ny / 2);
    return hipGetLastError() == hipSuccess ? FFT_SUCCESS : FFT_EXEC_FAILED;
}

FFTResult fft_inverse_shift_2d(void* output, const void* input, int nx, int ny)
{
    if (!output || !input || nx <= 0 || ny <= 0) return FFT_INVALID_VALUE;
    hipLaunchKernelGGL(shift2d_kernel, dim3(blocks_for(static_cast<long long>(nx) * ny)), dim3(256), 0, nullptr,
                       static_cast<const float2*>(input), static_cast<float2*>(output), nx, ny, (nx + 1) / 2,
                       (ny + 1) / 2);
    return hipGetLastError() == hipSuccess ? FFT_SUCCESS : FFT_EXEC_FAILED;
}

const char* fft_error_string(FFTResult result)
{
    switch (result) {
    case FFT_SUCCESS: return "Success";
    case FFT_INVALID_VALUE: return "Invalid value";
    case FFT_ALLOC_FAILED: return "Allocation failed";
    case FFT_EXEC_FAILED: return "Execution failed";
    case FFT_INTERNAL_ERROR: return "Internal error";
    default: return "Unknown error";
    }
}

int fft_version(void) { return 1; }

FFTResult fft_freq(float* freqs, int n, float d)
{
    if (!freqs || n <= 0) return FFT_INVALID_VALUE;
    std::vector<float> v(static_cast<size_t>(n));
    const float inv_n_d = 1.0f / (static_cast<float>(n) * d);
    const int num_pos = (n + 1) / 2;
    for (int i = 0; i < num_pos; ++i) v[i] = static_cast<float>(i) * inv_n_d;
    for (int i = 0; i < n - num_pos; ++i) v[num_pos + i] = -static_cast<float>(n - num_pos - i) * inv_n_d;
    return store_freqs(freqs, v);
}

FFTResult fft_rfftfreq(float* freqs, int n, float d)
{
    if (!freqs || n <= 0) return FFT_INVALID_VALUE;
    const int len = n / 2 + 1;
    std::vector<float> v(static_cast<size_t>(len));
    const float inv_n_d = 1.0f / (static_cast<float>(n) * d);
    for (int i = 0; i < len; ++i) v[i] = static_cast<float>(i) * inv_n_d;
    return store_freqs(freqs, v);
}

}  // extern "C"
