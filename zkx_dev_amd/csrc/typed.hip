// Element types of the reference's other instantiations (__half, int32,
// int16, int8; ZKX/cusignal_cpp/src/operators/radartools/radartools_typed.cpp:810-811,
// src/operators/filtering/filtering_typed.cu:683+): the reference loads every
// input sample to float (operator_compute_policy.h:42-45,64-67) and runs its
// FP32 kernels, so the typed operators here are a conversion pass in front of
// the float kernels plus the few rules that depend on the element type:
//   - integer CA-CFAR detections compare the integer input with the floored
//     threshold (radartools_kernels.cuh:23-37);
//   - resample_poly scales custom coefficients by up in the coefficients' own
//     type, integers wrapping (filtering_kernels.cuh:460-497);
//   - cfar_alpha_device computes alpha in FP32 on the device
//     (radartools_kernels.cuh:136-148) before the FP64 storage.
#include <hip/hip_fp16.h>

#include <cmath>
#include <limits>
#include <type_traits>

#include "runtime.hpp"

namespace zkx {
namespace {

template <class T>
__device__ __forceinline__ float load_f32(T v)
{
    if constexpr (std::is_same_v<T, __half>)
        return __half2float(v);
    else
        return static_cast<float>(v);
}

template <class T>
__global__ void to_f32_kernel(const T* __restrict__ in, size_t n, float* __restrict__ out)
{
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[i] = load_f32(in[i]);
}

// det = x > floor(T) in the input's integer type, with T outside the type's
// range, infinite or NaN decided without the conversion.
template <class T>
__global__ void cfar_detect_int_kernel(const T* __restrict__ x, const float* __restrict__ thr, size_t n,
                                       uint8_t* __restrict__ det)
{
    const float low = static_cast<float>(std::numeric_limits<T>::lowest());
    const float high = static_cast<float>(std::numeric_limits<T>::max());
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const float t = thr[i];
        bool d;
        if (isnan(t))
            d = false;
        else if (isinf(t))
            d = t < 0.0f;
        else if (t < low)
            d = true;
        else if (t >= high)
            d = false;
        else
            d = x[i] > static_cast<T>(floorf(t));
        det[i] = d ? 1 : 0;
    }
}

// value * up in the signed integer type W (two's-complement wrap-around),
// as a float
template <class W, class T>
__device__ __forceinline__ float wrapped_product(T value, int up)
{
    using U = std::make_unsigned_t<W>;
    const U a = static_cast<U>(static_cast<W>(value));
    const U p = static_cast<U>(static_cast<unsigned long long>(a) * static_cast<unsigned long long>(up));
    constexpr unsigned long long sign = 1ULL << (8 * sizeof(W) - 1);
    const long long v = static_cast<unsigned long long>(p) >= sign
                            ? static_cast<long long>(p) - static_cast<long long>(2 * sign)
                            : static_cast<long long>(p);
    return static_cast<float>(v);
}

template <class T>
__device__ __forceinline__ float scaled_tap(T value, int up)
{
    if constexpr (std::is_same_v<T, float>) {
        return value * static_cast<float>(up);
    } else if constexpr (std::is_same_v<T, __half>) {
        const float v = __half2float(value) * static_cast<float>(up);
        return up <= 255 ? __half2float(__float2half_rn(v)) : v;
    } else if constexpr (std::is_same_v<T, int32_t>) {
        return wrapped_product<int32_t>(value, up);
    } else if constexpr (std::is_same_v<T, int16_t>) {
        return up < 32768 ? wrapped_product<int16_t>(value, up) : wrapped_product<int32_t>(value, up);
    } else {
        if (up <= 127) return wrapped_product<int8_t>(value, up);
        return up < 32768 ? wrapped_product<int16_t>(value, up) : wrapped_product<int32_t>(value, up);
    }
}

template <class T>
__global__ void scale_taps_kernel(const T* __restrict__ c, int n, int up, float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = scaled_tap(c[i], up);
}

__global__ void cfar_alpha_f32_kernel(float pfa, int reference, double* out)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const float n = static_cast<float>(reference);
        const float a = n * (powf(pfa, -1.0f / n) - 1.0f);
        out[0] = static_cast<double>(a);
    }
}

unsigned grid1(size_t n)
{
    const size_t b = (n + 255) / 256;
    return static_cast<unsigned>(b < 8192 ? (b ? b : 1) : 8192);
}

template <class F>
int by_dtype(int dtype, F&& f)
{
    switch (dtype) {
    case ZKX_DTYPE_F32: return f(float{});
    case ZKX_DTYPE_F16: return f(__half{});
    case ZKX_DTYPE_I32: return f(int32_t{});
    case ZKX_DTYPE_I16: return f(int16_t{});
    case ZKX_DTYPE_I8: return f(int8_t{});
    default: ZKX_FAIL(ZKX_INVALID_DTYPE, "unsupported element type code");
    }
}

}  // namespace
}  // namespace zkx

using namespace zkx;

extern "C" {

int zkx_convert_to_f32(const void* in, int dtype, size_t count, float* out, void* stream)
{
    clear_last_error();
    if (count == 0) return ZKX_SUCCESS;
    if (!in || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "convert: null pointer");
    return by_dtype(dtype, [&](auto tag) {
        using T = decltype(tag);
        hipLaunchKernelGGL(to_f32_kernel<T>, dim3(grid1(count)), dim3(256), 0, (hipStream_t)stream,
                           static_cast<const T*>(in), count, out);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    });
}

int zkx_cfar_detect_typed(const void* x, int dtype, const float* threshold, size_t count, uint8_t* detections,
                          void* stream)
{
    clear_last_error();
    if (count == 0) return ZKX_SUCCESS;
    if (!x || !threshold || !detections) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "cfar detect: null pointer");
    if (dtype != ZKX_DTYPE_I32 && dtype != ZKX_DTYPE_I16 && dtype != ZKX_DTYPE_I8)
        ZKX_FAIL(ZKX_INVALID_DTYPE, "cfar detect: integer element type expected");
    return by_dtype(dtype, [&](auto tag) {
        using T = decltype(tag);
        if constexpr (std::is_integral_v<T>) {
            hipLaunchKernelGGL(cfar_detect_int_kernel<T>, dim3(grid1(count)), dim3(256), 0, (hipStream_t)stream,
                               static_cast<const T*>(x), threshold, count, detections);
            ZKX_LAUNCH_CHECK();
        }
        return ZKX_SUCCESS;
    });
}

int zkx_resample_poly_scale_taps(const void* coefficients, int dtype, int count, int up, float* taps, void* stream)
{
    clear_last_error();
    if (count < 1 || up < 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "resample_poly coefficient shape");
    if (!coefficients || !taps) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "resample_poly taps: null pointer");
    return by_dtype(dtype, [&](auto tag) {
        using T = decltype(tag);
        hipLaunchKernelGGL(scale_taps_kernel<T>, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                           static_cast<const T*>(coefficients), count, up, taps);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    });
}

int zkx_cfar_alpha_f32_device(float pfa, int reference, double* alpha, void* stream)
{
    clear_last_error();
    if (!alpha) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "cfar_alpha: null output");
    if (reference < 1 || !std::isfinite(pfa) || pfa <= 0.0f || pfa > 1.0f)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "cfar_alpha requires reference >= 1 and pfa in (0,1]");
    hipLaunchKernelGGL(cfar_alpha_f32_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, pfa, reference, alpha);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // extern "C"
