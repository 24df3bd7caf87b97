// Shared definitions for the gfx950 radar kernels: error codes, complex
// helpers, launch checks.  Error numbers are the reference's public ErrorCode
// values (ZKX/cusignal_cpp/include/cusignal/runtime/errors.h:18-41).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/zkx_radar.h"

namespace zkx {

__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__host__ __device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// |a|^2 as the reference's power kernel evaluates re*re + im*im under FMA
// contraction: fma(re, re, im*im) (Task1/pipeline/src/steps/step4/step4.cu:18;
// its archived Task1 power map equals this form in every cell).
__device__ __forceinline__ float power_of(float2 a) { return fmaf(a.x, a.x, a.y * a.y); }

namespace buf {
// Raw buffer access: 32-bit byte offsets from a base address.  A lane whose
// offset is kOff (beyond every buffer's size) reads zeros and its stores are
// dropped by the hardware range check, so edge columns need no branches.
constexpr unsigned kOff = 0x80000000u;
constexpr int kBufFlags = 0x00020000;  // raw 32-bit data, gfx9 word 3

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), static_cast<short>(0), static_cast<int>(bytes),
                                             kBufFlags);
}
typedef float f2v __attribute__((ext_vector_type(2)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));
// (whole-vector bit casts: subscripting the builtin's u2v result element by
// element has been narrowed to a single dword load by this compiler)
// Cache policy of an access (the instruction's sc0 / nt / sc1 bits): kCached
// is the default; kSystem (sc0 | sc1) writes through to memory and reads miss
// in both cache levels, so a value one CU stores this way is seen by another
// CU that loads it this way without any cache maintenance.
// kStream (nt) marks data that is read or written once: its lines are not
// kept in L2 or the memory-side cache, which then keep what a later kernel
// reads again (a pipeline stage's output read by the next stage).
constexpr int kCached = 0;
constexpr int kStream = 2;
constexpr int kSystem = 1 | 16;
template <int AUX = kCached>
__device__ __forceinline__ float2 ld2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    const f2v v = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
    return make_float2(v.x, v.y);
}
template <int AUX = kCached>
__device__ __forceinline__ void st2(__amdgpu_buffer_rsrc_t r, float2 v, unsigned voff, unsigned soff)
{
    const f2v w = {v.x, v.y};
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, w), r, voff, soff, AUX);
}
template <int AUX = kCached>
__device__ __forceinline__ void st1(__amdgpu_buffer_rsrc_t r, float v, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, AUX);
}
template <int AUX = kCached>
__device__ __forceinline__ void stb(__amdgpu_buffer_rsrc_t r, unsigned char v, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b8(v, r, voff, soff, AUX);
}
}  // namespace buf

// Map a HIP runtime status onto the public error numbers.
inline int hip_status_to_code(hipError_t e)
{
    if (e == hipSuccess) return ZKX_SUCCESS;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return ZKX_OUT_OF_MEMORY;
    return ZKX_RUNTIME_ERROR;
}

// Thread-local record of the last failure, read by zkx_last_error_message().
void set_last_error(int code, const char* msg);

#define ZKX_LAUNCH_CHECK()                                                        \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) {                                                   \
            ::zkx::set_last_error(::zkx::hip_status_to_code(e_), hipGetErrorString(e_)); \
            return ::zkx::hip_status_to_code(e_);                                 \
        }                                                                         \
    } while (0)

#define ZKX_HIP_CHECK(call)                                                       \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::zkx::set_last_error(::zkx::hip_status_to_code(e_), hipGetErrorString(e_)); \
            return ::zkx::hip_status_to_code(e_);                                 \
        }                                                                         \
    } while (0)

#define ZKX_FAIL(code, msg)                                                       \
    do {                                                                          \
        ::zkx::set_last_error((code), (msg));                                     \
        return (code);                                                            \
    } while (0)

// Copy `count` elements produced by load(e) into LDS dst[e] with U loads in
// flight per thread: all U global loads of a round are issued before any of
// their LDS stores, so the round trip is paid once per U elements.
template <int U, class V, class F>
__device__ __forceinline__ void stage_to_lds(V* dst, int count, F load)
{
    for (int base = threadIdx.x; base < count; base += U * static_cast<int>(blockDim.x)) {
        V tmp[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + u * static_cast<int>(blockDim.x);
            if (e < count) tmp[u] = load(e);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + u * static_cast<int>(blockDim.x);
            if (e < count) dst[e] = tmp[u];
        }
    }
}

__host__ __device__ constexpr int ilog2_c(int n) { return n <= 1 ? 0 : 1 + ilog2_c(n / 2); }

inline bool is_pow2(long long n) { return n > 0 && (n & (n - 1)) == 0; }
inline int ilog2(long long n)
{
    int l = 0;
    while ((1LL << l) < n) ++l;
    return l;
}

}  // namespace zkx
