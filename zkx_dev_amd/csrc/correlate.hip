// 1-D direct cross-correlation for gfx950, in every element type the
// reference instantiates, with its exact arithmetic.
//
// Follows correlate_device(in1, in2, out, mode) and correlate_typed_cpu
// (ZKX/cusignal_cpp/src/operators/convolution/convolution_typed.cpp:134-152,
// 285-334):
//   full = L1 + L2 - 1, out length n by mode (full / same = L1 / valid =
//   |L1 - L2| + 1), s = (full - n)/2 + k:
//   out[k] = sum_{j<L2, 0 <= s-(L2-1-j) < L1} in1[s - (L2-1-j)] * in2[j]
//   float and half: one FP32 fma chain in j order, stored in the input type;
//   int32/int16/int8: products summed modulo 2^bits in the unsigned type and
//   read back as two's complement (no saturation).
// A thread runs the chains of 8 consecutive outputs over one sliding window of
// in1 staged in LDS; each chain keeps the reference's order, so results are
// bit-identical to it.
#include "runtime.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstdint>
#include <limits>
#include <type_traits>

namespace {

template <class T>
__device__ __forceinline__ float load_f(T v) { return static_cast<float>(v); }
template <>
__device__ __forceinline__ float load_f<__half>(__half v) { return __half2float(v); }

template <class T>
__device__ __forceinline__ T store_f(float v) { return static_cast<T>(v); }
template <>
__device__ __forceinline__ __half store_f<__half>(float v) { return __float2half_rn(v); }

template <class T>
using Acc = std::conditional_t<std::is_integral_v<T>, uint32_t, float>;

template <class T>
__device__ __forceinline__ Acc<T> load_acc(T v)
{
    if constexpr (std::is_integral_v<T>) return static_cast<uint32_t>(static_cast<int32_t>(v));
    else return load_f(v);
}

template <class T>
__device__ __forceinline__ Acc<T> mac(Acc<T> x, Acc<T> w, Acc<T> acc)
{
    if constexpr (std::is_integral_v<T>) return acc + x * w;  // modulo 2^32, read back modulo 2^bits
    else return fmaf(x, w, acc);
}

template <class T>
__device__ __forceinline__ T store_acc(Acc<T> acc)
{
    if constexpr (std::is_integral_v<T>) return static_cast<T>(static_cast<std::make_unsigned_t<T>>(acc));
    else return store_f<T>(acc);
}

// R consecutive outputs per thread: output k0 + r takes in1[b0 + r + j] with
// in2[j] for j = 0..L2-1 (b0 = start + k0 - (L2-1)), so one window of in1
// slid along j serves all R chains and every loaded value feeds R products.
// Each chain keeps the reference's j order.
//
// A block of 256 threads covers kOB = 2048 consecutive outputs.  Its in1 span
// and a chunk of in2 are staged in LDS by coalesced loads; thread t reads its
// window at 8t + j (padded by one word per 32, so the 32 lanes an LDS read
// serves at a time hit distinct banks), one LDS read per tap for its R chains.
// Blocks whose every term lies inside in1 (all but the blocks at the ends) run
// without bounds tests; the others select per term, which leaves a chain
// exactly as skipping the term would (a zero product would not: 0 * inf and
// the sign of a zero sum differ).
constexpr int kR = 8;
constexpr int kTB = 256, kOB = kTB * kR, kJC = 256;
__device__ __forceinline__ int pad_idx(int i) { return i + (i >> 5); }

template <class T>
__global__ void __launch_bounds__(kTB)
correlate_tile_kernel(const T* __restrict__ a, long long l1, const T* __restrict__ b, long long l2, long long start,
                      long long n, T* __restrict__ out)
{
    // staged: in1 offsets [0, span); the window reads reach offset
    // kOB - 1 + jn <= kOB + kJC - 1 (words past span are never multiplied)
    __shared__ Acc<T> xs[kOB + kJC + (kOB + kJC) / 32 + 8];
    __shared__ Acc<T> cs[kJC];
    const long long kb = static_cast<long long>(blockIdx.x) * kOB;
    const long long bb = start + kb - (l2 - 1);  // in1 offset of the block's first window
    const bool interior = bb >= 0 && bb + kOB - 1 + (l2 - 1) < l1 && kb + kOB <= n;  // block-uniform
    const int t = threadIdx.x;
    const int w0 = kR * t;
    Acc<T> acc[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) acc[r] = Acc<T>(0);
    const int L2 = static_cast<int>(l2);
    for (int jc = 0; jc < L2; jc += kJC) {
        const int jn = min(kJC, L2 - jc);
        const int span = kOB - 1 + jn;  // in1[bb + jc + i], i < span
        __syncthreads();                // the previous chunk's readers are done
        if (interior) {
            for (int i = t; i < span; i += kTB) xs[pad_idx(i)] = load_acc(a[bb + jc + i]);
        } else {
            for (int i = t; i < span; i += kTB) {
                const long long idx = bb + jc + i;
                xs[pad_idx(i)] = (idx >= 0 && idx < l1) ? load_acc(a[idx]) : Acc<T>(0);
            }
        }
        for (int i = t; i < jn; i += kTB) cs[i] = load_acc(b[jc + i]);
        __syncthreads();
        if (!interior) {
            // edge block: term (r, j) exists iff its in1 index is inside in1
            const long long base = bb + jc + w0;
            for (int j = 0; j < jn; ++j) {
                const Acc<T> c = cs[j];
#pragma unroll
                for (int r = 0; r < kR; ++r) {
                    const long long idx = base + r + j;
                    const Acc<T> y = mac<T>(xs[pad_idx(w0 + r + j)], c, acc[r]);
                    acc[r] = (idx >= 0 && idx < l1) ? y : acc[r];
                }
            }
            continue;
        }
        Acc<T> w[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) w[r] = xs[pad_idx(w0 + r)];
        int j = 0;
        for (; j + kR <= jn; j += kR) {
            // w0 + kR + j is a multiple of 8, so its next 8 words share one
            // padding row: one address per step
            const Acc<T>* nx = xs + pad_idx(w0 + kR + j);
            Acc<T> nw[kR];
#pragma unroll
            for (int q = 0; q < kR; ++q) nw[q] = nx[q];
#pragma unroll
            for (int q = 0; q < kR; ++q) {
                // float taps: a wave-uniform global read, which the compiler
                // makes a scalar load (the FMA takes it from an SGPR)
                Acc<T> c;
                if constexpr (std::is_same_v<T, float>) c = b[jc + j + q];
                else c = cs[j + q];
#pragma unroll
                for (int r = 0; r < kR; ++r) acc[r] = mac<T>(r + q < kR ? w[r + q] : nw[r + q - kR], c, acc[r]);
            }
#pragma unroll
            for (int q = 0; q < kR; ++q) w[q] = nw[q];
        }
        for (; j < jn; ++j) {
            const Acc<T> c = cs[j];
#pragma unroll
            for (int r = 0; r < kR; ++r) acc[r] = mac<T>(w[r], c, acc[r]);
#pragma unroll
            for (int r = 0; r + 1 < kR; ++r) w[r] = w[r + 1];
            w[kR - 1] = xs[pad_idx(w0 + kR + j)];
        }
    }
    const long long k0 = kb + w0;
#pragma unroll
    for (int r = 0; r < kR; ++r)
        if (k0 + r < n) out[k0 + r] = store_acc<T>(acc[r]);
}

template <class T>
int launch(const void* a, long long l1, const void* b, long long l2, long long start, long long n, void* out,
           hipStream_t st)
{
    const long long blocks = (n + kOB - 1) / kOB;  // n < 2^31: at most 2^20 blocks
    hipLaunchKernelGGL((correlate_tile_kernel<T>), dim3(static_cast<unsigned>(blocks)), dim3(kTB), 0, st,
                       static_cast<const T*>(a), l1, static_cast<const T*>(b), l2, start, n, static_cast<T*>(out));
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

size_t elem_size(int dtype)
{
    switch (dtype) {
    case ZKX_DTYPE_F32: return 4;
    case ZKX_DTYPE_F16: return 2;
    case ZKX_DTYPE_I32: return 4;
    case ZKX_DTYPE_I16: return 2;
    case ZKX_DTYPE_I8: return 1;
    default: return 0;
    }
}

bool overlap(const void* p, size_t pb, const void* q, size_t qb)
{
    const auto a = reinterpret_cast<uintptr_t>(p), b = reinterpret_cast<uintptr_t>(q);
    return pb && qb && a < b + qb && b < a + pb;
}

}  // namespace

extern "C" {

long long zkx_correlate_out_len(long long l1, long long l2, int mode)
{
    if (mode < 0 || mode > 2 || l1 < 0 || l2 < 0) return -1;
    if (l1 == 0 || l2 == 0) return 0;
    if (mode == 0) return l1 + l2 - 1;
    if (mode == 1) return l1;
    return l1 >= l2 ? l1 - l2 + 1 : l2 - l1 + 1;
}

int zkx_correlate_direct(const void* in1, long long l1, const void* in2, long long l2, int dtype, int mode,
                         void* out, void* stream)
{
    zkx::clear_last_error();
    if (mode < 0 || mode > 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate mode must be full, same, or valid");
    const size_t es = elem_size(dtype);
    if (!es) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate element type");
    if (l1 < 0 || l2 < 0 || l1 > std::numeric_limits<int>::max() || l2 > std::numeric_limits<int>::max() ||
        l1 + l2 - 1 > std::numeric_limits<int>::max())
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate input exceeds 32-bit indexing");
    const long long n = zkx_correlate_out_len(l1, l2, mode);
    if (n == 0) return ZKX_SUCCESS;
    if (!in1 || !in2 || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate: null pointer");
    if (overlap(out, es * n, in1, es * l1) || overlap(out, es * n, in2, es * l2))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate output must not alias an input");
    const long long start = (l1 + l2 - 1 - n) / 2;
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (dtype) {
    case ZKX_DTYPE_F32: return launch<float>(in1, l1, in2, l2, start, n, out, st);
    case ZKX_DTYPE_F16: return launch<__half>(in1, l1, in2, l2, start, n, out, st);
    case ZKX_DTYPE_I32: return launch<int32_t>(in1, l1, in2, l2, start, n, out, st);
    case ZKX_DTYPE_I16: return launch<int16_t>(in1, l1, in2, l2, start, n, out, st);
    default: return launch<int8_t>(in1, l1, in2, l2, start, n, out, st);
    }
}

}  // extern "C"
