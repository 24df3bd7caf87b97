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
// in1 (in2 is a wave-uniform broadcast); each chain keeps the reference's
// order, so results are bit-identical to it.  Blocks wholly inside in1 read
// the window from LDS; the blocks at the ends read global memory with
// per-term bounds.
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
// Each chain keeps the reference's j order.  Threads whose R windows lie
// inside in1 run without bounds tests; the others select per term, which
// leaves a chain exactly as skipping the term would.
constexpr int kR = 8;

template <class T>
__global__ void __launch_bounds__(256)
correlate_direct_kernel(const T* __restrict__ a, long long l1, const T* __restrict__ b, long long l2,
                        long long start, long long n, T* __restrict__ out)
{
    const long long k0 = (blockIdx.x * 256LL + threadIdx.x) * kR;
    if (k0 >= n) return;
    const long long b0 = start + k0 - (l2 - 1);
    Acc<T> acc[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) acc[r] = Acc<T>(0);
    const bool inside = b0 >= 0 && b0 + (kR - 1) + (l2 - 1) < l1;
    if (inside) {
        // 32-bit offsets from b0: every index below stays under l1 < 2^31
        const T* ab = a + b0;
        const int L2 = static_cast<int>(l2);
        const int last = static_cast<int>(l1 - 1 - b0);  // largest offset inside in1
        Acc<T> w[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) w[r] = load_acc(ab[r]);
        int j = 0;
        // kR taps per step: window w = in1[b0 + j .. + kR), nw the next kR
        // values; term q of chain r is in1[b0 + r + j + q] (compile-time slot)
        for (; j + kR <= L2; j += kR) {
            Acc<T> c[kR], nw[kR];
#pragma unroll
            for (int q = 0; q < kR; ++q) c[q] = load_acc(b[j + q]);
#pragma unroll
            for (int q = 0; q < kR; ++q) nw[q] = load_acc(ab[min(j + kR + q, last)]);
#pragma unroll
            for (int q = 0; q < kR; ++q)
#pragma unroll
                for (int r = 0; r < kR; ++r) acc[r] = mac<T>(r + q < kR ? w[r + q] : nw[r + q - kR], c[q], acc[r]);
#pragma unroll
            for (int q = 0; q < kR; ++q) w[q] = nw[q];
        }
        for (; j < L2; ++j) {
            const Acc<T> c = load_acc(b[j]);
#pragma unroll
            for (int r = 0; r < kR; ++r) acc[r] = mac<T>(w[r], c, acc[r]);
#pragma unroll
            for (int r = 0; r + 1 < kR; ++r) w[r] = w[r + 1];
            w[kR - 1] = load_acc(ab[min(j + kR, last)]);
        }
    } else {
        for (long long j = 0; j < l2; ++j) {
            const Acc<T> c = load_acc(b[j]);
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                const long long idx = b0 + r + j;
                const bool ok = idx >= 0 && idx < l1;
                const Acc<T> x = load_acc(a[ok ? idx : 0]);
                const Acc<T> y = mac<T>(x, c, acc[r]);
                acc[r] = ok ? y : acc[r];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kR; ++r)
        if (k0 + r < n) out[k0 + r] = store_acc<T>(acc[r]);
}

// Blocks of 256 threads x kR outputs whose every term lies inside in1 (all
// but the blocks at the ends): the block's in1 span and a chunk of in2 are
// staged in LDS by coalesced loads; thread t reads its sliding window at
// 8t + j (padded by one word per 32, so the 32 lanes an LDS read serves at a
// time hit distinct banks), one LDS read per tap for its kR chains.  Same
// per-chain order as above.
constexpr int kTB = 256, kOB = kTB * kR, kJC = 256;
__device__ __forceinline__ int pad_idx(int i) { return i + (i >> 5); }

template <class T>
__global__ void __launch_bounds__(kTB)
correlate_tile_kernel(const T* __restrict__ a, long long l1, const T* __restrict__ b, long long l2, long long start,
                      long long n, T* __restrict__ out)
{
    __shared__ Acc<T> xs[kOB + kJC + (kOB + kJC) / 32 + 8];
    __shared__ Acc<T> cs[kJC];
    const long long kb = static_cast<long long>(blockIdx.x) * kOB;
    const long long bb = start + kb - (l2 - 1);  // in1 offset of the block's first window
    const int t = threadIdx.x;
    Acc<T> acc[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) acc[r] = Acc<T>(0);
    const int L2 = static_cast<int>(l2);
    for (int jc = 0; jc < L2; jc += kJC) {
        const int jn = min(kJC, L2 - jc);
        const int span = kOB - 1 + jn;  // in1[bb + jc + i], i < span (all inside in1)
        __syncthreads();                // the previous chunk's readers are done
        for (int i = t; i < span; i += kTB) xs[pad_idx(i)] = load_acc(a[bb + jc + i]);
        for (int i = t; i < jn; i += kTB) cs[i] = load_acc(b[jc + i]);
        __syncthreads();
        const int w0 = kR * t;
        Acc<T> w[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) w[r] = xs[pad_idx(w0 + r)];
        int j = 0;
        for (; j + kR <= jn; j += kR) {
            Acc<T> nw[kR];
#pragma unroll
            for (int q = 0; q < kR; ++q) nw[q] = xs[pad_idx(min(w0 + kR + j + q, span - 1))];
#pragma unroll
            for (int q = 0; q < kR; ++q) {
                const Acc<T> c = cs[j + q];
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
            w[kR - 1] = xs[pad_idx(min(w0 + kR + j, span - 1))];
        }
    }
    const long long k0 = kb + static_cast<long long>(kR) * t;
#pragma unroll
    for (int r = 0; r < kR; ++r)
        if (k0 + r < n) out[k0 + r] = store_acc<T>(acc[r]);
}

template <class T>
int launch(const void* a, long long l1, const void* b, long long l2, long long start, long long n, void* out,
           hipStream_t st)
{
    const T* A = static_cast<const T*>(a);
    const T* B = static_cast<const T*>(b);
    T* O = static_cast<T*>(out);
    // blocks [blo, bhi) have every window inside in1: bb >= 0 and
    // bb + kOB - 1 + l2 - 1 < l1 for bb = start + blk*kOB - (l2 - 1)
    long long blo = 0;
    const long long need_lo = (l2 - 1) - start;  // blk*kOB >= need_lo
    if (need_lo > 0) blo = (need_lo + kOB - 1) / kOB;
    long long bhi = -1;
    const long long lim = l1 - kOB - start;  // blk*kOB <= lim (and the block must be whole: (blk+1)*kOB <= n)
    if (lim >= 0) bhi = std::min(lim / kOB + 1, n / kOB);
    if (bhi > blo) {
        hipLaunchKernelGGL((correlate_tile_kernel<T>), dim3(static_cast<unsigned>(bhi - blo)), dim3(kTB), 0, st,
                           A, l1, B, l2, start + blo * kOB, n - blo * kOB, O + blo * kOB);
        ZKX_LAUNCH_CHECK();
    } else {
        blo = bhi = 0;
    }
    // the remaining outputs [0, blo*kOB) and [bhi*kOB, n) by the per-thread kernel
    auto rest = [&](long long k_begin, long long k_end) -> int {
        const long long cnt = k_end - k_begin;
        if (cnt <= 0) return ZKX_SUCCESS;
        const long long threads = (cnt + kR - 1) / kR;
        hipLaunchKernelGGL((correlate_direct_kernel<T>), dim3(static_cast<unsigned>((threads + 255) / 256)), dim3(256),
                           0, st, A, l1, B, l2, start + k_begin, cnt, O + k_begin);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    };
    int rc = rest(0, blo * kOB);
    if (rc) return rc;
    return rest(bhi * kOB, n);
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
