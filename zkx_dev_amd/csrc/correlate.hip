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
#include "conv_nd.hpp"
#include "fft_any.hpp"
#include "runtime.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
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

// ---------------------------------------------------------------------------
// 2-D direct correlation (convolution_typed.cpp:101-114, 203-236, 416-470):
// output (r, c) sits at (rs + r, cs + c) of the full correlation; its terms
// run over in2 in row-major (kr, kc) order with the sample at in1 row
// map(rs + r - (r2-1-kr)) and column map(cs + c - (c2-1-kc)), or the fill
// value where either maps outside (fill boundary).  Rows and columns map
// independently, so a block stages its whole in1 window with the boundary
// already applied and the inner loop has no index logic.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int map2d(long long index, int size, int boundary)
{
    if (index >= 0 && index < size) return static_cast<int>(index);
    if (boundary == 0) return -1;
    if (boundary == 1) {
        long long w = index % size;
        return static_cast<int>(w < 0 ? w + size : w);
    }
    const long long period = 2LL * size;
    long long w = index % period;
    if (w < 0) w += period;
    return static_cast<int>(w < size ? w : period - 1 - w);
}

struct Corr2d {
    int r1, c1, r2, c2;  // input shapes
    int rows, cols;      // output shape
    int rs, cs;          // output origin in the full correlation
    int boundary;        // 0 fill, 1 wrap, 2 symm
};

// tile[i * pitch + j] = sample at full-correlation row vr0 + i, column
// vc0 + j (boundary applied; fill outside for the fill boundary), for
// j < width; 0 in the padding columns.  Eight elements per thread per round
// with unconditional clamped loads, so that eight loads are in flight before
// the first wait.
template <class T>
__device__ __forceinline__ void stage2d(const T* __restrict__ a, const Corr2d& p, Acc<T> fill, long long vr0,
                                        long long vc0, int rows, int width, int pitch, Acc<T>* tile)
{
    constexpr int U = 8;
    const int total = rows * pitch;
    for (int e0 = threadIdx.x; e0 < total; e0 += 256 * U) {
        Acc<T> v[U];
        int j[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = min(e0 + u * 256, total - 1);
            const int i = e / pitch;
            j[u] = e - i * pitch;
            const int ir = map2d(vr0 + i, p.r1, p.boundary);
            const int ic = map2d(vc0 + min(j[u], width - 1), p.c1, p.boundary);
            in[u] = ir >= 0 && ic >= 0;
            v[u] = load_acc(a[in[u] ? static_cast<long long>(ir) * p.c1 + ic : 0]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * 256;
            if (e < total) tile[e] = j[u] >= width ? Acc<T>(0) : (in[u] ? v[u] : fill);
        }
    }
}

constexpr int k2Rows = 16, k2Cols = 64, k2RowsPerThread = 4;  // 16x64 outputs per 256 threads

template <class T>
__global__ void __launch_bounds__(256)
correlate2d_tile_kernel(const T* __restrict__ a, const T* __restrict__ b, Corr2d p, Acc<T> fill, T* __restrict__ out)
{
    extern __shared__ unsigned char smem[];
    Acc<T>* tile = reinterpret_cast<Acc<T>*>(smem);
    const int TR = k2Rows + p.r2 - 1, TC = k2Cols + p.c2 - 1;
    const int oy0 = blockIdx.y * k2Rows, ox0 = blockIdx.x * k2Cols;
    // tile[i][j] = sample at full-correlation row oy0+rs+i-(r2-1), column ox0+cs+j-(c2-1)
    const long long vr0 = static_cast<long long>(oy0) + p.rs - (p.r2 - 1);
    const long long vc0 = static_cast<long long>(ox0) + p.cs - (p.c2 - 1);
    stage2d(a, p, fill, vr0, vc0, TR, TC, TC, tile);
    __syncthreads();
    const int tx = threadIdx.x & (k2Cols - 1), ty = threadIdx.x / k2Cols;
    const int ly = ty * k2RowsPerThread;
    Acc<T> acc[k2RowsPerThread];
#pragma unroll
    for (int q = 0; q < k2RowsPerThread; ++q) acc[q] = Acc<T>(0);
    for (int kr = 0; kr < p.r2; ++kr) {
        const Acc<T>* trow = tile + (ly + kr) * TC + tx;
        const T* brow = b + static_cast<long long>(kr) * p.c2;
        for (int kc = 0; kc < p.c2; ++kc) {
            const Acc<T> h = load_acc(brow[kc]);
#pragma unroll
            for (int q = 0; q < k2RowsPerThread; ++q) acc[q] = mac<T>(trow[q * TC + kc], h, acc[q]);
        }
    }
    const int oc = ox0 + tx;
    if (oc >= p.cols) return;
#pragma unroll
    for (int q = 0; q < k2RowsPerThread; ++q) {
        const int orow = oy0 + ly + q;
        if (orow < p.rows) out[static_cast<long long>(orow) * p.cols + oc] = store_acc<T>(acc[q]);
    }
}

// Wide form: 8 x 256 outputs per block, each thread 2 rows x 4 consecutive
// columns (twice the blocks of 4-row threads: the kernel is latency-bound).  Tile rows are padded to a multiple of 4 words (plus 8) so that a
// thread reads the 8 words from its first column on with two aligned 16-byte
// loads per row and group of 4 taps; each loaded word feeds up to 4 chains.
constexpr int k2wCols = 256, k2wColsPerThread = 4, k2wRowsPerThread = 2, k2wRows = 4 * k2wRowsPerThread;

__host__ __device__ constexpr int wide_pitch(int c2) { return ((k2wCols + c2 - 1 + 3) / 4) * 4 + 8; }

template <class T>
__global__ void __launch_bounds__(256)
correlate2d_wide_kernel(const T* __restrict__ a, const T* __restrict__ b, Corr2d p, Acc<T> fill, T* __restrict__ out)
{
    extern __shared__ unsigned char smem[];
    Acc<T>* tile = reinterpret_cast<Acc<T>*>(smem);
    const int TR = k2wRows + p.r2 - 1, TCv = k2wCols + p.c2 - 1, PITCH = wide_pitch(p.c2);
    const int oy0 = blockIdx.y * k2wRows, ox0 = blockIdx.x * k2wCols;
    const long long vr0 = static_cast<long long>(oy0) + p.rs - (p.r2 - 1);
    const long long vc0 = static_cast<long long>(ox0) + p.cs - (p.c2 - 1);
    stage2d(a, p, fill, vr0, vc0, TR, TCv, PITCH, tile);  // padding columns are read but never used
    __syncthreads();
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int ly = ty * k2wRowsPerThread, lx = tx * k2wColsPerThread;
    Acc<T> acc[k2wRowsPerThread][k2wColsPerThread];
#pragma unroll
    for (int q = 0; q < k2wRowsPerThread; ++q)
#pragma unroll
        for (int j = 0; j < k2wColsPerThread; ++j) acc[q][j] = Acc<T>(0);
    for (int kr = 0; kr < p.r2; ++kr) {
        const T* brow = b + static_cast<long long>(kr) * p.c2;
        for (int kc0 = 0; kc0 < p.c2; kc0 += 4) {
            Acc<T> h[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) h[kk] = kc0 + kk < p.c2 ? load_acc(brow[kc0 + kk]) : Acc<T>(0);
            const int nk = min(4, p.c2 - kc0);
#pragma unroll
            for (int q = 0; q < k2wRowsPerThread; ++q) {
                // words lx + kc0 .. + 7 of tile row ly + q + kr (16-byte aligned)
                const Acc<T>* src = tile + (ly + q + kr) * PITCH + lx + kc0;
                Acc<T> w[8];
                *reinterpret_cast<uint4*>(&w[0]) = *reinterpret_cast<const uint4*>(src);
                *reinterpret_cast<uint4*>(&w[4]) = *reinterpret_cast<const uint4*>(src + 4);
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    if (kk < nk) {
#pragma unroll
                        for (int j = 0; j < k2wColsPerThread; ++j) acc[q][j] = mac<T>(w[j + kk], h[kk], acc[q][j]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < k2wRowsPerThread; ++q) {
        const int orow = oy0 + ly + q;
        if (orow >= p.rows) break;
#pragma unroll
        for (int j = 0; j < k2wColsPerThread; ++j) {
            const int oc = ox0 + lx + j;
            if (oc < p.cols) out[static_cast<long long>(orow) * p.cols + oc] = store_acc<T>(acc[q][j]);
        }
    }
}

// in2 too large for the staged window: one output per thread, terms mapped
// one by one from global memory (same order).
template <class T>
__global__ void __launch_bounds__(256)
correlate2d_direct_kernel(const T* __restrict__ a, const T* __restrict__ b, Corr2d p, Acc<T> fill,
                          T* __restrict__ out)
{
    const long long e = blockIdx.x * 256LL + threadIdx.x;
    if (e >= static_cast<long long>(p.rows) * p.cols) return;
    const int r = static_cast<int>(e / p.cols), c = static_cast<int>(e - static_cast<long long>(r) * p.cols);
    const long long fr = static_cast<long long>(p.rs) + r, fc = static_cast<long long>(p.cs) + c;
    Acc<T> acc = Acc<T>(0);
    for (int kr = 0; kr < p.r2; ++kr) {
        const int ir = map2d(fr - (p.r2 - 1 - kr), p.r1, p.boundary);
        for (int kc = 0; kc < p.c2; ++kc) {
            const int ic = map2d(fc - (p.c2 - 1 - kc), p.c1, p.boundary);
            const Acc<T> x = (ir < 0 || ic < 0) ? fill : load_acc(a[static_cast<long long>(ir) * p.c1 + ic]);
            acc = mac<T>(x, load_acc(b[static_cast<long long>(kr) * p.c2 + kc]), acc);
        }
    }
    out[e] = store_acc<T>(acc);
}

constexpr size_t k2TileLdsMax = 64 * 1024;

template <class T>
int launch2d(const void* a, const void* b, const Corr2d& p, Acc<T> fill, void* out, hipStream_t st)
{
    const bool rows_ok = (p.rows + k2Rows - 1) / k2Rows <= 65535;
    const size_t lds_w = sizeof(Acc<T>) * static_cast<size_t>(k2wRows + p.r2 - 1) * static_cast<size_t>(wide_pitch(p.c2));
    if (lds_w <= 48 * 1024 && (p.rows + k2wRows - 1) / k2wRows <= 65535) {
        const dim3 grid((p.cols + k2wCols - 1) / k2wCols, (p.rows + k2wRows - 1) / k2wRows);
        hipLaunchKernelGGL((correlate2d_wide_kernel<T>), grid, dim3(256), lds_w, st, static_cast<const T*>(a),
                           static_cast<const T*>(b), p, fill, static_cast<T*>(out));
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    }
    const size_t lds = sizeof(Acc<T>) * static_cast<size_t>(k2Rows + p.r2 - 1) * static_cast<size_t>(k2Cols + p.c2 - 1);
    if (lds <= k2TileLdsMax && rows_ok) {
        if (lds > 48 * 1024) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&correlate2d_tile_kernel<T>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     static_cast<int>(k2TileLdsMax));
            ZKX_HIP_CHECK(e);
        }
        const dim3 grid((p.cols + k2Cols - 1) / k2Cols, (p.rows + k2Rows - 1) / k2Rows);
        hipLaunchKernelGGL((correlate2d_tile_kernel<T>), grid, dim3(256), lds, st, static_cast<const T*>(a),
                           static_cast<const T*>(b), p, fill, static_cast<T*>(out));
    } else {
        const long long n = static_cast<long long>(p.rows) * p.cols;
        hipLaunchKernelGGL((correlate2d_direct_kernel<T>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                           st, static_cast<const T*>(a), static_cast<const T*>(b), p, fill, static_cast<T*>(out));
    }
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

// The fill value in the element type, widened to the accumulator
// (convolution_typed.cpp:28-37, 116-131): integers truncate toward zero and
// wrap modulo 2^bits; half rounds the float value to nearest; float rounds.
bool fill_bits(int dtype, double v, float* f, uint32_t* u)
{
    *f = 0.0f;
    *u = 0;
    switch (dtype) {
    case ZKX_DTYPE_F32: *f = static_cast<float>(v); return true;
    case ZKX_DTYPE_F16: *f = static_cast<float>(static_cast<_Float16>(static_cast<float>(v))); return true;
    default: break;
    }
    if (!std::isfinite(v)) return false;
    const int bits = dtype == ZKX_DTYPE_I32 ? 32 : dtype == ZKX_DTYPE_I16 ? 16 : 8;
    const double modulus = std::ldexp(1.0, bits);
    double w = std::fmod(std::trunc(v), modulus);
    if (w < 0.0) w += modulus;
    const uint32_t raw = static_cast<uint32_t>(w);
    // sign-extend the bits-wide value to 32 bits, as load_acc does for T
    const uint32_t sign = 1u << (bits - 1);
    *u = bits == 32 ? raw : ((raw & sign) ? (raw | ~((sign << 1) - 1u)) : raw);
    return true;
}

// ---------------------------------------------------------------------------
// FFT method (convolution_typed.h:307-327): the inputs converted to complex
// FP32, in2 reversed, the full convolution by the overlap-save FFT kernel,
// then the centred crop stored in the element type: float as is, half rounded
// to nearest, integers rounded half to even and wrapped modulo 2^bits (no
// saturation).  FP32 FFT arithmetic: not the direct sums bit for bit.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256)
to_c64_kernel(const T* __restrict__ x, long long n, int reverse, float2* __restrict__ y)
{
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n) y[i] = make_float2(load_f(x[reverse ? n - 1 - i : i]), 0.0f);
}

template <class T>
__device__ __forceinline__ T store_rounded(float v)
{
    if constexpr (std::is_integral_v<T>) {
        using U = std::make_unsigned_t<T>;
        constexpr double modulus = static_cast<double>(std::numeric_limits<U>::max()) + 1.0;
        if (!isfinite(v)) return T(0);
        double w = fmod(rint(static_cast<double>(v)), modulus);
        if (w < 0.0) w += modulus;
        return static_cast<T>(static_cast<U>(w));
    } else {
        return store_f<T>(v);
    }
}

template <class T>
__global__ void __launch_bounds__(256)
from_c64_kernel(const float2* __restrict__ y, long long start, long long n, T* __restrict__ out)
{
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n) out[i] = store_rounded<T>(y[start + i].x);
}

thread_local zkx::ThreadWorkspace t_corr_fft_ws;

template <class T>
int correlate_fft(const void* in1, long long l1, const void* in2, long long l2, long long start, long long n,
                  void* out, hipStream_t st)
{
    const long long full = l1 + l2 - 1;
    // both operands longer than the overlap-save kernel takes: one
    // whole-length transform pair instead
    const bool whole = std::min(l1, l2) > 4096;
    const size_t conv_bytes = whole ? zkx::fft_convolve_workspace_bytes(l1, l2) : 0;
    void* ws = nullptr;
    zkx::WorkspaceLease lease_(t_corr_fft_ws, st);
    const size_t io_bytes = (sizeof(float2) * static_cast<size_t>(l1 + l2 + full) + 255) / 256 * 256;
    int rc = lease_.get(io_bytes + conv_bytes, &ws);
    if (rc) return rc;
    float2* A = static_cast<float2*>(ws);
    float2* B = A + l1;
    float2* Y = B + l2;
    hipLaunchKernelGGL((to_c64_kernel<T>), dim3(static_cast<unsigned>((l1 + 255) / 256)), dim3(256), 0, st,
                       static_cast<const T*>(in1), l1, 0, A);
    ZKX_LAUNCH_CHECK();
    hipLaunchKernelGGL((to_c64_kernel<T>), dim3(static_cast<unsigned>((l2 + 255) / 256)), dim3(256), 0, st,
                       static_cast<const T*>(in2), l2, 1, B);
    ZKX_LAUNCH_CHECK();
    if (whole) {
        rc = zkx::fft_convolve_full(A, l1, B, l2, Y, static_cast<char*>(ws) + io_bytes, st);
        if (rc) return rc;
        hipLaunchKernelGGL((from_c64_kernel<T>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, Y,
                           start, n, static_cast<T*>(out));
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    }
    // the full convolution is symmetric in its operands: the longer one is
    // the signal, the shorter (at most 4096 samples) the kernel
    const bool swap = l2 > l1;
    rc = zkx_fftconvolve_c64(reinterpret_cast<const zkx_complex64*>(swap ? B : A), 1, static_cast<int>(swap ? l2 : l1),
                             reinterpret_cast<const zkx_complex64*>(swap ? A : B), static_cast<int>(swap ? l1 : l2), 0,
                             reinterpret_cast<zkx_complex64*>(Y), st);
    if (rc) return rc;
    hipLaunchKernelGGL((from_c64_kernel<T>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, Y, start,
                       n, static_cast<T*>(out));
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

// ---------------------------------------------------------------------------
// N-D correlate by FFT (correlate_device with a CorrelateNDWorkspace,
// ZKX/cusignal_cpp/src/operators/convolution/convolution_typed.cu:150-241;
// shape rules convolution_typed.cpp:153-199): every dimension padded to the
// power of two >= s1 + s2 - 1, in2 read backwards in every dimension so the
// full linear convolution is the full correlation, complex FP32 transforms,
// then the centred crop start (full - out)/2 stored with the 1-D FFT
// method's rounding (integers half to even, wrapped).
// ---------------------------------------------------------------------------
template <class T>
struct NdLoadTyped {
    const T* p;
    __device__ float2 operator()(long long i) const { return make_float2(load_f(p[i]), 0.0f); }
};

template <class T>
struct NdStoreTyped {
    T* p;
    __device__ void operator()(long long i, float2 v) const { p[i] = store_rounded<T>(v.x); }
};

template <class T>
__global__ void scalar_product_kernel(const T* a, const T* b, T* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = store_rounded<T>(load_f(a[0]) * load_f(b[0]));
}

thread_local zkx::ThreadWorkspace t_corr_nd_ws;

template <class T>
int correlate_nd_fft(const void* in1, const int* s1, const void* in2, const int* s2, int rank, const int* out_dims,
                     void* out, hipStream_t st)
{
    using namespace zkx;
    NdConvPlan p;
    int rc = nd_conv_plan(s1, s2, rank, (1u << rank) - 1u, &p);
    if (rc) return rc;
    void* wsv = nullptr;
    WorkspaceLease lease_(t_corr_nd_ws, st);
    rc = lease_.get(p.bytes, &wsv);
    if (rc) return rc;
    char* ws = static_cast<char*>(wsv);
    // both operands are real: one packed transform of in1 + i s reverse(in2)
    if ((rc = nd_pack_real_pair(p, ws, NdLoadTyped<T>{static_cast<const T*>(in1)}, s1,
                                NdLoadTyped<T>{static_cast<const T*>(in2)}, s2, 1, st)))
        return rc;
    float2* Y = nullptr;
    p.keep_lo0 = (s1[0] + s2[0] - 1 - out_dims[0]) / 2;  // the crop's rows of dimension 0
    p.keep_hi0 = p.keep_lo0 + out_dims[0];
    if ((rc = nd_conv_run_real_pair(p, ws, &Y, st))) return rc;
    NdCrop c{};
    c.rank = rank;
    nd_strides(p.g, rank, c.ystride);
    long long rows = 1;
    for (int d = 0; d < rank; ++d) {
        c.o[d] = out_dims[d];
        c.start[d] = (s1[d] + s2[d] - 1 - out_dims[d]) / 2;
        if (d + 1 < rank) rows *= out_dims[d];
    }
    hipLaunchKernelGGL(nd_crop_kernel<NdStoreTyped<T>>, nd_grid(rows, out_dims[rank - 1]), dim3(256), 0, st,
                       Y, c, NdStoreTyped<T>{static_cast<T*>(out)});
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

// Output extents of the N-D correlate (convolution_typed.cpp:153-199):
// full s1 + s2 - 1, same s1, valid |s1 - s2| + 1 per dimension; 'valid'
// needs one input at least as large as the other in every dimension.
// Returns the element count (0 when an input is empty), or minus the error
// code of a failure recorded in the last error.
long long correlate_nd_shape(const int* shape1, const int* shape2, int rank, int mode, int* out_dims)
{
    if (mode < 0 || mode > 2) {
        zkx::set_last_error(ZKX_INVALID_ARGUMENT, "typed correlate mode must be full, same, or valid");
        return -ZKX_INVALID_ARGUMENT;
    }
    if (rank < 0 || rank > zkx::kMaxNdRank || (rank > 0 && (!shape1 || !shape2))) {
        zkx::set_last_error(ZKX_INVALID_SHAPE, "typed correlate rank must be 0..8 with both shapes given");
        return -ZKX_INVALID_SHAPE;
    }
    long long n1 = 1, n2 = 1;
    for (int d = 0; d < rank; ++d) {
        if (shape1[d] < 0 || shape2[d] < 0) {
            zkx::set_last_error(ZKX_INVALID_SHAPE, "typed correlate negative dimension");
            return -ZKX_INVALID_SHAPE;
        }
        n1 *= shape1[d];
        n2 *= shape2[d];
        if (n1 > std::numeric_limits<int>::max() || n2 > std::numeric_limits<int>::max()) {
            zkx::set_last_error(ZKX_INVALID_SHAPE, "typed correlate input exceeds 32-bit indexing");
            return -ZKX_INVALID_SHAPE;
        }
    }
    if (n1 == 0 || n2 == 0) {
        if (out_dims && rank > 0) out_dims[0] = 0;
        return 0;
    }
    bool dom1 = true, dom2 = true;
    long long n = 1;
    for (int d = 0; d < rank; ++d) {
        const int a = shape1[d], b = shape2[d];
        dom1 = dom1 && a >= b;
        dom2 = dom2 && b >= a;
        const long long e = mode == 0 ? static_cast<long long>(a) + b - 1 : mode == 1 ? a : std::abs(a - b) + 1;
        if (e > std::numeric_limits<int>::max()) {
            zkx::set_last_error(ZKX_INVALID_SHAPE, "typed correlate full shape exceeds 32-bit indexing");
            return -ZKX_INVALID_SHAPE;
        }
        if (out_dims) out_dims[d] = static_cast<int>(e);
        n *= e;
    }
    if (mode == 2 && !dom1 && !dom2) {
        zkx::set_last_error(ZKX_INVALID_SHAPE,
                            "typed correlate valid mode requires one input to dominate in every dimension");
        return -ZKX_INVALID_SHAPE;
    }
    if (n > std::numeric_limits<int>::max()) {
        zkx::set_last_error(ZKX_INVALID_SHAPE, "typed correlate output exceeds 32-bit indexing");
        return -ZKX_INVALID_SHAPE;
    }
    return n;
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

int zkx_correlate_fft(const void* in1, long long l1, const void* in2, long long l2, int dtype, int mode, void* out,
                      void* stream)
{
    zkx::clear_last_error();
    if (mode < 0 || mode > 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate mode must be full, same, or valid");
    const size_t es = elem_size(dtype);
    if (!es) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate element type");
    if (l1 < 0 || l2 < 0 || l1 + l2 - 1 > std::numeric_limits<int>::max())
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate input exceeds 32-bit indexing");
    const long long n = zkx_correlate_out_len(l1, l2, mode);
    if (n == 0) return ZKX_SUCCESS;
    if (l1 + l2 - 1 > zkx::kMaxAnyFftPow2)
        ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "typed correlate fft method needs l1 + l2 - 1 <= 2^26");
    if (!in1 || !in2 || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate: null pointer");
    if (overlap(out, es * n, in1, es * l1) || overlap(out, es * n, in2, es * l2))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate output must not alias an input");
    const long long start = (l1 + l2 - 1 - n) / 2;
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (dtype) {
    case ZKX_DTYPE_F32: return correlate_fft<float>(in1, l1, in2, l2, start, n, out, st);
    case ZKX_DTYPE_F16: return correlate_fft<__half>(in1, l1, in2, l2, start, n, out, st);
    case ZKX_DTYPE_I32: return correlate_fft<int32_t>(in1, l1, in2, l2, start, n, out, st);
    case ZKX_DTYPE_I16: return correlate_fft<int16_t>(in1, l1, in2, l2, start, n, out, st);
    default: return correlate_fft<int8_t>(in1, l1, in2, l2, start, n, out, st);
    }
}

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

int zkx_correlate_nd_out_shape(const int* shape1, const int* shape2, int rank, int mode, int* out_shape,
                               long long* elements)
{
    zkx::clear_last_error();
    if (!elements || (rank > 0 && !out_shape)) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate: null output shape");
    const long long n = correlate_nd_shape(shape1, shape2, rank, mode, out_shape);
    if (n < 0) return static_cast<int>(-n);
    *elements = rank == 0 ? 1 : n;
    return ZKX_SUCCESS;
}

int zkx_correlate_nd(const void* in1, const int* shape1, const void* in2, const int* shape2, int rank, int dtype,
                     int mode, void* out, void* stream)
{
    zkx::clear_last_error();
    const size_t es = elem_size(dtype);
    if (!es) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate element type");
    int dims[zkx::kMaxNdRank] = {};
    const long long n = correlate_nd_shape(shape1, shape2, rank, mode, dims);
    if (n < 0) return static_cast<int>(-n);
    if (rank > 0 && n == 0) return ZKX_SUCCESS;
    if (!in1 || !in2 || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate: null pointer");
    long long n1 = 1, n2 = 1;
    for (int d = 0; d < rank; ++d) {
        n1 *= shape1[d];
        n2 *= shape2[d];
    }
    const long long nout = rank == 0 ? 1 : n;
    if (overlap(out, es * nout, in1, es * n1) || overlap(out, es * nout, in2, es * n2))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate output must not alias an input");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (rank == 0) {
        // scalars: the product stored in the element type
        switch (dtype) {
        case ZKX_DTYPE_F32: hipLaunchKernelGGL(scalar_product_kernel<float>, dim3(1), dim3(64), 0, st,
                                               static_cast<const float*>(in1), static_cast<const float*>(in2),
                                               static_cast<float*>(out)); break;
        case ZKX_DTYPE_F16: hipLaunchKernelGGL(scalar_product_kernel<__half>, dim3(1), dim3(64), 0, st,
                                               static_cast<const __half*>(in1), static_cast<const __half*>(in2),
                                               static_cast<__half*>(out)); break;
        case ZKX_DTYPE_I32: hipLaunchKernelGGL(scalar_product_kernel<int32_t>, dim3(1), dim3(64), 0, st,
                                               static_cast<const int32_t*>(in1), static_cast<const int32_t*>(in2),
                                               static_cast<int32_t*>(out)); break;
        case ZKX_DTYPE_I16: hipLaunchKernelGGL(scalar_product_kernel<int16_t>, dim3(1), dim3(64), 0, st,
                                               static_cast<const int16_t*>(in1), static_cast<const int16_t*>(in2),
                                               static_cast<int16_t*>(out)); break;
        default: hipLaunchKernelGGL(scalar_product_kernel<int8_t>, dim3(1), dim3(64), 0, st,
                                    static_cast<const int8_t*>(in1), static_cast<const int8_t*>(in2),
                                    static_cast<int8_t*>(out));
        }
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    }
    if (rank == 1) return zkx_correlate_fft(in1, shape1[0], in2, shape2[0], dtype, mode, out, stream);
    switch (dtype) {
    case ZKX_DTYPE_F32: return correlate_nd_fft<float>(in1, shape1, in2, shape2, rank, dims, out, st);
    case ZKX_DTYPE_F16: return correlate_nd_fft<__half>(in1, shape1, in2, shape2, rank, dims, out, st);
    case ZKX_DTYPE_I32: return correlate_nd_fft<int32_t>(in1, shape1, in2, shape2, rank, dims, out, st);
    case ZKX_DTYPE_I16: return correlate_nd_fft<int16_t>(in1, shape1, in2, shape2, rank, dims, out, st);
    default: return correlate_nd_fft<int8_t>(in1, shape1, in2, shape2, rank, dims, out, st);
    }
}

int zkx_correlate2d_out_shape(long long r1, long long c1, long long r2, long long c2, int mode, long long* rows,
                              long long* cols)
{
    zkx::clear_last_error();
    if (!rows || !cols) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate2d: null shape pointer");
    *rows = *cols = 0;
    if (mode < 0 || mode > 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate2d mode must be full, same, or valid");
    constexpr long long kMax = std::numeric_limits<int>::max();
    if (r1 < 0 || c1 < 0 || r2 < 0 || c2 < 0) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate2d negative dimension");
    if (r1 * c1 > kMax || r2 * c2 > kMax) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate2d input exceeds 32-bit indexing");
    if (r1 == 0 || c1 == 0 || r2 == 0 || c2 == 0) return ZKX_SUCCESS;
    long long R, C;
    if (mode == 0) {
        R = r1 + r2 - 1;
        C = c1 + c2 - 1;
        if (R > kMax || C > kMax) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate2d full shape exceeds 32-bit indexing");
    } else if (mode == 1) {
        R = r1;
        C = c1;
    } else {
        const bool d1 = r1 >= r2 && c1 >= c2, d2 = r2 >= r1 && c2 >= c1;
        if (!d1 && !d2)
            ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate2d valid mode requires one input to dominate in every dimension");
        R = std::llabs(r1 - r2) + 1;
        C = std::llabs(c1 - c2) + 1;
    }
    if (R * C > kMax) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate2d output exceeds 32-bit indexing");
    *rows = R;
    *cols = C;
    return ZKX_SUCCESS;
}

int zkx_correlate2d(const void* in1, long long r1, long long c1, const void* in2, long long r2, long long c2, int dtype,
                    int mode, int boundary, double fillvalue, void* out, void* stream)
{
    long long rows = 0, cols = 0;
    int rc = zkx_correlate2d_out_shape(r1, c1, r2, c2, mode, &rows, &cols);
    if (rc) return rc;
    if (boundary < 0 || boundary > 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate2d boundary must be fill, wrap, or symm");
    const size_t es = elem_size(dtype);
    if (!es) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate2d element type");
    float ff = 0.0f;
    uint32_t fu = 0;
    if (mode != 2 && boundary == 0 && !fill_bits(dtype, fillvalue, &ff, &fu))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate2d integral fillvalue must be finite");
    if (rows * cols == 0) return ZKX_SUCCESS;
    if (!in1 || !in2 || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "correlate2d: null pointer");
    if (overlap(out, es * rows * cols, in1, es * r1 * c1) || overlap(out, es * rows * cols, in2, es * r2 * c2))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "typed correlate2d output must not alias an input");
    Corr2d p{static_cast<int>(r1), static_cast<int>(c1), static_cast<int>(r2), static_cast<int>(c2),
             static_cast<int>(rows), static_cast<int>(cols), 0, 0, mode == 2 ? 0 : boundary};
    if (mode == 1) {
        p.rs = p.r2 / 2;
        p.cs = p.c2 / 2;
    } else if (mode == 2) {
        p.rs = std::min(p.r1, p.r2) - 1;
        p.cs = std::min(p.c1, p.c2) - 1;
        ff = 0.0f;
        fu = 0;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (dtype) {
    case ZKX_DTYPE_F32: return launch2d<float>(in1, in2, p, ff, out, st);
    case ZKX_DTYPE_F16: return launch2d<__half>(in1, in2, p, ff, out, st);
    case ZKX_DTYPE_I32: return launch2d<int32_t>(in1, in2, p, fu, out, st);
    case ZKX_DTYPE_I16: return launch2d<int16_t>(in1, in2, p, fu, out, st);
    default: return launch2d<int8_t>(in1, in2, p, fu, out, st);
    }
}

}  // extern "C"
