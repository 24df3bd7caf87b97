// Polyphase channelizer for gfx950.
//
// Follows channelize_poly_device / channelize_poly_typed_cpu
// (ZKX/cusignal_cpp/src/operators/filtering/filtering_typed.cpp:443-466 plan,
//  the CPU form after it; cuSignal filtering/channelize_poly.py):
//
//   M = n_chans channels, T = floor(len(h)/M) <= 32 taps per phase,
//   P = floor(N/M) points (a short last group of x and of h is ignored):
//   F[ph][p] = sum_{t<T, p-t>=0} x[(p-t)M + M-1-ph] * h[ph + tM]
//   Y[c][p]  = sum_{ph<M} F[ph][p] e^{+2 pi i c ph / M}       (channel-major)
//
// Three forms:
//   * M in {1, 2, 4, 8}: one thread per point filters all M phases from an
//     LDS copy of the block's input rows and runs the M-point inverse DFT in
//     registers (the twiddles are folded at compile time), then stores
//     Y[c][p] for every c -- consecutive threads store consecutive points of
//     a channel.
//   * M in {16, 32, 64, 128, 256}: one kernel filters by phase over runs of
//     points, inverts each point's row with the LDS FFT and stores through an
//     LDS transpose tile (no workspace).
//   * any other M: the filtered phases go to a [P][M] complex workspace, the
//     library's row FFT inverts each row (LDS Stockham for powers of two,
//     Bluestein otherwise), and a 32x32 LDS transpose writes Y channel-major.
// All arithmetic is FP32; the reference's GPU path is FP32 as well (its CPU
// form sums in FP64).
#include "runtime.hpp"
#include "fft_lds.hpp"
#include "../../include/zkx_fft.h"

#include <limits>

namespace zkx {
int fft_rows_dispatch(int n, const float2* x, long long rows, bool inverse, float scale, float2* y,
                      hipStream_t st);
}

namespace {

constexpr int kMaxTaps = 32;

// A block covers 256 consecutive points: the x groups they need (T - 1
// before the first point onwards) and the taps are staged in LDS once, rows
// padded to M + 1 floats so that the lanes (consecutive points) hit distinct
// banks.  Groups before the start of x are zero: their products add exact
// zeros, so every sum equals the one that skips them.
constexpr int kSmallPoints = 256;

template <int M>
__global__ void __launch_bounds__(kSmallPoints)
channelize_small_kernel(const float* __restrict__ x, const float* __restrict__ h, int T, int P,
                        float2* __restrict__ y)
{
    extern __shared__ float sm[];
    constexpr int RS = M + 1;
    const int rows = kSmallPoints + T - 1;
    float* xs = sm;               // [rows][RS]
    float* hs = sm + rows * RS;   // [T*M]
    const int p0 = blockIdx.x * kSmallPoints;
    const long long g0 = static_cast<long long>(p0) - (T - 1);
    for (int e = threadIdx.x; e < rows * M; e += kSmallPoints) {
        const int r = e / M, j = e - r * M;
        const long long g = g0 + r;
        xs[r * RS + j] = (g >= 0 && g < P) ? x[g * M + j] : 0.0f;
    }
    for (int e = threadIdx.x; e < T * M; e += kSmallPoints) hs[e] = h[e];
    __syncthreads();
    const int l = threadIdx.x;
    const int p = p0 + l;
    if (p >= P) return;  // after the only barrier
    float f[M];
#pragma unroll
    for (int ph = 0; ph < M; ++ph) f[ph] = 0.0f;
    for (int t = 0; t < T; ++t) {
        const float* xr = xs + (l + T - 1 - t) * RS;
        const float* hr = hs + t * M;
#pragma unroll
        for (int ph = 0; ph < M; ++ph) f[ph] = fmaf(xr[M - 1 - ph], hr[ph], f[ph]);
    }
    float2 v[M];
#pragma unroll
    for (int ph = 0; ph < M; ++ph) v[ph] = make_float2(f[ph], 0.0f);
    zkx::fft::Dft<M, true>::run(v);
#pragma unroll
    for (int c = 0; c < M; ++c) y[static_cast<long long>(c) * P + p] = v[c];
}

// F[p][ph] as a complex row per point (imaginary part zero).  A thread runs
// one phase over kRun consecutive points: the terms x[(p - t)M + M-1-ph]
// h[ph + tM] of outputs p0..p0+kRun-1 for one t are kRun consecutive rows of
// the phase's column, so a window of kRun rows slid down one row per tap
// feeds kRun chains per load (lanes take consecutive phases: coalesced).
// Rows before the start of x are skipped per term, as the reference skips
// them; rows past P only feed outputs that are not stored.
constexpr int kRun = 16;

__device__ __forceinline__ void fir_run(const float* __restrict__ x, const float* __restrict__ h, int M, int T,
                                        int P, int ph, int p0, float (&acc)[kRun])
{
    const float* col = x + (M - 1 - ph);
    float w[kRun];
#pragma unroll
    for (int i = 0; i < kRun; ++i) {
        w[i] = col[static_cast<long long>(min(p0 + i, P - 1)) * M];
        acc[i] = 0.0f;
    }
    if (p0 >= T - 1) {
        // every row p0 + i - t (t < T) is inside x
        for (int t = 0; t < T; ++t) {
            const float ht = h[ph + static_cast<long long>(t) * M];
            const float nx = col[static_cast<long long>(max(p0 - t - 1, 0)) * M];
#pragma unroll
            for (int i = 0; i < kRun; ++i) acc[i] = fmaf(w[i], ht, acc[i]);
#pragma unroll
            for (int i = kRun - 1; i > 0; --i) w[i] = w[i - 1];
            w[0] = nx;
        }
    } else {
        for (int t = 0; t < T; ++t) {
            const float ht = h[ph + static_cast<long long>(t) * M];
            const float nx = col[static_cast<long long>(max(p0 - t - 1, 0)) * M];
#pragma unroll
            for (int i = 0; i < kRun; ++i) {
                const float y = fmaf(w[i], ht, acc[i]);
                acc[i] = (p0 + i - t >= 0) ? y : acc[i];
            }
#pragma unroll
            for (int i = kRun - 1; i > 0; --i) w[i] = w[i - 1];
            w[0] = nx;
        }
    }
}

__global__ void __launch_bounds__(256)
channelize_fir_kernel(const float* __restrict__ x, const float* __restrict__ h, int M, int T, int P,
                      long long threads, float2* __restrict__ ws)
{
    const long long e = blockIdx.x * 256LL + threadIdx.x;
    if (e >= threads) return;
    const int run = static_cast<int>(e / M);
    const int ph = static_cast<int>(e - static_cast<long long>(run) * M);
    const int p0 = run * kRun;
    float acc[kRun];
    fir_run(x, h, M, T, P, ph, p0, acc);
#pragma unroll
    for (int i = 0; i < kRun; ++i)
        if (p0 + i < P) ws[static_cast<long long>(p0 + i) * M + ph] = make_float2(acc[i], 0.0f);
}

// M in {32, 64, 128, 256}: the whole general form in one block of 256
// threads for FPB = 4096/M points.  The FIR (a thread per phase and run of
// 16 points) writes the real rows into each point's LDS slot; each row's
// M-point unscaled inverse FFT runs in LDS (M/16 threads per row); the
// spectra go through an LDS [M][FPB+1] tile so that the stores run along
// points (channel-major output).  One HBM read of x, one write of y.
template <int M>
__global__ void __launch_bounds__(256)
channelize_fused_kernel(const float* __restrict__ x, const float* __restrict__ h, int T, int P,
                        const float2* __restrict__ tw, float2* __restrict__ y)
{
    using C = zkx::fft::RowCfg<M>;
    static_assert(C::THREADS == 256 && C::FPB % kRun == 0 && (C::FPB / kRun) * M == 256, "launch shape");
    static_assert(M * (C::FPB + 1) <= C::SLOT_FLOATS * C::FPB, "transpose tile fits the FFT slots");
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int p0 = blockIdx.x * C::FPB;
    {
        const int ph = tid % M, run = tid / M;
        float acc[kRun];
        fir_run(x, h, M, T, P, ph, p0 + run * kRun, acc);
#pragma unroll
        for (int i = 0; i < kRun; ++i) lds[(run * kRun + i) * 2 * C::SLOT_FLOATS + ph] = acc[i];
    }
    __syncthreads();
    const int slot = tid / C::T, t = tid % C::T;
    float* base = lds + slot * 2 * C::SLOT_FLOATS;
    zkx::fft::RowLds L{reinterpret_cast<float2*>(base)};
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) v[k] = make_float2(base[t + k * C::T], 0.0f);
    zkx::fft::transform<M, C::E, true>(v, L, tw, t);  // syncs before its first LDS store
    __syncthreads();                                  // the last pass's LDS reads are done
    float2* yt = reinterpret_cast<float2*>(lds);      // [M][FPB + 1]
    constexpr int YS = C::FPB + 1;
#pragma unroll
    for (int k = 0; k < C::E; ++k) yt[(t + k * C::T) * YS + slot] = v[k];
    __syncthreads();
#pragma unroll 4
    for (int e = tid; e < M * C::FPB; e += 256) {
        const int c = e / C::FPB, s = e % C::FPB;
        if (p0 + s < P) y[static_cast<long long>(c) * P + p0 + s] = yt[c * YS + s];
    }
}

// y[c][p] = scale * g[p][c] through a 32x32 tile.
__global__ void __launch_bounds__(256)
channelize_transpose_kernel(const float2* __restrict__ g, int P, int M, float scale, float2* __restrict__ y)
{
    __shared__ float2 tile[32][33];
    const int p0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows of 32
    for (int r = ty; r < 32; r += 8) {
        const int p = p0 + r, c = c0 + tx;
        if (p < P && c < M) tile[r][tx] = g[static_cast<long long>(p) * M + c];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int c = c0 + r, p = p0 + tx;
        if (p < P && c < M) {
            const float2 v = tile[tx][r];
            y[static_cast<long long>(c) * P + p] = make_float2(v.x * scale, v.y * scale);
        }
    }
}

template <int M>
int launch_fused(const float* x, const float* h, int T, int P, float2* y, hipStream_t st)
{
    using C = zkx::fft::RowCfg<M>;
    const float2* tw = nullptr;
    const int rc = zkx::twiddles(M, &tw);
    if (rc) return rc;
    hipLaunchKernelGGL((channelize_fused_kernel<M>), dim3((P + C::FPB - 1) / C::FPB), dim3(256), C::LDS_BYTES, st, x,
                       h, T, P, tw, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

template <int M>
int launch_small(const float* x, const float* h, int T, int P, float2* y, hipStream_t st)
{
    const size_t lds = sizeof(float) * (static_cast<size_t>(kSmallPoints + T - 1) * (M + 1) + static_cast<size_t>(T) * M);
    hipLaunchKernelGGL((channelize_small_kernel<M>), dim3((P + kSmallPoints - 1) / kSmallPoints), dim3(kSmallPoints),
                       lds, st, x, h, T, P, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

thread_local zkx::ThreadWorkspace t_chan_ws;

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

}  // namespace

extern "C" {

int zkx_channelize_poly_f32(const float* x, long long x_count, const float* h, long long h_count, int n_chans,
                            zkx_complex64* out, void* stream)
{
    zkx::clear_last_error();
    if (n_chans <= 0) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "channelize_poly n_chans");
    if (x_count < 0 || h_count < 0 || x_count > std::numeric_limits<int>::max() ||
        h_count > std::numeric_limits<int>::max())
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "channelize_poly input exceeds int indexing");
    const int M = n_chans;
    const int T = static_cast<int>(h_count / M);
    if (T > kMaxTaps) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "channelize_poly taps per phase exceed 32");
    const int P = static_cast<int>(x_count / M);
    const long long total = static_cast<long long>(M) * P;
    if (total > std::numeric_limits<int>::max()) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "channelize_poly output exceeds int indexing");
    if (total == 0) return ZKX_SUCCESS;
    if (!x || !out || (T > 0 && !h)) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "channelize_poly: null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    float2* y = reinterpret_cast<float2*>(out);
    switch (M) {
    case 1: return launch_small<1>(x, h, T, P, y, st);
    case 2: return launch_small<2>(x, h, T, P, y, st);
    case 4: return launch_small<4>(x, h, T, P, y, st);
    case 8: return launch_small<8>(x, h, T, P, y, st);
    case 16: return launch_fused<16>(x, h, T, P, y, st);
    case 32: return launch_fused<32>(x, h, T, P, y, st);
    case 64: return launch_fused<64>(x, h, T, P, y, st);
    case 128: return launch_fused<128>(x, h, T, P, y, st);
    case 256: return launch_fused<256>(x, h, T, P, y, st);
    default: break;
    }
    if (!pow2(M) && M > 4096)
        ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "channelize_poly channel counts that are not powers of two need M <= 4096");
    if (M > 8192) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "channelize_poly needs M <= 8192");
    void* ws = nullptr;
    zkx::WorkspaceLease lease_(t_chan_ws, st);
    int rc = lease_.get(2 * sizeof(float2) * static_cast<size_t>(total), &ws);
    if (rc) return rc;
    float2* f = static_cast<float2*>(ws);
    float2* g = f + total;
    const long long fir_threads = static_cast<long long>(M) * ((P + kRun - 1) / kRun);
    hipLaunchKernelGGL(channelize_fir_kernel, dim3(static_cast<unsigned>((fir_threads + 255) / 256)), dim3(256), 0, st,
                       x, h, M, T, P, fir_threads, f);
    ZKX_LAUNCH_CHECK();
    float scale = 1.0f;
    if (pow2(M)) {
        rc = zkx::fft_rows_dispatch(M, f, P, true, 1.0f, g, st);  // unscaled inverse
        if (rc) return rc;
    } else {
        FFTPlanHandle plan = nullptr;
        if (fft_plan_create_1d(&plan, M, P, FFT_TYPE_C2C, FFT_INVERSE) != FFT_SUCCESS)
            ZKX_FAIL(ZKX_BACKEND_ERROR, "channelize_poly FFT plan");
        const FFTResult fr = fft_execute_on_stream(plan, g, f, stream);  // inverse scaled by 1/M
        // the plan's tables must outlive the queued transform
        const hipError_t se = hipStreamSynchronize(st);
        fft_plan_destroy(plan);
        ZKX_HIP_CHECK(se);
        if (fr != FFT_SUCCESS) ZKX_FAIL(ZKX_BACKEND_ERROR, "channelize_poly FFT");
        scale = static_cast<float>(M);
    }
    hipLaunchKernelGGL(channelize_transpose_kernel, dim3((P + 31) / 32, (M + 31) / 32), dim3(256), 0, st, g, P, M,
                       scale, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // extern "C"
