// fftconvolve on gfx950: linear convolution of complex64 rows with one
// complex64 kernel by overlap-save over the LDS Stockham FFT.
//
// Semantics follow cuSignal's fftconvolve for 1-D inputs
// (ZKX/cusignal-23.08.00/python/cusignal/convolution/convolve.py:159-320 and
// convolution_utils.py:94-100 _centered): 'full' gives n1+n2-1 samples,
// 'same' the n1 samples centred in the full result (start (n2-1)/2), 'valid'
// the |n1-n2|+1 samples that use no zero padding (start min(n1,n2)-1).  The
// reference's C++ library has no fftconvolve (operator_contracts.cmake:10-63).
//
// One N-point transform per (row, segment): segment s covers full-output
// indices [s*L, s*L+L) with L = N - n2 + 1, reads in1[s*L-(n2-1) .. s*L+L-1]
// (zero outside the row), multiplies by the kernel spectrum and keeps the L
// circularly-clean outputs.  A short signal is a single segment.
#include "fft_lds.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <limits>

namespace zkx {
int fft_rows_dispatch(int n, const float2* x, long long rows, bool inverse, float scale, float2* y,
                      hipStream_t st);

namespace {

template <int N>
__global__ void __launch_bounds__(fft::RowCfg<N>::THREADS)
ols_kernel(const float2* __restrict__ x, long long rows, int n1, int n2, long long seg0, long long segs, int L,
           const float2* __restrict__ S, const float2* __restrict__ tw, long long out_start, long long out_len,
           float2* __restrict__ y)
{
    using C = fft::RowCfg<N>;
    extern __shared__ float lds[];
    const int slot = threadIdx.x / C::T, t = threadIdx.x % C::T;
    const long long item = static_cast<long long>(blockIdx.x) * C::FPB + slot;
    const bool active = item < rows * segs;
    const long long row = active ? item / segs : 0;
    const long long seg = seg0 + (active ? item % segs : 0);
    fft::RowLds Ld{reinterpret_cast<float2*>(lds + slot * 2 * C::SLOT_FLOATS)};
    const float2* src = x + row * n1;
    const long long in0 = seg * L - (n2 - 1);
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const long long u = in0 + t + k * C::T;
        v[k] = (active && u >= 0 && u < n1) ? src[u] : make_float2(0.f, 0.f);
    }
    fft::transform<N, C::E, false>(v, Ld, tw, t);
#pragma unroll
    for (int k = 0; k < C::E; ++k) v[k] = cmul(v[k], S[t + k * C::T]);
    fft::transform<N, C::E, true>(v, Ld, tw, t);
    if (!active) return;
    const float sc = 1.0f / static_cast<float>(N);
    float2* dst = y + row * out_len;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int i = t + k * C::T;  // circular position; clean for i >= n2-1
        if (i < n2 - 1) continue;
        const long long f = seg * L + (i - (n2 - 1));  // full-convolution index
        const long long o = f - out_start;
        if (o >= 0 && o < out_len) dst[o] = cscale(v[k], sc);
    }
}

__global__ void pad_kernel(const float2* __restrict__ in, int n, int N, float2* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) out[i] = i < n ? in[i] : make_float2(0.f, 0.f);
}

template <int N>
int launch_ols(const float2* x, long long rows, int n1, int n2, long long out_start, long long out_len,
               float2* S, float2* y, hipStream_t st)
{
    using C = fft::RowCfg<N>;
    const float2* tw = nullptr;
    int rc = twiddles(N, &tw);
    if (rc) return rc;
    const int L = N - n2 + 1;
    const long long full = static_cast<long long>(n1) + n2 - 1;
    // only the segments that reach [out_start, out_start + out_len) are needed
    const long long seg_lo = out_start / L;
    const long long seg_hi = (std::min(full, out_start + out_len) + L - 1) / L;
    const long long segs = seg_hi - seg_lo;
    const long long items = rows * segs;
    const long long blocks = (items + C::FPB - 1) / C::FPB;
    if (blocks > std::numeric_limits<int>::max()) ZKX_FAIL(ZKX_INVALID_SHAPE, "fftconvolve too large");
    if (C::LDS_BYTES > 64 * 1024)
        ZKX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(ols_kernel<N>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          static_cast<int>(C::LDS_BYTES)));
    hipLaunchKernelGGL((ols_kernel<N>), dim3(static_cast<unsigned>(blocks)), dim3(C::THREADS), C::LDS_BYTES, st, x,
                       rows, n1, n2, seg_lo, segs, L, S, tw, out_start, out_len, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

thread_local ThreadWorkspace t_conv_ws;

}  // namespace
}  // namespace zkx

using namespace zkx;

extern "C" {

long long zkx_fftconvolve_out_len(long long n1, long long n2, int mode)
{
    if (n1 < 1 || n2 < 1) return -1;
    if (mode == 0) return n1 + n2 - 1;
    if (mode == 1) return n1;
    if (mode == 2) return (n1 > n2 ? n1 - n2 : n2 - n1) + 1;
    return -1;
}

int zkx_fftconvolve_c64(const zkx_complex64* in1, int batch, int n1, const zkx_complex64* in2, int n2, int mode,
                        zkx_complex64* out, void* stream)
{
    clear_last_error();
    hipStream_t st = (hipStream_t)stream;
    if (batch < 1 || n1 < 1 || n2 < 1) ZKX_FAIL(ZKX_INVALID_SHAPE, "fftconvolve inputs must be non-empty");
    if (mode < 0 || mode > 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve mode must be full, same or valid");
    if (!in1 || !in2 || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve: null pointer");
    if (n2 > 4096) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "fftconvolve kernel longer than 4096 samples");
    const long long full = static_cast<long long>(n1) + n2 - 1;
    const long long out_len = zkx_fftconvolve_out_len(n1, n2, mode);
    const long long out_start = mode == 0 ? 0 : mode == 1 ? (n2 - 1) / 2 : std::min(n1, n2) - 1;
    // transform length: the whole convolution when it fits, else overlap-save
    // segments of 8192 points (at least half of each segment is fresh output)
    long long N = 2;
    while (N < full && N < 8192) N *= 2;
    while (N < 2LL * n2 && N < 8192) N *= 2;
    void* ws = nullptr;
    WorkspaceLease lease_(t_conv_ws, st);
    int rc = lease_.get(sizeof(float2) * static_cast<size_t>(N), &ws);
    if (rc) return rc;
    float2* S = static_cast<float2*>(ws);
    hipLaunchKernelGGL(pad_kernel, dim3((N + 255) / 256), dim3(256), 0, st, reinterpret_cast<const float2*>(in2), n2,
                       static_cast<int>(N), S);
    ZKX_LAUNCH_CHECK();
    rc = fft_rows_dispatch(static_cast<int>(N), S, 1, false, 1.0f, S, st);
    if (rc) return rc;
    const float2* x = reinterpret_cast<const float2*>(in1);
    float2* y = reinterpret_cast<float2*>(out);
    switch (N) {
    case 2: return launch_ols<2>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 4: return launch_ols<4>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 8: return launch_ols<8>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 16: return launch_ols<16>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 32: return launch_ols<32>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 64: return launch_ols<64>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 128: return launch_ols<128>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 256: return launch_ols<256>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 512: return launch_ols<512>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 1024: return launch_ols<1024>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 2048: return launch_ols<2048>(x, batch, n1, n2, out_start, out_len, S, y, st);
    case 4096: return launch_ols<4096>(x, batch, n1, n2, out_start, out_len, S, y, st);
    default: return launch_ols<8192>(x, batch, n1, n2, out_start, out_len, S, y, st);
    }
}

}  // extern "C"
