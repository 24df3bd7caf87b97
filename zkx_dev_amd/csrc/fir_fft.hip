// Long FIR filters with up = 1 (firfilter, and upfirdn decimating by D) as
// frequency-domain overlap-save on the LDS FFT:
//
//   y[m] = sum_k h[k] x[m D - k] = sum_{r<D} (h_r * x_r)[m],
//   h_r[j] = h[j D + r],  x_r[q] = x[q D - r],
//
// so an output block m0 .. m0+L-1 is IFFT_N( sum_r FFT_N(x_r segment) H_r ) / N
// with H_r = FFT_N(h_r) and segments starting Kmax-1 samples early
// (Kmax = ceil(K / D) taps per phase): D forward transforms, D spectral
// multiply-adds and one inverse transform per block, all in registers and
// LDS.  Real lines carry two consecutive blocks in one complex row (real and
// imaginary parts), which the real taps keep apart.  The direct polyphase
// kernels (polyphase.hip) remain the path for short filters, up > 1 and
// strided lines; this path turns the tap loop (K FMAs per output) into
// ~(D+1) 5 log2 N flops per output, so long filters run at HBM speed.
#include "fft_lds.hpp"
#include "runtime.hpp"

#include <algorithm>

namespace zkx {
int fft_rows_dispatch(int n, const float2* x, long long rows, bool inverse, float scale, float2* y,
                      hipStream_t st);

namespace {

using fft::RowCfg;

template <class V>
struct Lane2;  // how a line sample enters / leaves the complex row
template <>
struct Lane2<float2> {
    static constexpr int BLOCKS = 1;
};
template <>
struct Lane2<float> {
    static constexpr int BLOCKS = 2;
};

// Hs[r][f]: phase r's taps zero-padded to N (transformed afterwards in place)
__global__ void phase_taps_kernel(const float* __restrict__ h, int K, int D, int N, float2* __restrict__ Hs)
{
    const long long total = static_cast<long long>(D) * N;
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int r = static_cast<int>(g / N), j = static_cast<int>(g % N);
        const long long k = static_cast<long long>(j) * D + r;
        Hs[g] = make_float2(k < K ? h[k] : 0.f, 0.f);
    }
}

// Index arithmetic is 32-bit: the caller keeps lines * n and lines * out_n
// below 2^31.  No __restrict__ on Hs / tw: with it the scheduler hoists the
// spectrum loads above the transforms and the kernel needs 256 VGPRs.
template <int N, class V>
__global__ void __launch_bounds__(RowCfg<N>::THREADS)
fir_fft_kernel(const V* __restrict__ x, int rows, int n, int out_n, int D, int Kmax, int L, int nblk,
               const float2* Hs, const float2* tw, V* __restrict__ y)
{
    using C = RowCfg<N>;
    constexpr int BL = Lane2<V>::BLOCKS;
    extern __shared__ float lds[];
    const int slot = threadIdx.x / C::T;
    const int t = threadIdx.x % C::T;
    const int row = blockIdx.x * C::FPB + slot;
    const bool active = row < rows;
    const int rpl = (nblk + BL - 1) / BL;  // rows per line
    const int line = active ? row / rpl : 0;
    const int b0 = (active ? row - line * rpl : 0) * BL;  // first block of this row
    fft::RowLds Ls{reinterpret_cast<float2*>(lds + slot * 2 * C::SLOT_FLOATS)};
    const V* xl = x + static_cast<long long>(line) * n;
    const int base = (b0 * L - (Kmax - 1)) * D;  // input index of segment sample 0, phase 0
    float2 acc[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) acc[k] = make_float2(0.f, 0.f);
    for (int r = 0; r < D; ++r) {
        float2 v[C::E];
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            const int u = base + (t + k * C::T) * D - r;
            float2 s = make_float2(0.f, 0.f);
            if (active) {
                if constexpr (BL == 1) {
                    if (u >= 0 && u < n) s = xl[u];
                } else {
                    const int ub = u + L * D;
                    if (u >= 0 && u < n) s.x = xl[u];
                    if (b0 + 1 < nblk && ub >= 0 && ub < n) s.y = xl[ub];
                }
            }
            v[k] = s;
        }
        // an opaque copy of t per phase keeps the compiler from hoisting the
        // transform's twiddle loads out of the phase loop into registers
        int tt = t;
        asm volatile("" : "+v"(tt));
        fft::transform<N, C::E, false>(v, Ls, tw, tt);
        const float2* H = Hs + r * N;
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            const float2 hv = H[t + k * C::T];
            acc[k].x += v[k].x * hv.x - v[k].y * hv.y;
            acc[k].y += v[k].x * hv.y + v[k].y * hv.x;
        }
    }
    fft::transform<N, C::E, true>(acc, Ls, tw, t);
    if (!active) return;
    const float sc = 1.0f / static_cast<float>(N);
    V* yl = y + static_cast<long long>(line) * out_n;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int i = t + k * C::T;
        if (i < Kmax - 1) continue;
        const int m = b0 * L + i - (Kmax - 1);
        if constexpr (BL == 1) {
            if (m < out_n) yl[m] = cscale(acc[k], sc);
        } else {
            if (m < out_n) yl[m] = acc[k].x * sc;
            if (b0 + 1 < nblk && m + L < out_n) yl[m + L] = acc[k].y * sc;
        }
    }
}

thread_local ThreadWorkspace t_spec_ws;

template <int N, class V>
int launch_fir_fft(const float* h, int K, const V* x, int lines, long long n, int D, long long out_n, V* y,
                   hipStream_t st)
{
    using C = RowCfg<N>;
    constexpr int BL = Lane2<V>::BLOCKS;
    const int Kmax = (K + D - 1) / D;
    const int L = N - Kmax + 1;
    void* ws = nullptr;
    WorkspaceLease lease_(t_spec_ws, st);
    int rc = lease_.get(sizeof(float2) * static_cast<size_t>(D) * N, &ws);
    if (rc) return rc;
    float2* Hs = static_cast<float2*>(ws);
    hipLaunchKernelGGL(phase_taps_kernel, dim3(std::min<long long>((static_cast<long long>(D) * N + 255) / 256, 4096)),
                       dim3(256), 0, st, h, K, D, N, Hs);
    ZKX_LAUNCH_CHECK();
    if ((rc = fft_rows_dispatch(N, Hs, D, false, 1.0f, Hs, st))) return rc;
    const float2* tw = nullptr;
    if ((rc = twiddles(N, &tw))) return rc;
    const long long nblk = (out_n + L - 1) / L;
    const long long rows = static_cast<long long>(lines) * ((nblk + BL - 1) / BL);
    const long long blocks = (rows + C::FPB - 1) / C::FPB;
    if (rows >= (1LL << 31) || static_cast<long long>(lines) * n >= (1LL << 31) ||
        static_cast<long long>(lines) * out_n >= (1LL << 31) || (nblk * L + N) * D >= (1LL << 31))
        ZKX_FAIL(ZKX_INVALID_SHAPE, "frequency-domain FIR: line too long for 32-bit indexing");
    if (C::LDS_BYTES > 64 * 1024)
        ZKX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(fir_fft_kernel<N, V>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(C::LDS_BYTES)));
    hipLaunchKernelGGL((fir_fft_kernel<N, V>), dim3(static_cast<unsigned>(blocks)), dim3(C::THREADS), C::LDS_BYTES, st,
                       x, static_cast<int>(rows), static_cast<int>(n), static_cast<int>(out_n), D, Kmax, L,
                       static_cast<int>(nblk), Hs, tw, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // namespace

// The frequency-domain path for `lines` contiguous lines of n samples:
// y[m] = sum_k h[k] x[m D - k], m < out_n.  Sets *done = false (and does
// nothing) when the filter is too short for the path to pay or too long for
// an 8192-point segment.
template <class V>
int fir_fft_run(const float* h, int K, const V* x, int lines, long long n, int D, long long out_n, V* y,
                hipStream_t st, bool* done)
{
    *done = false;
    const int Kmax = (K + D - 1) / D;
    if (Kmax < 48 || lines < 1 || out_n < 1) return ZKX_SUCCESS;
    // segment: at most 1/8 of it spent on the history
    int N = 1024;
    while (N < 8192 && (Kmax - 1) * 8 > N) N <<= 1;
    if ((Kmax - 1) * 8 > N) return ZKX_SUCCESS;
    if (out_n < N / 2) return ZKX_SUCCESS;
    int rc;
    switch (N) {
    case 1024: rc = launch_fir_fft<1024, V>(h, K, x, lines, n, D, out_n, y, st); break;
    case 2048: rc = launch_fir_fft<2048, V>(h, K, x, lines, n, D, out_n, y, st); break;
    case 4096: rc = launch_fir_fft<4096, V>(h, K, x, lines, n, D, out_n, y, st); break;
    default: rc = launch_fir_fft<8192, V>(h, K, x, lines, n, D, out_n, y, st); break;
    }
    if (!rc) *done = true;
    return rc;
}

template int fir_fft_run<float>(const float*, int, const float*, int, long long, int, long long, float*, hipStream_t,
                                bool*);
template int fir_fft_run<float2>(const float*, int, const float2*, int, long long, int, long long, float2*, hipStream_t,
                                 bool*);

}  // namespace zkx
