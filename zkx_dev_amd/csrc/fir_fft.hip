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
extern int g_fir_fft;
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

// Segment form for D in {1, 2, 4, 8}: each row is an overlap-save segment of
// NF = 4096 contiguous input samples (real lines: two consecutive segments
// as the real and imaginary parts of one row).  One forward transform, the
// product with H = FFT_NF(h), then the spectrum folded to NI = NF/D points:
// the D aliases k' + q*NI summed give the transform of every D-th output,
//   y[D j] = (1/NF) sum_{k'<NI} (sum_{q<D} F[k' + q NI]) W_NI^{-k' j},
// so one NI-point inverse transform yields the decimated outputs.  A thread
// holds F[t + 256 k] (k < 16), whose aliases are its own elements k' + (16/D) q,
// so the fold is in registers.  Segment s starts at input s*Lm*D - KD*D
// (KD = ceil((K-1)/D) output rows of history, Lm = NI - KD outputs).
template <int D, class V>
__global__ void __launch_bounds__(256)
fir_seg_kernel(const V* __restrict__ x, int rpl, int n, int out_n, int KD, int Lm, int nseg,
               const float2* __restrict__ H, const float2* __restrict__ twF, const float2* __restrict__ twI,
               V* __restrict__ y)
{
    constexpr int NF = 4096, NI = NF / D, E = 16, EI = E / D, T = 256;
    constexpr int BL = Lane2<V>::BLOCKS;
    using C = RowCfg<NF>;
    static_assert(C::T == T && C::FPB == 1 && EI >= 2, "segment shape");
    extern __shared__ float lds[];
    const int t = threadIdx.x;
    const int row = blockIdx.x;
    const int line = row / rpl;
    const int s0 = (row - line * rpl) * BL;
    const fft::RowLds L{reinterpret_cast<float2*>(lds)};
    float2* p16 = reinterpret_cast<float2*>(lds + 2 * C::SLOT_FLOATS);
    fft::fill_pass16_twiddles<NF>(p16, twF, t, T);
    const fft::Pass16Twiddles TW{twF, p16};
    const auto rx = buf::rsrc(x + static_cast<long long>(line) * n, static_cast<unsigned>(n) * sizeof(V));
    const int b = (s0 * Lm - KD) * D;  // input index of sample 0 of segment s0
    float2 v[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int i = b + t + k * T;
        if constexpr (BL == 1) {
            v[k] = buf::ld2(rx, static_cast<unsigned>(i) < static_cast<unsigned>(n) ? i * 8u : buf::kOff, 0u);
        } else {
            const int i2 = i + Lm * D;  // the next segment
            const unsigned o1 = static_cast<unsigned>(i) < static_cast<unsigned>(n) ? i * 4u : buf::kOff;
            const unsigned o2 = static_cast<unsigned>(i2) < static_cast<unsigned>(n) ? i2 * 4u : buf::kOff;
            v[k] = make_float2(__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, o1, 0u, 0)),
                               __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, o2, 0u, 0)));
        }
    }
    fft::transform<NF, E, false>(v, L, TW, t);
    const auto rh = buf::rsrc(H, NF * 8u);
#pragma unroll
    for (int k = 0; k < E; ++k) v[k] = cmul(v[k], buf::ld2(rh, t * 8u, k * T * 8u));
    float2 z[EI];
#pragma unroll
    for (int k = 0; k < EI; ++k) {
        z[k] = v[k];
#pragma unroll
        for (int q = 1; q < D; ++q) z[k] = make_float2(z[k].x + v[k + EI * q].x, z[k].y + v[k + EI * q].y);
    }
    if constexpr (D == 1)
        fft::transform<NI, EI, true>(z, L, TW, t);
    else
        fft::transform<NI, EI, true>(z, L, twI, t);
    const float sc = 1.0f / static_cast<float>(NF);
    // outputs j = t + 256 k' >= KD of the segment; stores past the line's end
    // fall outside its buffer
    const auto ry = buf::rsrc(y + static_cast<long long>(line) * out_n, static_cast<unsigned>(out_n) * sizeof(V));
#pragma unroll
    for (int k = 0; k < EI; ++k) {
        const int j = t + k * T;
        const int m = s0 * Lm + j - KD;
        if constexpr (BL == 1) {
            buf::st2(ry, cscale(z[k], sc), j >= KD ? m * 8u : buf::kOff, 0u);
        } else {
            buf::st1(ry, z[k].x * sc, j >= KD ? m * 4u : buf::kOff, 0u);
            buf::st1(ry, z[k].y * sc, (j >= KD && s0 + 1 < nseg) ? (m + Lm) * 4u : buf::kOff, 0u);
        }
    }
}

// H = FFT_NF(h zero-padded to NF), one block.
template <int NF>
__global__ void __launch_bounds__(RowCfg<NF>::THREADS)
fir_spectrum_kernel(const float* __restrict__ h, int K, const float2* __restrict__ tw, float2* __restrict__ H)
{
    using C = RowCfg<NF>;
    static_assert(C::FPB == 1, "one row per block");
    extern __shared__ float lds[];
    const int t = threadIdx.x;
    const fft::RowLds L{reinterpret_cast<float2*>(lds)};
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int i = t + k * C::T;
        v[k] = make_float2(i < K ? h[i] : 0.f, 0.f);
    }
    fft::transform<NF, C::E, false>(v, L, tw, t);
#pragma unroll
    for (int k = 0; k < C::E; ++k) H[t + k * C::T] = v[k];
}

// Rate-changing form (up > 1, Pn = PN residues after gcd reduction): output
// n = PN*T + r is y_r[T] = sum_i g_r[i] x[D T + c - i], where g_r holds the
// taps of residue r's phase shifted so every residue reads the same input
// alignment c.  A segment's forward transform is shared by the PN residues:
// for each, the product with G_r = FFT(g_r), the fold by D and one
// (4096/D)-point inverse; residue r's outputs land at stride PN.
struct PolyTaps {
    int phase[4];  // phi_r: first tap of residue r's phase
    int shift[4];  // c - Q_r: leading zeros of g_r
};

// G[r] = FFT_NF(g_r), block r.
template <int NF>
__global__ void __launch_bounds__(RowCfg<NF>::THREADS)
fir_poly_spectra_kernel(const float* __restrict__ h, int K, int up, PolyTaps pt, const float2* __restrict__ tw,
                        float2* __restrict__ G)
{
    using C = RowCfg<NF>;
    static_assert(C::FPB == 1, "one row per block");
    extern __shared__ float lds[];
    const int t = threadIdx.x, r = blockIdx.x;
    const fft::RowLds L{reinterpret_cast<float2*>(lds)};
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int i = t + k * C::T - pt.shift[r];
        const long long tap = pt.phase[r] + static_cast<long long>(up) * i;
        v[k] = make_float2(i >= 0 && tap < K ? h[tap] : 0.f, 0.f);
    }
    fft::transform<NF, C::E, false>(v, L, tw, t);
#pragma unroll
    for (int k = 0; k < C::E; ++k) G[static_cast<long long>(r) * NF + t + k * C::T] = v[k];
}

template <int D, int PN, class V>
__global__ void __launch_bounds__(256)
fir_poly_kernel(const V* __restrict__ x, int rpl, int n, int out_n, int KD, int Lm, int nseg, int c,
                const float2* __restrict__ G, const float2* __restrict__ twF, const float2* __restrict__ twI,
                V* __restrict__ y)
{
    constexpr int NF = 4096, NI = NF / D, E = 16, EI = E / D, T = 256;
    constexpr int BL = Lane2<V>::BLOCKS;
    using C = RowCfg<NF>;
    static_assert(C::T == T && C::FPB == 1 && EI >= 2, "segment shape");
    extern __shared__ float lds[];
    const int t = threadIdx.x;
    const int row = blockIdx.x;
    const int line = row / rpl;
    const int s0 = (row - line * rpl) * BL;
    const fft::RowLds L{reinterpret_cast<float2*>(lds)};
    float2* p16 = reinterpret_cast<float2*>(lds + 2 * C::SLOT_FLOATS);
    fft::fill_pass16_twiddles<NF>(p16, twF, t, T);
    const fft::Pass16Twiddles TW{twF, p16};
    const auto rx = buf::rsrc(x + static_cast<long long>(line) * n, static_cast<unsigned>(n) * sizeof(V));
    const int b = (s0 * Lm - KD) * D + c;  // input index of sample 0 of segment s0
    float2 v[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int i = b + t + k * T;
        if constexpr (BL == 1) {
            v[k] = buf::ld2(rx, static_cast<unsigned>(i) < static_cast<unsigned>(n) ? i * 8u : buf::kOff, 0u);
        } else {
            const int i2 = i + Lm * D;  // the next segment
            const unsigned o1 = static_cast<unsigned>(i) < static_cast<unsigned>(n) ? i * 4u : buf::kOff;
            const unsigned o2 = static_cast<unsigned>(i2) < static_cast<unsigned>(n) ? i2 * 4u : buf::kOff;
            v[k] = make_float2(__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, o1, 0u, 0)),
                               __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, o2, 0u, 0)));
        }
    }
    fft::transform<NF, E, false>(v, L, TW, t);
    const float sc = 1.0f / static_cast<float>(NF);
    const auto ry = buf::rsrc(y + static_cast<long long>(line) * out_n, static_cast<unsigned>(out_n) * sizeof(V));
    const auto rg = buf::rsrc(G, PN * NF * 8u);
#pragma unroll 1
    for (int r = 0; r < PN; ++r) {
        float2 z[EI];
#pragma unroll
        for (int k = 0; k < EI; ++k) z[k] = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const float2 p = cmul(v[k], buf::ld2(rg, t * 8u, (r * NF + k * T) * 8u));
            z[k % EI] = make_float2(z[k % EI].x + p.x, z[k % EI].y + p.y);
        }
        if constexpr (D == 1)
            fft::transform<NI, EI, true>(z, L, TW, t);
        else
            fft::transform<NI, EI, true>(z, L, twI, t);
#pragma unroll
        for (int k = 0; k < EI; ++k) {
            const int j = t + k * T;
            const int o = PN * (s0 * Lm + j - KD) + r;
            if constexpr (BL == 1) {
                buf::st2(ry, cscale(z[k], sc), j >= KD ? o * 8u : buf::kOff, 0u);
            } else {
                buf::st1(ry, z[k].x * sc, j >= KD ? o * 4u : buf::kOff, 0u);
                buf::st1(ry, z[k].y * sc, (j >= KD && s0 + 1 < nseg) ? (o + PN * Lm) * 4u : buf::kOff, 0u);
            }
        }
    }
}

thread_local ThreadWorkspace t_spec_ws;

// Launchers of the segment forms return kNotTaken, before any device work,
// when a line is too long for their 32-bit offsets; the caller then takes
// the direct kernels.
constexpr int kNotTaken = -1;

template <int D, class V>
int launch_fir_seg(const float* h, int K, const V* x, int lines, long long n, long long out_n, V* y, hipStream_t st)
{
    constexpr int NF = 4096, NI = NF / D;
    constexpr int BL = Lane2<V>::BLOCKS;
    using C = RowCfg<NF>;
    const int KD = (K - 1 + D - 1) / D;
    const int Lm = NI - KD;
    const long long nseg = (out_n + Lm - 1) / Lm;
    const long long rpl = (nseg + BL - 1) / BL;
    const long long rows = static_cast<long long>(lines) * rpl;
    // 32-bit offsets within a line, row index and segment starts
    if (rows >= (1LL << 31) || n * static_cast<long long>(sizeof(V)) >= (1LL << 31) ||
        (out_n + Lm) * static_cast<long long>(sizeof(V)) >= (1LL << 31) || (nseg + 1) * Lm * D + NF >= (1LL << 31))
        return kNotTaken;
    void* ws = nullptr;
    WorkspaceLease lease_(t_spec_ws, st);
    int rc = lease_.get(sizeof(float2) * NF, &ws);
    if (rc) return rc;
    float2* Hs = static_cast<float2*>(ws);
    const float2 *twF = nullptr, *twI = nullptr;
    if ((rc = twiddles(NF, &twF)) || (rc = twiddles(NI, &twI))) return rc;
    hipLaunchKernelGGL(fir_spectrum_kernel<NF>, dim3(1), dim3(C::THREADS), C::LDS_BYTES, st, h, K, twF, Hs);
    ZKX_LAUNCH_CHECK();
    constexpr size_t lds = C::LDS_BYTES + sizeof(float2) * fft::Pass16Table<NF>::ENTRIES;  // 36.7 KB
    static_assert(lds <= 64 * 1024, "default dynamic LDS limit");
    hipLaunchKernelGGL((fir_seg_kernel<D, V>), dim3(static_cast<unsigned>(rows)), dim3(256), lds, st, x,
                       static_cast<int>(rpl), static_cast<int>(n), static_cast<int>(out_n), KD, Lm,
                       static_cast<int>(nseg), Hs, twF, twI, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

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
    // segment form: history at most 1/8 of the 4096-point segment
    const int KD = (K - 1 + D - 1) / D;
    if (g_fir_fft == 2 && (D == 1 || D == 2 || D == 4 || D == 8) && KD * D * 8 <= 4096 && out_n >= 4096 / D) {
        int rc;
        switch (D) {
        case 1: rc = launch_fir_seg<1, V>(h, K, x, lines, n, out_n, y, st); break;
        case 2: rc = launch_fir_seg<2, V>(h, K, x, lines, n, out_n, y, st); break;
        case 4: rc = launch_fir_seg<4, V>(h, K, x, lines, n, out_n, y, st); break;
        default: rc = launch_fir_seg<8, V>(h, K, x, lines, n, out_n, y, st); break;
        }
        if (rc == kNotTaken) return ZKX_SUCCESS;
        if (!rc) *done = true;
        return rc;
    }
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

namespace {

constexpr int kPolySegMinPhaseTaps = 64;
thread_local ThreadWorkspace t_poly_spec_ws;

template <int D, int PN, class V>
int launch_fir_poly(const float* h, int K, int up, const PolyTaps& pt, int KD, int c, const V* x, int lines,
                    long long n, long long out_n, V* y, hipStream_t st)
{
    constexpr int NF = 4096, NI = NF / D;
    constexpr int BL = Lane2<V>::BLOCKS;
    using C = RowCfg<NF>;
    const int Lm = NI - KD;
    const long long T_count = (out_n + PN - 1) / PN;  // outputs per residue (at most)
    const long long nseg = (T_count + Lm - 1) / Lm;
    const long long rpl = (nseg + BL - 1) / BL;
    const long long rows = static_cast<long long>(lines) * rpl;
    if (rows >= (1LL << 31) || n * static_cast<long long>(sizeof(V)) >= (1LL << 31) ||
        (out_n + PN * static_cast<long long>(Lm) + PN) * static_cast<long long>(sizeof(V)) >= (1LL << 31) ||
        (nseg + 1) * Lm * D + NF + c >= (1LL << 31))
        return kNotTaken;
    void* ws = nullptr;
    WorkspaceLease lease_(t_poly_spec_ws, st);
    int rc = lease_.get(sizeof(float2) * NF * PN, &ws);
    if (rc) return rc;
    float2* G = static_cast<float2*>(ws);
    const float2 *twF = nullptr, *twI = nullptr;
    if ((rc = twiddles(NF, &twF)) || (rc = twiddles(NI, &twI))) return rc;
    hipLaunchKernelGGL(fir_poly_spectra_kernel<NF>, dim3(PN), dim3(C::THREADS), C::LDS_BYTES, st, h, K, up, pt, twF,
                       G);
    ZKX_LAUNCH_CHECK();
    constexpr size_t lds = C::LDS_BYTES + sizeof(float2) * fft::Pass16Table<NF>::ENTRIES;
    static_assert(lds <= 64 * 1024, "default dynamic LDS limit");
    hipLaunchKernelGGL((fir_poly_kernel<D, PN, V>), dim3(static_cast<unsigned>(rows)), dim3(256), lds, st, x,
                       static_cast<int>(rpl), static_cast<int>(n), static_cast<int>(out_n), KD, Lm,
                       static_cast<int>(nseg), c, G, twF, twI, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

template <int D, class V>
int dispatch_fir_poly(int PN, const float* h, int K, int up, const PolyTaps& pt, int KD, int c, const V* x,
                      int lines, long long n, long long out_n, V* y, hipStream_t st)
{
    switch (PN) {
    case 2: return launch_fir_poly<D, 2, V>(h, K, up, pt, KD, c, x, lines, n, out_n, y, st);
    case 3: return launch_fir_poly<D, 3, V>(h, K, up, pt, KD, c, x, lines, n, out_n, y, st);
    default: return launch_fir_poly<D, 4, V>(h, K, up, pt, KD, c, x, lines, n, out_n, y, st);
    }
}

}  // namespace

// The rate-changing segment form for `lines` contiguous lines:
// y[o] = sum_k h[k] x_up[o*down + m_off - k], x_up[u] = x[u/up] when up | u.
// Sets *done = false (and does nothing) unless up/g in {2, 3, 4} and
// down/g in {1, 2, 4} (g = gcd(up, down)), the phases hold at least 64
// taps, the history fits 1/8 of a segment and the line holds at least one
// segment of outputs.
template <class V>
int fir_poly_run(const float* h, int K, int up, int down, long long m_off, const V* x, int lines, long long n,
                 long long out_n, V* y, hipStream_t st, bool* done)
{
    *done = false;
    if (g_fir_fft != 2 || lines < 1 || n < 1 || out_n < 1 || m_off < 0) return ZKX_SUCCESS;
    int g = up, b = down;
    while (b) {
        const int tmp = g % b;
        g = b;
        b = tmp;
    }
    const int PN = up / g, D = down / g;
    if (PN < 2 || PN > 4 || !(D == 1 || D == 2 || D == 4)) return ZKX_SUCCESS;
    // the direct kernels' work grows with the taps per phase, this path's
    // does not: at 41 taps per phase (resample 3/2, 121 taps) the direct
    // kernel is still faster (142 vs 154 us on 16 x 2^20 complex64)
    const int Jmax = (K + up - 1) / up;
    if (Jmax < kPolySegMinPhaseTaps) return ZKX_SUCCESS;
    PolyTaps pt{};
    long long Q[4], c = 0;
    for (int r = 0; r < PN; ++r) {
        const long long m = static_cast<long long>(r) * down + m_off;
        pt.phase[r] = static_cast<int>(m % up);
        Q[r] = m / up;
        c = std::max(c, Q[r]);
    }
    long long Kg = 1;
    for (int r = 0; r < PN; ++r) {
        const long long sh = c - Q[r];
        if (sh > 4096) return ZKX_SUCCESS;
        pt.shift[r] = static_cast<int>(sh);
        const long long taps_r = pt.phase[r] < K ? (K - pt.phase[r] + up - 1) / up : 0;
        Kg = std::max(Kg, sh + taps_r);
    }
    const long long KD = (Kg - 1 + D - 1) / D;
    if (KD * D * 8 > 4096 || c >= (1LL << 30) || out_n < static_cast<long long>(PN) * (4096 / D)) return ZKX_SUCCESS;
    int rc;
    switch (D) {
    case 1: rc = dispatch_fir_poly<1, V>(PN, h, K, up, pt, static_cast<int>(KD), static_cast<int>(c), x, lines, n,
                                         out_n, y, st); break;
    case 2: rc = dispatch_fir_poly<2, V>(PN, h, K, up, pt, static_cast<int>(KD), static_cast<int>(c), x, lines, n,
                                         out_n, y, st); break;
    default: rc = dispatch_fir_poly<4, V>(PN, h, K, up, pt, static_cast<int>(KD), static_cast<int>(c), x, lines, n,
                                          out_n, y, st); break;
    }
    if (rc == kNotTaken) return ZKX_SUCCESS;
    if (!rc) *done = true;
    return rc;
}

template int fir_poly_run<float>(const float*, int, int, int, long long, const float*, int, long long, long long,
                                 float*, hipStream_t, bool*);
template int fir_poly_run<float2>(const float*, int, int, int, long long, const float2*, int, long long, long long,
                                  float2*, hipStream_t, bool*);

template int fir_fft_run<float>(const float*, int, const float*, int, long long, int, long long, float*, hipStream_t,
                                bool*);
template int fir_fft_run<float2>(const float*, int, const float2*, int, long long, int, long long, float2*, hipStream_t,
                                 bool*);

}  // namespace zkx
