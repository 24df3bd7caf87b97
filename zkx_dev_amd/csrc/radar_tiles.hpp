// Tile bodies shared by the single-shape radar kernels (radar.hip) and the
// heterogeneous batch engine (batch.hip):
//   doppler_cfar_tile<N>  windowed N-point Doppler FFT of NCM range bins,
//                         |Y|^2 and FP64 windowed CA-CFAR of the inner bins;
//   pc_segment_row<N>     pulse compression of one N-sample segment of a
//                         pulse by FFT -> x conj(H) -> IFFT (circular when
//                         N == R, overlap-save over the periodic pulse
//                         otherwise).
// Each body is written for `nthr` cooperating threads with local index `tid`
// and its own LDS region, so several tiles of the same N can share a block;
// every thread of the block must make the call (the bodies hold barriers).
#pragma once

#include "fft_lds.hpp"

namespace zkx {
namespace tiles {

constexpr double kPi = 3.141592653589793238462643383279502884;

// Symmetric window value in FP64 (kind 1 Hann, 2 Hamming), rounded by callers
// (radartools_typed.cpp:238-251 of the reference defines the windows).
__host__ __device__ inline double window_value(int kind, int i, int len)
{
    if (kind == ZKX_WINDOW_NONE || len <= 1) return 1.0;
    const double c = cos(2.0 * kPi * static_cast<double>(i) / static_cast<double>(len - 1));
    return kind == ZKX_WINDOW_HANN ? 0.5 - 0.5 * c : 0.54 - 0.46 * c;
}

// ---------------------------------------------------------------------------
// Doppler FFT + |.|^2 + windowed CA-CFAR tile
// ---------------------------------------------------------------------------
// A tile transforms NCM consecutive range bins: its TCO = NCM - 2*cm output
// bins plus the cm bins of CFAR halo on each side (their FFTs are recomputed
// by the neighbouring tiles), so the power map lives only in LDS.  The CFAR
// covers all N Doppler rows of the tile, so it needs no row halo (rows within
// g0+r0 of the edges are border cells with T = 0).  Window sums: FP64
// horizontal sums of the outer window (H_o) and of its side bands
// (H_s = H_o - H_inner, formed in FP64) kept as float; FP64 vertical sliding
// sums give annulus = (rows |df| <= rm of H_o) - (rows |df| <= g0 of H_o) +
// (rows |df| <= g0 of H_s), so a strong guard cell's rounded sums cancel
// exactly; FP64 threshold and comparison.
template <int N>
struct FusedCfg {
    static constexpr int E = fft::elems_for(N);
    static constexpr int T = N / E;
    static constexpr int NCM = 64;  // bins transformed per tile
    static constexpr int THREADS = NCM * T;
    static constexpr int FFT_FLOATS = 2 * N * NCM + N;
    // LDS floats of one tile for a CFAR halo of cm bins
    static constexpr int floats(int cm)
    {
        const int cf = N * (NCM + 1) + 2 * (NCM - 2 * cm) * (N + 1);
        return FFT_FLOATS > cf ? FFT_FLOATS : cf;
    }
};

struct DcTile {
    const float2* src;  // the dwell, [P][R]
    int P, R, window;
    int g0, g1, r0, r1;
    double alpha_over_n;
    float* thr;           // [N][R] of this dwell (P*R < 2^31: 32-bit offsets)
    unsigned char* det;   // [N][R]
    float2* y;            // optional range-Doppler map [N][R]
    float* power;         // optional |Y|^2 [N][R]
};

// Raw (unwindowed) samples of a tile's NCM columns, zero outside the dwell:
// thread (slot s, row group t) gets rows t + k*T of column tile*TCO - cm + s.
template <int N>
__device__ __forceinline__ void dc_load(const DcTile& a, int tile, int tid, float2 (&buf)[FusedCfg<N>::E])
{
    using C = FusedCfg<N>;
    const int cm = a.g1 + a.r1;
    const int TCO = C::NCM - 2 * cm;
    const int s = tid % C::NCM, t = tid / C::NCM;
    const int col = tile * TCO - cm + s;
    const bool inb = col >= 0 && col < a.R;
    const int used = a.P < N ? a.P : N;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int p = t + k * C::T;
        buf[k] = (inb && p < used) ? a.src[p * a.R + col] : make_float2(0.f, 0.f);
    }
}

// The tile body on samples already in `buf` (from dc_load, issued before the
// window-table barrier so the loads are in flight while the table is built).
// CG0..CR1 >= 0 fix the CFAR cells at compile time (every window loop then
// unrolls with constant LDS offsets); -1 reads them from `a`.
template <int N, int CG0 = -1, int CG1 = -1, int CR0 = -1, int CR1 = -1>
__device__ __forceinline__ void doppler_cfar_tile_buf(const DcTile& a, int tile, int tid, int nthr, float* lds,
                                                      const float2* tw, const float2 (&buf)[FusedCfg<N>::E])
{
    using C = FusedCfg<N>;
    const int P = a.P, R = a.R;
    const int g0 = CG0 >= 0 ? CG0 : a.g0, g1 = CG1 >= 0 ? CG1 : a.g1;
    const int rm = g0 + (CR0 >= 0 ? CR0 : a.r0), cm = g1 + (CR1 >= 0 ? CR1 : a.r1);
    const int TCO = C::NCM - 2 * cm;
    const int c0 = tile * TCO;  // first output bin of the tile
    const int cb = c0 - cm;     // bin of slot 0
    float* wtab = lds + 2 * N * C::NCM;
    const int used = P < N ? P : N;
    // a block that runs several tiles in a row must not overwrite the window
    // table (inside the previous tile's sum buffers) while that tile's last
    // pass still reads them
    __syncthreads();
    for (int p = tid; p < used; p += nthr) wtab[p] = static_cast<float>(window_value(a.window, p, P));
    __syncthreads();
    const int s = tid % C::NCM, t = tid / C::NCM;
    const int col = cb + s;
    const bool inb = col >= 0 && col < R;
    fft::ColLds<C::NCM> L{lds, lds + N * C::NCM, s};
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int p = t + k * C::T;
        v[k] = p < used ? cscale(buf[k], wtab[p]) : make_float2(0.f, 0.f);
    }
    fft::transform<N, C::E, false>(v, L, tw, t);
    // |Y|^2 tile pw[N][NCM+1] over the FFT buffers, then the horizontal sums
    // stored column-major, ho/hs[TCO][N+1], so that every LDS stride the
    // vertical pass uses is a compile-time constant (immediate offsets).
    // Global offsets are 32-bit from the dwell's base (P*R < 2^31).
    constexpr int PS = C::NCM + 1;
    constexpr int NS = N + 1;
    float* pw = lds;
    float* ho = pw + N * PS;
    float* hs = ho + (C::NCM - 2 * cm) * NS;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int f = t + k * C::T;
        const float pv = power_of(v[k]);
        pw[f * PS + s] = pv;
        if (inb && s >= cm && s < cm + TCO) {
            const int o = f * R + col;
            if (a.y) a.y[o] = v[k];
            if (a.power) a.power[o] = pv;
        }
    }
    __syncthreads();
    // outputs per horizontal task: with the cells compiled in, a quarter of
    // the tile's outputs, so the N * 4 tasks are one per thread (THREADS = 4N)
    constexpr int RUN = (CG1 >= 0 && CR1 >= 0) ? (C::NCM - 2 * (CG1 + CR1) + 3) / 4 : 8;
    const int runs = (TCO + RUN - 1) / RUN;
    for (int task = tid; task < N * runs; task += nthr) {
        const int f = task % N, o0 = (task / N) * RUN;
        const float* pr = pw + f * PS + o0;  // slot o0 .. : window of output o is slots o .. o+2cm
        double so = 0.0, si = 0.0;
        for (int d = 0; d <= 2 * cm; ++d) so += pr[d];
        for (int d = cm - g1; d <= cm + g1; ++d) si += pr[d];
        // all window reads of the run before any store, then the stores
        float vo[RUN], vi[RUN];
#pragma unroll
        for (int j = 0; j < RUN; ++j) {
            if (j > 0) {
                so += static_cast<double>(pr[j + 2 * cm]) - static_cast<double>(pr[j - 1]);
                si += static_cast<double>(pr[j + cm + g1]) - static_cast<double>(pr[j - 1 + cm - g1]);
            }
            vo[j] = static_cast<float>(so);
            vi[j] = static_cast<float>(so - si);
        }
        float* hoc = ho + o0 * NS + f;
        float* hsc = hs + o0 * NS + f;
#pragma unroll
        for (int j = 0; j < RUN; ++j) {
            if (o0 + j < TCO) {
                hoc[j * NS] = vo[j];
                hsc[j * NS] = vi[j];
            }
        }
    }
    __syncthreads();
    // vertical FP64 sliding sums: a task is one output bin and VR consecutive
    // Doppler rows; x and the window reads are issued up front and the rows
    // are unrolled, so no row waits on its own LDS load
    constexpr int VR = N < 16 ? N : 16;
    constexpr int vruns = N / VR;
    for (int task = tid; task < TCO * vruns; task += nthr) {
        const int o = task % TCO, f0 = (task / TCO) * VR;
        const int c = c0 + o;
        if (c >= R) continue;
        float xv[VR];
        const float* pc = pw + f0 * PS + o + cm;
#pragma unroll
        for (int j = 0; j < VR; ++j) xv[j] = pc[j * PS];
        const bool colin = c >= cm && c < R - cm;
        const int fi = f0 > rm ? f0 : rm;
        const int fe = (f0 + VR) < (N - rm) ? (f0 + VR) : (N - rm);
        const float* hoc = ho + o * NS;
        const float* hsc = hs + o * NS;
        double so = 0.0, sg = 0.0, sb = 0.0;
        if (colin && fi < fe) {
            for (int d = -rm; d <= rm; ++d) so += hoc[fi + d];
            for (int d = -g0; d <= g0; ++d) {
                sg += hoc[fi + d];
                sb += hsc[fi + d];
            }
        }
        int off = f0 * R + c;
#pragma unroll
        for (int j = 0; j < VR; ++j) {
            const int f = f0 + j;
            const bool live = colin && f >= fi && f < fe;
            if (live && f > fi) {
                so += static_cast<double>(hoc[f + rm]) - static_cast<double>(hoc[f - rm - 1]);
                sg += static_cast<double>(hoc[f + g0]) - static_cast<double>(hoc[f - g0 - 1]);
                sb += static_cast<double>(hsc[f + g0]) - static_cast<double>(hsc[f - g0 - 1]);
            }
            const double T = live ? a.alpha_over_n * ((so - sg) + sb) : 0.0;
            a.thr[off] = static_cast<float>(T);
            a.det[off] = static_cast<double>(xv[j]) > T ? 1 : 0;
            off += R;
        }
    }
}

// One tile, loading its own samples.
template <int N, int CG0 = -1, int CG1 = -1, int CR0 = -1, int CR1 = -1>
__device__ __forceinline__ void doppler_cfar_tile(const DcTile& a, int tile, int tid, int nthr, float* lds,
                                                  const float2* __restrict__ tw)
{
    float2 buf[FusedCfg<N>::E];
    dc_load<N>(a, tile, tid, buf);
    doppler_cfar_tile_buf<N, CG0, CG1, CR0, CR1>(a, tile, tid, nthr, lds, tw, buf);
}

// ---------------------------------------------------------------------------
// Pulse-compression segment
// ---------------------------------------------------------------------------
// Output bins [k0, k0 + count) of the circular correlation of one pulse of
// period R (its first `valid` samples in memory, zeros after them) with a
// template of at most N - count + 1 taps:
//   y[k] = sum_t x[(k + t) mod R] conj(h[t])
// The N samples x[(k0 + i) mod R], i < N, are transformed, multiplied by
// conj(H) (H = FFT_N of the windowed, normalised template) and transformed
// back; outputs i < count are exact (an overlap-save segment).  With N == R
// and k0 == 0 this is the plain circular correlation of length R.
// `t` indexes the T = N/E threads of the row; `row_lds` is the row's region;
// tw is a twiddle source of fft::transform (the N-point table, or it with an
// LDS table for the stride-16 pass).
// (No __restrict__ on these pointers: with it the scheduler hoists the
// spectrum loads above the forward transform and the row needs ~200 VGPRs.)
template <int N, class TW>
__device__ __forceinline__ void pc_segment_row(const float2* pulse, int R, int valid, int k0, int count,
                                               bool active, const float2* H, float2* out,
                                               int t, float* row_lds, TW tw)
{
    using C = fft::RowCfg<N>;
    fft::RowLds L{reinterpret_cast<float2*>(row_lds)};
    float2 v[C::E];
    // positions k0 + t + k*T reduced mod R by one conditional subtraction per
    // step: callers guarantee k0 < R and T = N/E <= R
    int pos = k0 + t;
    pos -= pos >= R ? R : 0;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        v[k] = (active && pos < valid) ? pulse[pos] : make_float2(0.f, 0.f);
        pos += C::T;
        pos -= pos >= R ? R : 0;
    }
    fft::transform<N, C::E, false>(v, L, tw, t);
#pragma unroll
    for (int k = 0; k < C::E; ++k) v[k] = cmulc(v[k], H[t + k * C::T]);
    fft::transform<N, C::E, true>(v, L, tw, t);
    if (active) {
        const float sc = 1.0f / static_cast<float>(N);
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            const int i = t + k * C::T;
            if (i < count) out[k0 + i] = cscale(v[k], sc);
        }
    }
}

}  // namespace tiles
}  // namespace zkx
