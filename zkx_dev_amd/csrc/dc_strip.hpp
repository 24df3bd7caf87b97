// Strip-streaming Doppler FFT + |.|^2 + FP64 windowed CA-CFAR.
//
// One 256-thread block walks a strip of range bins of one dwell in chunks of
// CW = 256 / T columns (T = N / 16 threads per Doppler column).  Each chunk is
// transformed once: its |Y|^2 columns enter a ring in LDS that also carries
// the columns of the previous chunk the next outputs still need, so the CFAR
// window of every output column is complete without transforming any column
// twice (only the columns either side of a strip are read by both
// neighbours).  The samples of the next chunk are loaded into registers while
// the current chunk is transformed and its CFAR computed (barriers wait on
// LDS only, so the loads stay in flight across them).
//
// Per chunk, with the transform leading the outputs by LEAD columns (cm
// rounded up to 16, so chunk loads start on 128-byte lines) and
// CARRY = LEAD + cm:
//   ring slots [0, CARRY)           columns c_k - CARRY .. c_k - 1 (carried)
//   ring slots [CARRY, CARRY + CW)  chunk k, columns c_k .. c_k + CW - 1
//   output chunk k - 1              columns c_k - LEAD .. c_k - LEAD + CW - 1
// with c_k = S0 + LEAD - CW + k*CW for a strip starting at S0 (chunk 0 only
// fills the carry); output i's window is ring slots i .. i + 2cm.
//
// The statistic: FP64 prefix sums along range give each row's outer window
// sum H_o and side bands H_s = H_o - H_inner (formed in FP64 before rounding
// to float); FP64 prefix sums down the columns then give
//   annulus = sum over rows g0 < |df| <= rm of H_o + sum over |df| <= g0 of H_s,
// a sum of rounded positive terms in which no strong guard cell has to cancel;
// T = (alpha/N) * annulus in FP64, strict x > T, T = 0 on border cells.
// Horizontal runs cover 16 outputs aligned at multiples of 16 and vertical
// runs 16 rows aligned at multiples of 16, so the result of a cell does not
// depend on the strip or batch decomposition.
#pragma once

#include "fft_lds.hpp"
#include "radar_tiles.hpp"

namespace zkx {
namespace strip {

// Column-interleaved FFT exchange: element i of column `slot` at row
// (i ^ bit4(i)) of TC floats.  With TC = 32 a wave holds two row groups t and
// t + 1; the swizzle puts the rows they touch in one access (i and i + 16, or
// i and i + 1) at opposite parity, so the two half-waves use disjoint banks.
template <int TC>
struct ColLdsSw {
    float* re;
    float* im;
    int slot;
    __device__ __forceinline__ int a(int i) const { return (i ^ ((i >> 4) & 1)) * TC + slot; }
    __device__ __forceinline__ float2 load(int i) const
    {
        const int k = a(i);
        return make_float2(re[k], im[k]);
    }
    __device__ __forceinline__ void store(int i, float2 v) const
    {
        const int k = a(i);
        re[k] = v.x;
        im[k] = v.y;
    }
    // element base + c, where c is a multiple of 16, or base is a multiple of
    // 16 and c < 16 (every access of the FFT passes): the swizzle bit of
    // base + c is then bit 4 of base flipped by bit 4 of c, or bit 4 of base,
    // so the address is one of two per-thread bases plus the constant c*TC
    __device__ __forceinline__ int a2(int base, int c) const
    {
        const int pb = (base >> 4) & 1;
        if ((c & 15) == 0) {
            // (base + c) ^ bit == (base ^ bit) + c, bit = pb ^ bit 4 of c
            return ((base ^ (pb ^ ((c >> 4) & 1))) + c) * TC + slot;
        }
        // base % 16 == 0, c < 16: (base + c) ^ pb == base + (c ^ pb), and
        // c ^ pb is c + pb for even c, c - pb for odd c
        return (base + ((c & 1) ? -pb : pb) + c) * TC + slot;
    }
    __device__ __forceinline__ float2 load(int base, int c) const
    {
        const int k = a2(base, c);
        return make_float2(re[k], im[k]);
    }
    __device__ __forceinline__ void store(int base, int c, float2 v) const
    {
        const int k = a2(base, c);
        re[k] = v.x;
        im[k] = v.y;
    }
};

template <int N, int G0, int G1, int R0, int R1, int NT = 256>
struct StripCfg {
    static constexpr int E = fft::elems_for(N);
    static constexpr int T = N / E;
    static constexpr int THREADS = NT;
    static constexpr int CW = THREADS / T;      // columns per chunk
    static constexpr int CM = G1 + R1;          // range half-window
    static constexpr int RM = G0 + R0;          // Doppler half-window
    static constexpr int RUN = 16;              // outputs per horizontal task, rows per vertical task
    // the transform leads the outputs by LEAD >= CM columns, a multiple of 16
    // (128 B of complex64), so every chunk's loads start on a cache line
    static constexpr int LEAD = (CM + 15) / 16 * 16;
    static constexpr int CARRY = LEAD + CM;       // ring columns carried from one chunk to the next
    // Row strides are 2 mod 4 floats: every run starts on 8 bytes, so the
    // prefix passes read their values as 8-byte pairs (ds_read2_b64), and the
    // stride is odd in 8-byte units, so a wave's rows (columns) fall on
    // distinct banks.
    static constexpr int RS = ((CW + CARRY + 1) / 2 | 1) * 2;      // ring row stride
    static constexpr int HS = ((N + 2 * RM + 3) / 2 | 1) * 2;      // window-sum column stride, RM zero rows either side
    static constexpr int RING_FLOATS = N * RS;
    static constexpr int H_FLOATS = 2 * CW * HS;
    static constexpr int FFT_FLOATS = 2 * N * CW;
    static constexpr int TW_FLOATS = 3 * N;      // LDS copies of the N-point twiddle table and the window
    static constexpr int LDS_FLOATS = RING_FLOATS + (H_FLOATS > FFT_FLOATS ? H_FLOATS : FFT_FLOATS) + TW_FLOATS;
    static constexpr size_t LDS_BYTES = sizeof(float) * LDS_FLOATS;
    static_assert(E == 16 && N % RUN == 0 && CW % RUN == 0, "strip shape");
    static_assert(CW >= CARRY, "the carried columns must come from one chunk");
    static_assert(N * CW == THREADS * RUN, "one horizontal task per thread");
    static_assert(RS % 4 == 2 && HS % 4 == 2 && HS >= N + 2 * RM + 2 && (RUN + 2 * CM) % 2 == 0, "strides");
};

struct StripArgs {
    const float2* x;   // [batch][N][R] (P == N)
    int R, window;
    int SW;            // outputs per strip (multiple of CW)
    double alpha_over_n;
    const float2* tw;  // exp(-2 pi i k / N), k < N
    float* thr;        // [batch][N][R]
    unsigned char* det;
    float2* y;         // optional range-Doppler map [batch][N][R]
    float* power;      // optional |Y|^2 [batch][N][R]
};

// raw buffer access (zkx_hip_common.hpp)
using buf::kOff;
using buf::ld2;
using buf::rsrc;
using buf::st1;
using buf::st2;
using buf::stb;

// Samples of column `col` (rows t + e*T) of the dwell behind `src`; zeros
// outside [0, R).
template <int N, int E, int T, int AUX>
__device__ __forceinline__ void strip_load(__amdgpu_buffer_rsrc_t src, int R, int col, int t, float2 (&buf)[E])
{
    const unsigned voff = (col >= 0 && col < R) ? (static_cast<unsigned>(t) * R + col) * 8u : kOff;
#pragma unroll
    for (int e = 0; e < E; ++e) buf[e] = ld2<AUX>(src, voff, static_cast<unsigned>(e) * T * R * 8u);
}

// The whole strip of dwell `dwell`, outputs [S0, S1); tid is the thread's
// index in the block; LOAD_AUX the cache policy of the sample loads.  Every
// thread of the block must call this (it holds barriers).  The launcher
// guarantees N * R * 8 < 2^31 (all byte offsets fit the buffer range check).
template <int N, int G0, int G1, int R0, int R1, int NTH = 256, int LOAD_AUX = buf::kCached>
__device__ __forceinline__ void doppler_cfar_strip(const StripArgs& a, int dwell, int S0, int S1, float* lds,
                                                   int tid)
{
    using C = StripCfg<N, G0, G1, R0, R1, NTH>;
    constexpr int E = C::E, T = C::T, CW = C::CW, CM = C::CM, RM = C::RM, RS = C::RS, HS = C::HS;
    constexpr int RUN = C::RUN, NT = C::THREADS, LEAD = C::LEAD, CARRY = C::CARRY;
    const int R = a.R;
    const long long cells = static_cast<long long>(N) * R;
    const unsigned ncell = static_cast<unsigned>(cells);
    const auto src = rsrc(a.x + dwell * cells, ncell * 8u);
    const auto rthr = rsrc(a.thr + dwell * cells, ncell * 4u);
    const auto rdet = rsrc(a.det + dwell * cells, ncell);
    const bool maps = a.y != nullptr || a.power != nullptr;
    const auto ry = rsrc(a.y ? a.y + dwell * cells : nullptr, a.y ? ncell * 8u : 0u);
    const auto rpw = rsrc(a.power ? a.power + dwell * cells : nullptr, a.power ? ncell * 4u : 0u);

    // window sums first (LDS offset 0 keeps their ds_read2 offsets in range),
    // then the ring, then the twiddle and window tables
    float* hreg = lds;  // window sums; also the FFT exchange buffer
    float* ring = lds + (C::H_FLOATS > C::FFT_FLOATS ? C::H_FLOATS : C::FFT_FLOATS);
    float* ho = hreg;                    // outer horizontal sums
    float* hs = hreg + CW * HS;          // side bands: outer minus inner horizontal sums
    float2* twl = reinterpret_cast<float2*>(lds + C::LDS_FLOATS - C::TW_FLOATS);
    float* wtab = lds + C::LDS_FLOATS - N;

    const int s = tid % CW, t = tid / CW;
    // the window (built once per block in FP64 and rounded) and the twiddle
    // table, both kept in LDS for the whole strip
    for (int p = tid; p < N; p += NT) {
        wtab[p] = static_cast<float>(tiles::window_value(a.window, p, N));
        twl[p] = a.tw[p];
    }
    __syncthreads();

    const int nout = (S1 - S0 + CW - 1) / CW;
    const ColLdsSw<CW> L{hreg, hreg + N * CW, s};
    const fft::LdsTwiddles TWL{twl};

    // vertical task of this thread: column vo, rows f0 .. f0 + RUN - 1; with
    // CW = 32 the two half-waves take adjacent runs, 16 ring rows apart
    // (16 * RS = 32 mod 64 floats: their ring reads use disjoint banks)
    int vo, vrun;
    if constexpr (CW == 32) {
        vo = tid & 31;
        vrun = tid >> 5;
    } else {
        vo = tid % CW;
        vrun = tid / CW;
    }
    const int f0 = vrun * RUN;
    // horizontal task: row hf, outputs hrun*RUN ..
    const int hf = tid % N, hrun = tid / N;
    // pad rows of the window sums this thread clears each chunk: ZPT of the
    // 2*RM pad rows of one (plane, column)
    constexpr int ZTPC = NT / (2 * CW);                    // threads per (plane, column)
    constexpr int ZPT = (2 * RM + ZTPC - 1) / ZTPC;        // pad rows per thread
    const int zpc = tid / ZTPC, zpart = tid % ZTPC;
    float* zbase = hreg + zpc * HS;  // (plane, column) zpc: plane * CW * HS + column * HS
    // carried columns this thread copies: CPT of the CARRY of one row
    constexpr int CTPR = NT / N;                           // threads per ring row
    constexpr int CPT = (CARRY + CTPR - 1) / CTPR;
    const int crow = tid / CTPR, cpart = tid % CTPR;
    float* cbase = ring + crow * RS + cpart * CPT;
    static_assert(NT % (2 * CW) == 0 && NT % N == 0, "pad/carry split");

    float2 buf[E];
    strip_load<N, E, T, LOAD_AUX>(src, R, S0 + LEAD - CW + s, t, buf);
    for (int k = 0; k <= nout; ++k) {
        const int ck = S0 + LEAD - CW + k * CW;  // first column of chunk k
        float2 v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = cscale(buf[e], wtab[t + e * T]);
        // the samples are consumed before the next chunk's loads reuse their
        // registers (keeps the scheduler from hoisting the loads and copying)
        __builtin_amdgcn_sched_barrier(0);
        // (after the last chunk the column is out of range: zeros, no phi)
        strip_load<N, E, T, LOAD_AUX>(src, R, k < nout ? ck + CW + s : -1, t, buf);
        fft::transform<N, E, false>(v, L, TWL, t);
        // |Y|^2 into ring slots CARRY + s
#pragma unroll
        for (int e = 0; e < E; ++e) ring[(t + e * T) * RS + CARRY + s] = power_of(v[e]);
        if (maps) {  // optional map/power outputs for the strip's own columns
            const int col = ck + s;
            const bool own = col >= S0 && col < S1;
            const unsigned vo8 = own ? (static_cast<unsigned>(t) * R + col) * 8u : kOff;
            const unsigned vo4 = own ? (static_cast<unsigned>(t) * R + col) * 4u : kOff;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const unsigned rowoff = static_cast<unsigned>(e) * T * R;
                st2<buf::kStream>(ry, v[e], vo8, rowoff * 8u);
                st1<buf::kStream>(rpw, power_of(v[e]), vo4, rowoff * 4u);
            }
        }
        __syncthreads();
        if (k > 0) {
            // ---- horizontal FP64 window sums of output chunk k-1 ------------
            // zero the RM rows either side of every column's sums (the FFT
            // exchange above reused this region)
#pragma unroll
            for (int j = 0; j < ZPT; ++j) {
                const int z = zpart * ZPT + j;
                if (z < 2 * RM) zbase[z < RM ? z : N + z] = 0.f;
            }
            {
                // output i's window: ring values rr[i .. i + 2CM] (8-byte aligned)
                constexpr int HN = RUN + 2 * CM;
                const float2* rr = reinterpret_cast<const float2*>(ring + hf * RS + hrun * RUN);
                float2 rv[HN / 2];
#pragma unroll
                for (int j = 0; j < HN / 2; ++j) rv[j] = rr[j];
                // FP64 prefix sums of the run's window values: P[q] = sum of
                // the first q values, so every window sum is one difference
                double pre[HN + 1];
                pre[0] = 0.0;
#pragma unroll
                for (int q = 0; q < HN; ++q)
                    pre[q + 1] = pre[q] + static_cast<double>((q & 1) ? rv[q >> 1].y : rv[q >> 1].x);
                float* hoc = ho + (hrun * RUN) * HS + RM + hf;
                float* hsc = hs + (hrun * RUN) * HS + RM + hf;
#pragma unroll
                for (int i = 0; i < RUN; ++i) {
                    const double so = pre[i + 2 * CM + 1] - pre[i];
                    const double si = pre[i + CM + G1 + 1] - pre[i + CM - G1];
                    // the side bands are formed in FP64 before rounding, so a
                    // strong cell inside the guard columns cancels exactly
                    hoc[i * HS] = static_cast<float>(so);
                    hsc[i * HS] = static_cast<float>(so - si);
                }
            }
            __syncthreads();
            // ---- vertical FP64 sliding sums, threshold, detection ------------
            // annulus = (rows |df| <= RM of H_outer) - (rows |df| <= G0 of
            // H_outer) + (rows |df| <= G0 of the side bands); every rounded
            // horizontal sum that holds a guard cell enters both of the first
            // two terms and cancels.
            {
                // row f0 - RM + q of the outer sums at hq[q]; row f0 - G0 + q
                // of the side bands at hz[ZS + q] (both read as 8-byte pairs
                // from an 8-byte aligned start, ZS the skipped leading value)
                constexpr int QN = RUN + 2 * RM, ZN = RUN + 2 * G0;
                constexpr int ZS = (RM - G0) & 1, ZP = (ZS + ZN + 1) / 2;
                const float2* hoc = reinterpret_cast<const float2*>(ho + vo * HS + f0);
                const float2* hsc = reinterpret_cast<const float2*>(hs + vo * HS + f0 + RM - G0 - ZS);
                float2 hq2[QN / 2], hz2[ZP];
#pragma unroll
                for (int j = 0; j < QN / 2; ++j) hq2[j] = hoc[j];
#pragma unroll
                for (int j = 0; j < ZP; ++j) hz2[j] = hsc[j];
                const float* hq = reinterpret_cast<const float*>(hq2);
                const float* hz = reinterpret_cast<const float*>(hz2) + ZS;
                // FP64 prefix sums down the column of outer sums (Q) and of
                // side bands (Z)
                double qv[QN + 1], zv[ZN + 1];
                qv[0] = 0.0;
                zv[0] = 0.0;
#pragma unroll
                for (int q = 0; q < QN; ++q) qv[q + 1] = qv[q] + static_cast<double>(hq[q]);
#pragma unroll
                for (int q = 0; q < ZN; ++q) zv[q + 1] = zv[q] + static_cast<double>(hz[q]);
                const int c = ck - LEAD + vo;  // output column whose window is ring slots vo .. vo + 2CM
                const bool colin = c >= CM && c < R - CM;
                const bool store = c < S1 && c < R;
                const unsigned voff4 = store ? (static_cast<unsigned>(f0) * R + c) * 4u : kOff;
                const unsigned voff1 = store ? static_cast<unsigned>(f0) * R + c : kOff;
                const float* xr = ring + f0 * RS + CM + vo;
#pragma unroll
                for (int i = 0; i < RUN; ++i) {
                    // rows of H_outer at |df| in (G0, RM] (the two bands above
                    // and below the guard rows) plus the guard rows' side bands
                    const double bands = (qv[i + 2 * RM + 1] - qv[i + RM + G0 + 1]) + (qv[i + RM - G0] - qv[i]);
                    const double side = zv[i + 2 * G0 + 1] - zv[i];
                    const int f = f0 + i;
                    const bool live = colin && f >= RM && f < N - RM;
                    const double th = live ? a.alpha_over_n * (bands + side) : 0.0;
                    const float xv = xr[i * RS];
                    const unsigned row = static_cast<unsigned>(i) * R;
                    st1<buf::kStream>(rthr, static_cast<float>(th), voff4, row * 4u);
                    stb<buf::kStream>(rdet, static_cast<double>(xv) > th ? 1 : 0, voff1, row);
                }
            }
        }
        __syncthreads();
        // carry the last CARRY columns of the ring to slots [0, CARRY)
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            if (cpart * CPT + j < CARRY) cbase[j] = cbase[CW + j];
        }
    }
}

}  // namespace strip
}  // namespace zkx
