// Doppler DFT as a dense contraction on the gfx950 matrix cores, with
// split-BF16 operands and FP32 accumulation.
//
// For each channel c and range bin n, Y[k, n] = sum_p w[p] e^{-2 pi i k p / nfft} X[p, n]
// over the first min(P, nfft) pulses (the transform pulse_doppler_device
// computes, radartools_typed.cpp:714-808 of the reference, there by FFT; any
// nfft works here, so this is also the library's path for lengths the LDS
// FFT does not take).  With complex numbers split into real rows, Y' = A X' where
//   X'[2p + j][n] = component j of X[p, n]      (j = 0 real, 1 imaginary)
//   Y'[2k + i][n] = component i of Y[k, n]
//   A[2k][2p] = Fr, A[2k][2p+1] = -Fi, A[2k+1][2p] = Fi, A[2k+1][2p+1] = Fr,
//   F = w[p] e^{-2 pi i k p / P} = Fr + i Fi,
// a real (2 nfft x 2P) by (2P x R) product per channel: 8 nfft P R flops.
//
// Split-BF16: every operand v is carried as hi = bf16(v) and
// lo = bf16(v - hi); the product is hi*hi + hi*lo + lo*hi (the lo*lo term,
// ~2^-16 relative, is dropped), each a v_mfma_f32_32x32x16_bf16 into one FP32
// accumulator.  `terms = 1` keeps only hi*hi (plain BF16) for comparison.
// A is built once per (device, nfft, P, window) in FP64 and split on the
// device (per call for explicit window values).
//
// MFMA 32x32x16 bf16 lane layout (checked on the device): lane l supplies
// row (A) / column (B) l % 32 and k = 8 (l / 32) + e for its 8 values;
// accumulator v of lane l is D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
// Rows 2k and 2k+1 (real, imaginary) therefore sit in one lane as
// accumulators v and v+1, so each lane stores whole complex values.
//
// Tiling: a 256-thread block computes 128 rows of Y' (64 Doppler bins) by
// 128 range bins; its four waves take 64 x 64 each (2 x 2 MFMA tiles).  The
// K loop walks 16 pulses (32 rows of X') at a time: A tiles come from the
// precomputed table (L2-resident), X tiles are read from HBM, split to
// hi/lo and staged in LDS; the next chunk's global loads are issued before
// the current chunk's MFMAs.  Blocks of one range tile and all Doppler tiles
// are consecutive and go to one XCD, so X is read from HBM once.
#include "runtime.hpp"

#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

namespace zkx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int LDA = BK + 8;  // bf16 per LDS row (padding staggers banks)
constexpr int kThreads = 256;
constexpr double kTwoPi = 6.283185307179586476925286766559;

// A (2 nfft x K2) as bf16 hi and lo tables, row-major: the DFT of length
// nfft over the first `used` pulses, window of length P folded in (explicit
// window values when given), columns from 2*used to K2 zero.
__global__ void dft_table_kernel(int nfft, int used, int P, int K2, int window, const float* __restrict__ ew,
                                 __bf16* __restrict__ hi, __bf16* __restrict__ lo)
{
    const long long rows = 2LL * nfft;
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < rows * K2;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int m = static_cast<int>(g / K2), kk = static_cast<int>(g % K2);
        const int k = m >> 1, i = m & 1, p = kk >> 1, j = kk & 1;
        if (p >= used) {
            hi[g] = static_cast<__bf16>(0.0f);
            lo[g] = static_cast<__bf16>(0.0f);
            continue;
        }
        double w = 1.0;
        if (ew) {
            w = static_cast<double>(ew[p]);
        } else if (window != ZKX_WINDOW_NONE && P > 1) {
            const double c = cos(kTwoPi * p / (P - 1));
            w = window == ZKX_WINDOW_HANN ? 0.5 - 0.5 * c : 0.54 - 0.46 * c;
        }
        const long long kp = (static_cast<long long>(k) * p) % nfft;
        const double th = -kTwoPi * static_cast<double>(kp) / nfft;
        const double fr = w * cos(th), fi = w * sin(th);
        const double a = i == 0 ? (j == 0 ? fr : -fi) : (j == 0 ? fi : fr);
        const __bf16 h = static_cast<__bf16>(static_cast<float>(a));
        const __bf16 l = static_cast<__bf16>(static_cast<float>(a - static_cast<double>(static_cast<float>(h))));
        hi[g] = h;
        lo[g] = l;
    }
}

template <int TERMS>
__global__ void __launch_bounds__(kThreads)
doppler_dft_kernel(const float2* __restrict__ x, int P, int used, int nfft, int K2, int R, int mtiles, int ntiles,
                   const __bf16* __restrict__ ahi, const __bf16* __restrict__ alo, float2* __restrict__ y,
                   float* __restrict__ power)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[2][BM * LDA];  // hi, lo
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][BN * LDA];  // [n][kk]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // block -> (channel, range tile, Doppler tile); the Doppler tiles of one
    // range tile are consecutive and, after the XCD remap, on one XCD
    const int total = static_cast<int>(gridDim.x);
    const int lin = blockIdx.x;
    const int full = (total / 8) * 8;
    const int logical = lin < full ? (lin % 8) * (full / 8) + lin / 8 : lin;
    const int mt = logical % mtiles;
    const int rest = logical / mtiles;
    const int nt = rest % ntiles;
    const int ch = rest / ntiles;
    const int M2 = 2 * nfft;
    const int m0 = mt * BM, n0 = nt * BN;
    const float2* xc = x + static_cast<long long>(ch) * P * R;

    // global -> register staging: A rows (128 x 32 bf16, hi and lo): 512
    // 16-byte pieces per table, 2 per thread per table; X: 128 columns x 4
    // groups of 4 pulses, 2 per thread
    uint4 ra[2][2];
    float2 rb[2][4];
    auto fetch = [&](int kk0) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int piece = tid + u * kThreads;  // 0..511
            const int r = piece >> 2, q = piece & 3;
            const int m = m0 + r;
            const bool ok = m < M2;
            const long long off = static_cast<long long>(ok ? m : 0) * K2 + kk0 + q * 8;
            ra[0][u] = ok ? *reinterpret_cast<const uint4*>(ahi + off) : make_uint4(0, 0, 0, 0);
            ra[1][u] = ok ? *reinterpret_cast<const uint4*>(alo + off) : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int task = tid + u * kThreads;  // 0..511
            const int nn = task & (BN - 1), g = task >> 7;
            const int n = n0 + nn;
            const int p0 = (kk0 >> 1) + g * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                rb[u][q] = (n < R && p0 + q < used) ? xc[static_cast<long long>(p0 + q) * R + n] : make_float2(0.f, 0.f);
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int piece = tid + u * kThreads;
            const int r = piece >> 2, q = piece & 3;
            *reinterpret_cast<uint4*>(&As[0][r * LDA + q * 8]) = ra[0][u];
            *reinterpret_cast<uint4*>(&As[1][r * LDA + q * 8]) = ra[1][u];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int task = tid + u * kThreads;
            const int nn = task & (BN - 1), g = task >> 7;
            bf16x8 h, l;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float v0 = rb[u][q].x, v1 = rb[u][q].y;
                const __bf16 h0 = static_cast<__bf16>(v0), h1 = static_cast<__bf16>(v1);
                h[2 * q] = h0;
                h[2 * q + 1] = h1;
                l[2 * q] = static_cast<__bf16>(v0 - static_cast<float>(h0));
                l[2 * q + 1] = static_cast<__bf16>(v1 - static_cast<float>(h1));
            }
            *reinterpret_cast<bf16x8*>(&Bs[0][nn * LDA + g * 8]) = h;
            *reinterpret_cast<bf16x8*>(&Bs[1][nn * LDA + g * 8]) = l;
        }
    };

    const int wm = wave >> 1, wn = wave & 1;  // wave tile: rows wm*64.., cols wn*64..
    f32x16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

    fetch(0);
    for (int kk0 = 0; kk0 < K2; kk0 += BK) {
        __syncthreads();  // previous chunk's operand reads are done
        stash();
        __syncthreads();
        if (kk0 + BK < K2) fetch(kk0 + BK);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            const int kof = ks + 8 * (lane >> 5);
            bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int r = wm * 64 + a * 32 + (lane & 31);
                ah[a] = *reinterpret_cast<const bf16x8*>(&As[0][r * LDA + kof]);
                al[a] = *reinterpret_cast<const bf16x8*>(&As[1][r * LDA + kof]);
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int c = wn * 64 + b * 32 + (lane & 31);
                bh[b] = *reinterpret_cast<const bf16x8*>(&Bs[0][c * LDA + kof]);
                bl[b] = *reinterpret_cast<const bf16x8*>(&Bs[1][c * LDA + kof]);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
                    if constexpr (TERMS == 3) {
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                    }
                }
        }
    }

    // epilogue: accumulators v, v+1 (v % 4 in {0, 2}) are rows 2k, 2k+1
    float2* yc = y ? y + static_cast<long long>(ch) * nfft * R : nullptr;
    float* pc = power ? power + static_cast<long long>(ch) * nfft * R : nullptr;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int n = n0 + wn * 64 + b * 32 + (lane & 31);
            if (n >= R) continue;
#pragma unroll
            for (int v = 0; v < 16; v += 2) {
                const int m = m0 + wm * 64 + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (m >= M2) continue;
                const int k = m >> 1;
                const float2 val = make_float2(acc[a][b][v], acc[a][b][v + 1]);
                const long long o = static_cast<long long>(k) * R + n;
                if (yc) yc[o] = val;
                if (pc) pc[o] = zkx::power_of(val);
            }
        }
}

struct Table {
    __bf16* hi = nullptr;
    __bf16* lo = nullptr;
};
std::mutex g_mu;
std::map<std::tuple<int, int, int, int, int>, Table> g_tables;

int build_table(int nfft, int used, int P, int K2, int window, const float* ew, Table* t, hipStream_t st)
{
    const size_t n = 2ull * nfft * K2;
    ZKX_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&t->hi), n * sizeof(__bf16)));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&t->lo), n * sizeof(__bf16));
    if (e != hipSuccess) {
        (void)hipFree(t->hi);
        ZKX_HIP_CHECK(e);
    }
    hipLaunchKernelGGL(dft_table_kernel, dim3(1024), dim3(256), 0, st, nfft, used, P, K2, window, ew, t->hi, t->lo);
    e = hipGetLastError();
    if (e != hipSuccess) {
        (void)hipFree(t->hi);
        (void)hipFree(t->lo);
        ZKX_HIP_CHECK(e);
    }
    return ZKX_SUCCESS;
}

// Cached per (device, nfft, used, P, window); built on the null stream and
// synchronised once.
int dft_table(int nfft, int used, int P, int K2, int window, Table* out)
{
    int dev = 0;
    ZKX_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_tuple(dev, nfft, used, P, window);
    auto it = g_tables.find(key);
    if (it != g_tables.end()) {
        *out = it->second;
        return ZKX_SUCCESS;
    }
    Table t;
    int rc = build_table(nfft, used, P, K2, window, nullptr, &t, nullptr);
    if (rc) return rc;
    const hipError_t e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        (void)hipFree(t.hi);
        (void)hipFree(t.lo);
        ZKX_HIP_CHECK(e);
    }
    g_tables[key] = t;
    *out = t;
    return ZKX_SUCCESS;
}

}  // namespace

// Doppler DFT of length nfft over channels x [P, R] (the first min(P, nfft)
// pulses, window of length P or explicit values ew) into out [channels, nfft, R].
int doppler_dft(const float2* x, int channels, int P, int R, int nfft, int window, const float* ew, int terms,
                float2* out, float* power, hipStream_t st)
{
    const int used = std::min(P, nfft);
    const int K2 = (2 * used + BK - 1) / BK * BK;
    Table t;
    int rc = ZKX_SUCCESS;
    bool owned = false;
    if (ew) {
        // explicit window values: a table for this call only (freed after the
        // stream has consumed it)
        rc = build_table(nfft, used, P, K2, window, ew, &t, st);
        owned = true;
    } else {
        rc = dft_table(nfft, used, P, K2, window, &t);
    }
    if (rc) return rc;
    const int mtiles = (2 * nfft + BM - 1) / BM;
    const int ntiles = (R + BN - 1) / BN;
    const long long blocks = static_cast<long long>(mtiles) * ntiles * channels;
    if (blocks > (1LL << 31) - 1) {
        if (owned) {
            (void)hipStreamSynchronize(st);
            (void)hipFree(t.hi);
            (void)hipFree(t.lo);
        }
        ZKX_FAIL(ZKX_INVALID_SHAPE, "doppler_dft: problem too large for one launch");
    }
    if (terms == 3)
        hipLaunchKernelGGL(doppler_dft_kernel<3>, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, st, x, P,
                           used, nfft, K2, R, mtiles, ntiles, t.hi, t.lo, out, power);
    else
        hipLaunchKernelGGL(doppler_dft_kernel<1>, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, st, x, P,
                           used, nfft, K2, R, mtiles, ntiles, t.hi, t.lo, out, power);
    hipError_t e = hipGetLastError();
    if (owned) {
        const hipError_t e2 = hipStreamSynchronize(st);
        (void)hipFree(t.hi);
        (void)hipFree(t.lo);
        if (e == hipSuccess) e = e2;
    }
    ZKX_HIP_CHECK(e);
    return ZKX_SUCCESS;
}

}  // namespace zkx

extern "C" int zkx_doppler_dft_mfma_c64(const zkx_complex64* x, int channels, int num_pulses, int range_bins,
                                        int window, int terms, zkx_complex64* out, float* power, void* stream)
{
    zkx::clear_last_error();
    if (!x || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "doppler_dft: input and output must be non-null");
    if (channels < 1 || range_bins < 1 || num_pulses < 1 || num_pulses > 4096)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "doppler_dft: pulses must be in [1, 4096]");
    if (window < ZKX_WINDOW_NONE || window > ZKX_WINDOW_HAMMING)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "doppler_dft: window must be none, hann or hamming");
    if (terms != 1 && terms != 3) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "doppler_dft: terms must be 1 or 3");
    return zkx::doppler_dft(reinterpret_cast<const float2*>(x), channels, num_pulses, range_bins, num_pulses, window,
                            nullptr, terms, reinterpret_cast<float2*>(out), power, static_cast<hipStream_t>(stream));
}
