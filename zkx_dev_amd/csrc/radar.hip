// Radar operators on gfx950: LDS Stockham FFT, fused pulse compression,
// Doppler FFT with a fused |.|^2 epilogue, CA-CFAR (bit-identical integral
// image and FP64 windowed variants), and the batched dwell pipeline.
//
// Reference behaviour followed (paths relative to the reference repository):
//   pulse_compression  ZKX/cusignal_cpp/src/operators/radartools/radartools_typed.cpp:618-712
//   pulse_doppler      radartools_typed.cpp:714-808
//   cfar_alpha         radartools_typed.cpp:592-617
//   ca_cfar            radartools_typed.cpp:148-236,465-591; radartools_kernels.cuh:23-134
//   windows            radartools_typed.cpp:238-251 (symmetric Hann/Hamming)
//   Task1 chain        ZKX/Task1/pipeline/src/steps/step{2,3,4}/step{2,3,4}.cu
#include "fft_lds.hpp"
#include "radar_tiles.hpp"
#include "dc_strip.hpp"
#include "fft_any.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <mutex>
#include <string>

namespace zkx {
namespace {

using tiles::kPi;
using tiles::window_value;

// ---------------------------------------------------------------------------
// row FFT kernels (one signal per slot, FPB slots per block)
// ---------------------------------------------------------------------------
using fft::RowCfg;

// Template spectrum H = FFT_N(h~), h~ = window*h / ||window*h|| (normalize
// over all taps) keeping the first `keep` taps, zero-padded to N.  One block,
// FPB slots but only slot 0 works.
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::THREADS)
pc_template_kernel(const float2* __restrict__ h, int taps, int keep, int normalize, int window,
                   const float* __restrict__ explicit_window, const float2* __restrict__ tw,
                   float2* __restrict__ spectrum)
{
    using C = RowCfg<N>;
    extern __shared__ float lds[];
    __shared__ double red[1024];
    const int tid = threadIdx.x;
    const bool expl = window == ZKX_WINDOW_EXPLICIT_FP32 || window == ZKX_WINDOW_EXPLICIT_FP64;
    // ||w*h||^2 over all taps in FP64
    double acc = 0.0;
    for (int i = tid; i < taps; i += blockDim.x) {
        const double w = expl ? static_cast<double>(explicit_window[i]) : window_value(window, i, taps);
        const double re = h[i].x * w, im = h[i].y * w;
        acc += re * re + im * im;
    }
    red[tid] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double scale = normalize ? 1.0 / sqrt(red[0]) : 1.0;
    const int slot = tid / C::T, t = tid % C::T;
    fft::RowLds L{reinterpret_cast<float2*>(lds + slot * 2 * C::SLOT_FLOATS)};
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int i = t + k * C::T;
        float2 val = make_float2(0.f, 0.f);
        if (slot == 0 && i < taps && i < keep) {
            const double w = expl ? static_cast<double>(explicit_window[i]) : window_value(window, i, taps);
            val = make_float2(static_cast<float>(h[i].x * w * scale), static_cast<float>(h[i].y * w * scale));
        }
        v[k] = val;
    }
    fft::transform<N, C::E, false>(v, L, tw, t);
    if (slot == 0) {
#pragma unroll
        for (int k = 0; k < C::E; ++k) spectrum[t + k * C::T] = v[k];
    }
}

// Fused pulse compression of rows: y = IFFT_N(FFT_N(x_pad) * conj(H)) / N.
// rows = total pulses (batch * P); each row reads min(S, N) samples.  The
// stride-16 pass takes its twiddles from a small LDS table after the slots
// (both transforms share it); full rows load, and every row stores, through
// raw buffer instructions with constant per-element offsets.
template <int N>
struct PcCfg {
    using C = RowCfg<N>;
    static constexpr size_t LDS_BYTES = C::LDS_BYTES + sizeof(float2) * fft::Pass16Table<N>::ENTRIES;
};

template <int N>
__global__ void __launch_bounds__(RowCfg<N>::THREADS)
pc_fused_kernel(const float2* __restrict__ x, long long rows, int S, const float2* __restrict__ H,
                const float2* __restrict__ tw, float2* __restrict__ y)
{
    using C = RowCfg<N>;
    // one row per block from N = 4096 up: the row (and the buffer bases) is
    // uniform across the block
    constexpr bool ONE = C::FPB == 1;
    extern __shared__ float lds[];
    const int slot = ONE ? 0 : threadIdx.x / C::T, t = ONE ? threadIdx.x : threadIdx.x % C::T;
    const long long row = static_cast<long long>(blockIdx.x) * C::FPB + slot;
    const bool active = row < rows;
    const fft::RowLds L{reinterpret_cast<float2*>(lds + slot * 2 * C::SLOT_FLOATS)};
    float2* p16 = reinterpret_cast<float2*>(lds + 2 * C::SLOT_FLOATS * C::FPB);
    fft::fill_pass16_twiddles<N>(p16, tw, threadIdx.x, C::THREADS);
    const fft::Pass16Twiddles TW{tw, p16};
    const float2* src = x + (active ? row : 0) * static_cast<long long>(S);
    float2 v[C::E];
    if (ONE && S >= N && active) {
        const auto rs = buf::rsrc(src, N * 8u);
#pragma unroll
        // the pulses are read once: streamed, so the caches keep the
        // compressed output a Doppler stage reads next
        for (int k = 0; k < C::E; ++k) v[k] = buf::ld2<buf::kStream>(rs, t * 8u, k * C::T * 8u);
    } else {
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            const int i = t + k * C::T;
            v[k] = (active && i < S) ? src[i] : make_float2(0.f, 0.f);
        }
    }
    fft::transform<N, C::E, false>(v, L, TW, t);
    if constexpr (ONE) {
        const auto rh = buf::rsrc(H, N * 8u);
#pragma unroll
        for (int k = 0; k < C::E; ++k) v[k] = cmulc(v[k], buf::ld2(rh, t * 8u, k * C::T * 8u));
    } else {
#pragma unroll
        for (int k = 0; k < C::E; ++k) v[k] = cmulc(v[k], H[t + k * C::T]);
    }
    fft::transform<N, C::E, true>(v, L, TW, t);
    const float s = 1.0f / static_cast<float>(N);
    if constexpr (ONE) {
        // an inactive row's stores fall outside its zero-length buffer
        const auto ry = buf::rsrc(y + (active ? row : 0) * static_cast<long long>(N), active ? N * 8u : 0u);
#pragma unroll
        for (int k = 0; k < C::E; ++k) buf::st2(ry, cscale(v[k], s), t * 8u, k * C::T * 8u);
    } else if (active) {
        float2* dst = y + row * static_cast<long long>(N);
#pragma unroll
        for (int k = 0; k < C::E; ++k) dst[t + k * C::T] = cscale(v[k], s);
    }
}

// Pulse compression of any nfft by overlap-save segments: each of the
// P * nseg rows is one N-point segment (tiles::pc_segment_row) of a pulse's
// circular correlation over period nfft (the first `valid` = min(S, nfft)
// samples of the pulse, zeros after them); L outputs per segment.
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::THREADS)
pc_segments_kernel(const float2* __restrict__ x, long long rows, int nseg, int L, int nfft, int valid, int S,
                   const float2* __restrict__ H, const float2* __restrict__ tw, float2* __restrict__ y)
{
    using C = RowCfg<N>;
    extern __shared__ float lds[];
    const int slot = threadIdx.x / C::T, t = threadIdx.x % C::T;
    // rows < 2^31 (the launcher's limit): 32-bit row arithmetic
    const int row = static_cast<int>(blockIdx.x) * C::FPB + slot;
    const bool active = row < rows;
    const int r = active ? row : 0;
    const int p = r / nseg;
    const int seg = r - p * nseg;
    const int k0 = seg * L;
    const int count = min(L, nfft - k0);
    float2* p16 = reinterpret_cast<float2*>(lds + 2 * C::SLOT_FLOATS * C::FPB);
    fft::fill_pass16_twiddles<N>(p16, tw, threadIdx.x, C::THREADS);
    tiles::pc_segment_row<N>(x + static_cast<long long>(p) * S, nfft, valid, k0, count, active, H,
                             y + static_cast<long long>(p) * nfft, t, lds + slot * 2 * C::SLOT_FLOATS,
                             fft::Pass16Twiddles{tw, p16});
}

// Persistent form for one row per block (N >= 4096) and full rows (S >= N):
// each block walks rows blockIdx.x, + gridDim.x, ... with the next row's
// samples loaded into registers while the current row is transformed, and
// the template spectrum held in registers for the whole walk, so HBM latency
// overlaps the two transforms.
template <int N, int MINW>
__global__ void __launch_bounds__(RowCfg<N>::THREADS, MINW)
pc_persistent_kernel(const float2* __restrict__ x, long long rows, int S, const float2* __restrict__ H,
                     const float2* __restrict__ tw, float2* __restrict__ y)
{
    using C = RowCfg<N>;
    static_assert(C::FPB == 1, "one row per block");
    extern __shared__ float lds[];
    const int t = threadIdx.x;
    const fft::RowLds L{reinterpret_cast<float2*>(lds)};
    float2* p16 = reinterpret_cast<float2*>(lds + 2 * C::SLOT_FLOATS);
    fft::fill_pass16_twiddles<N>(p16, tw, t, C::THREADS);
    const fft::Pass16Twiddles TW{tw, p16};
    const auto rh = buf::rsrc(H, N * 8u);
    float2 h[C::E], nxt[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) h[k] = buf::ld2(rh, t * 8u, k * C::T * 8u);
    auto fetch = [&](long long r) {
        const auto rs = buf::rsrc(x + r * static_cast<long long>(S), N * 8u);
#pragma unroll
        for (int k = 0; k < C::E; ++k) nxt[k] = buf::ld2<buf::kStream>(rs, t * 8u, k * C::T * 8u);
    };
    long long row = blockIdx.x;
    if (row < rows) fetch(row);
    for (; row < rows; row += gridDim.x) {
        float2 v[C::E];
#pragma unroll
        for (int k = 0; k < C::E; ++k) v[k] = nxt[k];
        if (row + gridDim.x < rows) fetch(row + gridDim.x);
        // opaque copy of t: keeps the compiler from hoisting every pass's
        // twiddles out of the row loop into registers
        int tt = t;
        asm volatile("" : "+v"(tt));
        fft::transform<N, C::E, false>(v, L, TW, tt);
#pragma unroll
        for (int k = 0; k < C::E; ++k) v[k] = cmulc(v[k], h[k]);
        fft::transform<N, C::E, true>(v, L, TW, tt);
        const auto ry = buf::rsrc(y + row * static_cast<long long>(N), N * 8u);
        const float s = 1.0f / static_cast<float>(N);
#pragma unroll
        for (int k = 0; k < C::E; ++k) buf::st2(ry, cscale(v[k], s), t * 8u, k * C::T * 8u);
    }
}

// Plain batched C2C row FFT (forward or inverse scaled by `scale`).
template <int N, bool INV>
__global__ void __launch_bounds__(RowCfg<N>::THREADS)
fft_rows_kernel(const float2* __restrict__ x, long long rows, const float2* __restrict__ tw,
                float scale, float2* __restrict__ y)
{
    using C = RowCfg<N>;
    extern __shared__ float lds[];
    const int slot = threadIdx.x / C::T, t = threadIdx.x % C::T;
    const long long row = static_cast<long long>(blockIdx.x) * C::FPB + slot;
    const bool active = row < rows;
    fft::RowLds L{reinterpret_cast<float2*>(lds + slot * 2 * C::SLOT_FLOATS)};
    const float2* src = x + (active ? row : 0) * static_cast<long long>(N);
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) v[k] = active ? src[t + k * C::T] : make_float2(0.f, 0.f);
    fft::transform<N, C::E, INV>(v, L, tw, t);
    if (active) {
        float2* dst = y + row * static_cast<long long>(N);
#pragma unroll
        for (int k = 0; k < C::E; ++k) dst[t + k * C::T] = cscale(v[k], scale);
    }
}

// ---------------------------------------------------------------------------
// Doppler: column FFT along pulses, TC range bins per block
// ---------------------------------------------------------------------------
template <int N, int WIDE = 0>
struct ColCfg {
    static constexpr int E = fft::elems_for(N);
    static constexpr int T = N / E;
    // range bins per block: enough for coalesced rows, at most 1024 threads
    static constexpr int TC0 = T <= 16 ? (256 / T > 64 ? 64 : 256 / T)
                                       : (T >= 512 ? 1 : (1024 / T < 16 ? 1024 / T : 16));
    // the wide variant doubles the bins per block where the block stays <= 512 threads
    static constexpr int TC = (WIDE && TC0 * 2 * T <= 512 && TC0 * 2 <= 128) ? TC0 * 2 : TC0;
    static constexpr int THREADS = T * TC;
    static constexpr size_t LDS_BYTES = sizeof(float) * (2 * N * TC + N);
};

// x [batch, P, R] -> Y [batch, N, R] (optional), |Y|^2 [batch, N, R] (optional),
// and the axis-0 running FP32 sum of |Y|^2 (optional): the first half of the
// reference CFAR's integral image (radartools_typed.cpp:501-510), fused here so
// the power map never makes an extra HBM round trip for it.
// Window over P pulses (symmetric), computed per block in FP64 into LDS.
template <int N, int WIDE>
__global__ void __launch_bounds__((ColCfg<N, WIDE>::THREADS))
doppler_kernel(const float2* __restrict__ x, int P, int R, int window,
               const float* __restrict__ explicit_window, const float2* __restrict__ tw,
               float2* __restrict__ y, float* __restrict__ power, float* __restrict__ colsum)
{
    using C = ColCfg<N, WIDE>;
    extern __shared__ float lds[];
    float* wtab = lds + 2 * N * C::TC;
    const int used = P < N ? P : N;
    const bool expl = window == ZKX_WINDOW_EXPLICIT_FP32 || window == ZKX_WINDOW_EXPLICIT_FP64;
    for (int p = threadIdx.x; p < used; p += blockDim.x)
        wtab[p] = expl ? explicit_window[p] : static_cast<float>(window_value(window, p, P));
    __syncthreads();
    const int s = threadIdx.x % C::TC, t = threadIdx.x / C::TC;
    const int col = blockIdx.x * C::TC + s;
    const long long dwell = blockIdx.y;
    const bool active = col < R;
    const float2* src = x + dwell * static_cast<long long>(P) * R;
    fft::ColLds<C::TC> L{lds, lds + N * C::TC, s};
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int p = t + k * C::T;
        v[k] = (active && p < used) ? cscale(src[static_cast<long long>(p) * R + col], wtab[p])
                                    : make_float2(0.f, 0.f);
    }
    fft::transform<N, C::E, false>(v, L, tw, t);
    const long long obase = dwell * static_cast<long long>(N) * R;
    float pw[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        pw[k] = zkx::power_of(v[k]);
        if (active) {
            const long long o = obase + static_cast<long long>(t + k * C::T) * R + col;
            if (y) y[o] = v[k];
            if (power) power[o] = pw[k];
        }
    }
    if (colsum) {
        // stage |Y|^2 as [f][slot] and let one thread per column run the
        // sequential FP32 sum down the N Doppler rows
        __syncthreads();
#pragma unroll
        for (int k = 0; k < C::E; ++k) lds[(t + k * C::T) * C::TC + s] = pw[k];
        __syncthreads();
        if (t == 0 && active) {
            constexpr int CH = N < 16 ? N : 16;
            float run = 0.0f;
            for (int f0 = 0; f0 < N; f0 += CH) {
                float v[CH];
#pragma unroll
                for (int j = 0; j < CH; ++j) v[j] = lds[(f0 + j) * C::TC + s];
#pragma unroll
                for (int j = 0; j < CH; ++j) {
                    run = __fadd_rn(run, v[j]);
                    colsum[obase + static_cast<long long>(f0 + j) * R + col] = run;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// CA-CFAR, reference-identical FP32 integral image
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

// Running FP32 sum down each column (axis 0), one thread per column.
__global__ void cfar_colscan_kernel(const float* __restrict__ x, int rows, int cols,
                                    float* __restrict__ I)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const long long base = static_cast<long long>(blockIdx.y) * rows * cols;
    float run = 0.0f;
    for (int r = 0; r < rows; ++r) {
        const long long i = base + static_cast<long long>(r) * cols + c;
        run += x[i];
        I[i] = run;
    }
}

// Running FP32 sum along each row (axis 1), in place.  64 rows per block, one
// thread per row; 64-column tiles staged through LDS so HBM traffic stays
// coalesced while every row keeps the reference's left-to-right order.
constexpr int kScanRows = 64;
constexpr int kScanCols = 64;
__global__ void __launch_bounds__(kScanRows) cfar_rowscan_kernel(float* __restrict__ I, int rows, int cols)
{
    __shared__ float tile[kScanRows][kScanCols + 1];
    const int lane = threadIdx.x;
    const int r0 = blockIdx.x * kScanRows;
    const long long base = static_cast<long long>(blockIdx.y) * rows * cols;
    float run = 0.0f;
    for (int c0 = 0; c0 < cols; c0 += kScanCols) {
        const int c = c0 + lane;
        for (int i = 0; i < kScanRows; ++i) {
            const int r = r0 + i;
            tile[i][lane] = (r < rows && c < cols) ? I[base + static_cast<long long>(r) * cols + c] : 0.0f;
        }
        __syncthreads();
        const int width = cols - c0 < kScanCols ? cols - c0 : kScanCols;
        if (width == kScanCols) {
#pragma unroll
            for (int j = 0; j < kScanCols; ++j) {
                run += tile[lane][j];
                tile[lane][j] = run;
            }
        } else {
            for (int j = 0; j < width; ++j) {
                run += tile[lane][j];
                tile[lane][j] = run;
            }
        }
        __syncthreads();
        for (int i = 0; i < kScanRows; ++i) {
            const int r = r0 + i;
            if (r < rows && c < cols) I[base + static_cast<long long>(r) * cols + c] = tile[i][lane];
        }
        __syncthreads();
    }
}

// Latency-bound scans for the reference's sequential FP32 order.  The serial
// chain is one add per element per row (or column); what costs is waiting on
// memory, so loads run SD elements ahead of the adds in straight-line code
// (unconditional, clamped addresses; the compiler then waits for exactly the
// load it needs instead of all of them).  Lanes past the edge work on the last
// column and store the same values the lane that owns it stores.
constexpr int SD = 8;

// Column prefix (axis 0) per thread; x and I must not overlap.
__global__ void __launch_bounds__(256)
cfar_colscan_stream_kernel(const float* __restrict__ x, int rows, int cols, float* __restrict__ I)
{
    const int c = min(static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), cols - 1);
    const long long base = static_cast<long long>(blockIdx.y) * rows * cols + c;
    float buf[SD];
#pragma unroll
    for (int d = 0; d < SD; ++d) buf[d] = x[base + static_cast<long long>(min(d, rows - 1)) * cols];
    const int rmain = rows - rows % SD;
    float run = 0.0f;
    for (int r0 = 0; r0 < rmain; r0 += SD) {
#pragma unroll
        for (int d = 0; d < SD; ++d) {
            run = __fadd_rn(run, buf[d]);
            I[base + static_cast<long long>(r0 + d) * cols] = run;
            buf[d] = x[base + static_cast<long long>(min(r0 + d + SD, rows - 1)) * cols];
        }
    }
#pragma unroll
    for (int d = 0; d < SD; ++d) {
        if (rmain + d < rows) {
            run = __fadd_rn(run, buf[d]);
            I[base + static_cast<long long>(rmain + d) * cols] = run;
        }
    }
}

// Row prefix (axis 1) in place with coalesced traffic.  A block owns RT_TR
// rows of one dwell and walks them in RT_TC-column chunks: its 256 threads
// move each chunk between HBM and an LDS tile with float4 accesses (RT_RD
// chunks in flight in registers), and the first RT_TR lanes of wave 0 scan the
// tile's rows left to right, one row per lane, four columns per LDS access.
// The tile row stride RT_TC+4 floats (4 banks past a multiple of the bank
// count) keeps the loaders' row-major float4 writes and the scanning lanes'
// float4 reads down the rows free of bank conflicts.  Rows past the end are
// clamped to the last row and, in the last chunk, columns past the end to the
// last float4 of the row: a clamped thread loads and stores the address its
// owner does and stores the value its owner stores, so every thread issues the
// same loads and stores in every chunk and the waits stay exact.  The columns
// past the end are scanned after the row's last value and never stored.
// Small row blocks spread one dwell's rows over many CUs; a single dwell is
// then bound by the scanning lanes (one dependent add per element, one LDS
// read and one write per four).
// Requires cols % 4 == 0.
constexpr int RT_TR = 16, RT_TC = 256, RT_TH = 256, RT_RD = 4;
typedef float rt_f4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(RT_TH)
cfar_rowscan_tile_kernel(float* I, int rows, int cols)
{
    constexpr int Q = RT_TC / 4;            // rt_f4 per tile row
    constexpr int PER = RT_TR * Q / RT_TH;  // rt_f4 per thread per chunk
    constexpr int LS = RT_TC + 4;           // tile row stride (floats)
    constexpr int G = 16;                   // rt_f4 a scanning lane holds at once
    static_assert(RT_TR * Q % RT_TH == 0 && RT_TR <= 64 && Q % G == 0, "tile shape");
    __shared__ __align__(16) float tile[2][RT_TR * LS];
    const int tid = threadIdx.x;
    float* dw = I + static_cast<long long>(blockIdx.y) * rows * cols;
    const int r0 = blockIdx.x * RT_TR;
    const int chunks = (cols + RT_TC - 1) / RT_TC;
    // element e of a chunk: tile row e / Q, rt_f4 column e % Q
    long long roff[PER];
    int q4[PER], tpos[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = tid + i * RT_TH;
        roff[i] = static_cast<long long>(min(r0 + e / Q, rows - 1)) * cols;
        q4[i] = 4 * (e % Q);
        tpos[i] = (e / Q) * LS;
    }
    auto col = [&](int i, int k) { return min(k * RT_TC + q4[i], cols - 4); };
    rt_f4 buf[RT_RD][PER];
#pragma unroll
    for (int d = 0; d < RT_RD; ++d)
#pragma unroll
        for (int i = 0; i < PER; ++i)
            buf[d][i] = *reinterpret_cast<const rt_f4*>(dw + roff[i] + col(i, min(d, chunks - 1)));
    float run = 0.0f;
    auto step = [&](rt_f4 (&cur)[PER], int k) __attribute__((always_inline)) {
        float* t = tile[k & 1];
#pragma unroll
        for (int i = 0; i < PER; ++i) *reinterpret_cast<rt_f4*>(t + tpos[i] + q4[i]) = cur[i];
        const int kn = min(k + RT_RD, chunks - 1);
#pragma unroll
        for (int i = 0; i < PER; ++i) cur[i] = *reinterpret_cast<const rt_f4*>(dw + roff[i] + col(i, kn));
        __syncthreads();
        if (tid < RT_TR) {
            // groups of G float4: the next group's reads are issued before
            // the current group's adds, so LDS latency hides behind the chain
            rt_f4* row = reinterpret_cast<rt_f4*>(t + tid * LS);
            rt_f4 v[2][G];
#pragma unroll
            for (int j = 0; j < G; ++j) v[0][j] = row[j];
#pragma unroll
            for (int g = 0; g < Q / G; ++g) {
                rt_f4 (&a)[G] = v[g & 1];
                if (g + 1 < Q / G) {
#pragma unroll
                    for (int j = 0; j < G; ++j) v[(g + 1) & 1][j] = row[(g + 1) * G + j];
                }
#pragma unroll
                for (int j = 0; j < G; ++j) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        run = __fadd_rn(run, a[j][u]);
                        a[j][u] = run;
                    }
                }
#pragma unroll
                for (int j = 0; j < G; ++j) row[g * G + j] = a[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int c = col(i, k);
            *reinterpret_cast<rt_f4*>(dw + roff[i] + c) =
                *reinterpret_cast<const rt_f4*>(t + tpos[i] + (c - k * RT_TC));
        }
    };
    // whole groups of RT_RD chunks: straight-line bodies, so every wait is exact
    const int kmain = chunks - chunks % RT_RD;
    for (int k0 = 0; k0 < kmain; k0 += RT_RD) {
#pragma unroll
        for (int d = 0; d < RT_RD; ++d) step(buf[d], k0 + d);
    }
#pragma unroll
    for (int d = 0; d < RT_RD; ++d)
        if (kmain + d < chunks) step(buf[d], kmain + d);
}

// Threshold and detection per cell from the integral image, in the reference's
// association: T = (alpha/N) * (outer - inner); borders T = 0; det = x > T.
__global__ void cfar_threshold_kernel(const float* __restrict__ x, const float* __restrict__ I,
                                      int rows, int cols, int g0, int g1, int r0, int r1,
                                      float alpha, int count, float* __restrict__ thr,
                                      unsigned char* __restrict__ det)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= cols) return;
    const long long base = static_cast<long long>(blockIdx.z) * rows * cols;
    const float* A = I + base;
    const int rm = g0 + r0, cm = g1 + r1;
    float T = 0.0f;
    if (r >= rm && r < rows - rm && c >= cm && c < cols - cm) {
        auto at = [&](int rr, int cc) { return A[static_cast<long long>(rr) * cols + cc]; };
        const int ot = r + rm, ol = r - rm - 1, orc = c + cm, olc = c - cm - 1;
        float outer = at(ot, orc);
        if (ol >= 0) outer -= at(ol, orc);
        if (olc >= 0) outer -= at(ot, olc);
        if (ol >= 0 && olc >= 0) outer += at(ol, olc);
        const float inner = at(r + g0, c + g1) - at(r - g0 - 1, c + g1) - at(r + g0, c - g1 - 1) +
                            at(r - g0 - 1, c - g1 - 1);
        T = (alpha / static_cast<float>(count)) * (outer - inner);
    }
    const long long i = base + static_cast<long long>(r) * cols + c;
    thr[i] = T;
    det[i] = (!isnan(T) && x[i] > T) ? 1 : 0;
}

// The same thresholds, TK rows of one column per thread: all integral-image
// and x loads of the TK cells are issued before any is used (clamped indices;
// border cells get T = 0 as in cfar_threshold_kernel), so each thread keeps
// 9*TK loads in flight instead of 9.
constexpr int TK = 4;
__global__ void __launch_bounds__(256)
cfar_threshold_rows_kernel(const float* __restrict__ x, const float* __restrict__ I, int rows, int cols, int g0,
                           int g1, int r0, int r1, float alpha, int count, float* __restrict__ thr,
                           unsigned char* __restrict__ det)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const int rb = blockIdx.y * TK;
    const long long base = static_cast<long long>(blockIdx.z) * rows * cols;
    const float* A = I + base;
    const int rm = g0 + r0, cm = g1 + r1;
    const bool cin = c >= cm && c < cols - cm;
    const int orc = min(c + cm, cols - 1), olc = max(c - cm - 1, 0);
    const int irc = min(c + g1, cols - 1), ilc = max(c - g1 - 1, 0);
    float v[TK][8], xv[TK];
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int r = min(rb + k, rows - 1);
        const long long ot = min(r + rm, rows - 1), ol = max(r - rm - 1, 0);
        const long long it = min(r + g0, rows - 1), il = max(r - g0 - 1, 0);
        v[k][0] = A[ot * cols + orc];
        v[k][1] = A[ol * cols + orc];
        v[k][2] = A[ot * cols + olc];
        v[k][3] = A[ol * cols + olc];
        v[k][4] = A[it * cols + irc];
        v[k][5] = A[il * cols + irc];
        v[k][6] = A[it * cols + ilc];
        v[k][7] = A[il * cols + ilc];
        xv[k] = x[base + static_cast<long long>(r) * cols + c];
    }
    const float scale = alpha / static_cast<float>(count);
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int r = rb + k;
        if (r >= rows) break;
        float T = 0.0f;
        if (cin && r >= rm && r < rows - rm) {
            float outer = v[k][0];
            if (r - rm - 1 >= 0) outer -= v[k][1];
            if (c - cm - 1 >= 0) outer -= v[k][2];
            if (r - rm - 1 >= 0 && c - cm - 1 >= 0) outer += v[k][3];
            const float inner = v[k][4] - v[k][5] - v[k][6] + v[k][7];
            T = scale * (outer - inner);
        }
        const long long o = base + static_cast<long long>(r) * cols + c;
        thr[o] = T;
        det[o] = (!isnan(T) && xv[k] > T) ? 1 : 0;
    }
}

// Row scan + threshold + detection, fused.  A block owns an output band of
// `band` rows of one dwell and scans the rows its windows need (the band plus
// g0+r0+1 rows above and g0+r0 below) left to right, one thread per row, so
// every integral-image value is the reference's sequential FP32 sum.  Columns
// stream through LDS in chunks of kBandCW: the next chunk's column prefixes
// are prefetched into registers while the current one is scanned; scanned
// values go to a ring of the last `ring` columns (a power of two covering
// kBandCW + 2*(g1+r1) + 1), from which all threads then evaluate the
// thresholds of the chunk's cells, reading x and writing T and the mask with
// coalesced accesses.
constexpr int kBandCW = 32;

constexpr int kBandMaxRows = 128;  // rows one block scans (band + halo)

template <int THREADS>
__global__ void __launch_bounds__(THREADS)
cfar_band_kernel(const float* __restrict__ x, const float* __restrict__ C, int rows, int cols, int g0, int g1,
                 int r0, int r1, float alpha, int count, int band, int ring, float* __restrict__ thr,
                 unsigned char* __restrict__ det)
{
    extern __shared__ float sm[];
    constexpr int MAXR = kBandMaxRows;
    constexpr int SW = kBandCW + 1;  // stage row stride (padded)
    const int RWS = ring + 1;        // ring row stride (padded)
    float* stage = sm;                        // [MAXR][SW]  column prefixes of the chunk
    float* irng = sm + MAXR * SW;             // [MAXR][RWS] integral-image ring
    float* xrng = irng + MAXR * RWS;          // [band][RWS] x ring of the output rows
    const int tid = threadIdx.x;
    const int rm = g0 + r0, cm = g1 + r1;
    const int ro0 = blockIdx.x * band;
    const int ro1 = min(rows, ro0 + band);
    const int BR = ro1 - ro0;
    const int sr0 = max(0, ro0 - rm - 1);
    const int sr1 = min(rows, ro1 + rm);
    const int SR = sr1 - sr0;
    const long long base = static_cast<long long>(blockIdx.y) * rows * cols;
    const float* Cb = C + base;
    const float* Xb = x + base;
    constexpr int LOADS = MAXR * kBandCW / THREADS;
    float pc[LOADS], px[LOADS];
    auto prefetch = [&](int c0) {
#pragma unroll
        for (int i = 0; i < LOADS; ++i) {
            const int e = tid + i * THREADS;
            const int rr = e / kBandCW, cc = e % kBandCW;
            const bool inb = c0 + cc < cols;
            pc[i] = (rr < SR && inb) ? Cb[static_cast<long long>(sr0 + rr) * cols + c0 + cc] : 0.0f;
            px[i] = (rr < BR && inb) ? Xb[static_cast<long long>(ro0 + rr) * cols + c0 + cc] : 0.0f;
        }
    };
    auto I = [&](int r, int c) { return irng[(r - sr0) * RWS + (c & (ring - 1))]; };
    auto X = [&](int r, int c) { return xrng[(r - ro0) * RWS + (c & (ring - 1))]; };
    const float scale = __fdiv_rn(alpha, static_cast<float>(count));
    auto cell = [&](int r, int c) {
        float T = 0.0f;
        if (r >= rm && r < rows - rm && c >= cm && c < cols - cm) {
            const int ot = r + rm, ol = r - rm - 1, orc = c + cm, olc = c - cm - 1;
            float outer = I(ot, orc);
            if (ol >= 0) outer = __fsub_rn(outer, I(ol, orc));
            if (olc >= 0) outer = __fsub_rn(outer, I(ot, olc));
            if (ol >= 0 && olc >= 0) outer = __fadd_rn(outer, I(ol, olc));
            float inner = __fsub_rn(I(r + g0, c + g1), I(r - g0 - 1, c + g1));
            inner = __fsub_rn(inner, I(r + g0, c - g1 - 1));
            inner = __fadd_rn(inner, I(r - g0 - 1, c - g1 - 1));
            T = __fmul_rn(scale, __fsub_rn(outer, inner));
        }
        const long long o = base + static_cast<long long>(r) * cols + c;
        thr[o] = T;
        det[o] = (!isnan(T) && X(r, c) > T) ? 1 : 0;
    };
    float run = 0.0f;
    prefetch(0);
    for (int c0 = 0; c0 < cols; c0 += kBandCW) {
#pragma unroll
        for (int i = 0; i < LOADS; ++i) {
            const int e = tid + i * THREADS;
            const int rr = e / kBandCW, cc = e % kBandCW;
            if (rr < SR) stage[rr * SW + cc] = pc[i];
            if (rr < BR) xrng[rr * RWS + ((c0 + cc) & (ring - 1))] = px[i];
        }
        __syncthreads();
        if (c0 + kBandCW < cols) prefetch(c0 + kBandCW);
        const int width = min(kBandCW, cols - c0);
        if (tid < SR) {
            const float* st = stage + tid * SW;
            float* rw = irng + tid * RWS;
            if (width == kBandCW) {
                // all reads first, then the dependent adds, then the writes
                float v[kBandCW];
#pragma unroll
                for (int j = 0; j < kBandCW; ++j) v[j] = st[j];
#pragma unroll
                for (int j = 0; j < kBandCW; ++j) {
                    run = __fadd_rn(run, v[j]);
                    v[j] = run;
                }
#pragma unroll
                for (int j = 0; j < kBandCW; ++j) rw[(c0 + j) & (ring - 1)] = v[j];
            } else {
                for (int j = 0; j < width; ++j) {
                    run = __fadd_rn(run, st[j]);
                    rw[(c0 + j) & (ring - 1)] = run;
                }
            }
        }
        __syncthreads();
        // cells of columns [c0 - cm, c0 + width - cm): every window they need is scanned
        const int chi = c0 + width - cm;
        for (int q = tid; q < BR * kBandCW; q += THREADS) {
            const int rr = q / kBandCW, cc = q % kBandCW;
            const int c = c0 - cm + cc;
            if (c >= 0 && c < chi) cell(ro0 + rr, c);
        }
        if (c0 + width >= cols) {
            // right border columns [max(0, chi), cols): threshold 0
            const int t0 = max(0, chi);
            const int tw = cols - t0;
            for (int q = tid; q < BR * tw; q += THREADS) cell(ro0 + q / tw, t0 + q % tw);
        }
    }
}

// Rank 1: sequential FP32 prefix sum (one thread), then per-cell threshold.
__global__ void cfar1d_scan_kernel(const float* __restrict__ x, int n, float* __restrict__ I)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float run = 0.0f;
    for (int i = 0; i < n; ++i) {
        run += x[i];
        I[i] = run;
    }
}

__global__ void cfar1d_threshold_kernel(const float* __restrict__ x, const float* __restrict__ I, int n,
                                        int g, int ref, float alpha, int count, float* __restrict__ thr,
                                        unsigned char* __restrict__ det)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = g + ref;
    float T = 0.0f;
    if (i >= m && i < n - m) {
        const int orr = i + m, ol = i - m - 1, ir = i + g, il = i - g - 1;
        const float outer = ol >= 0 ? I[orr] - I[ol] : I[orr];
        const float inner = I[ir] - I[il];
        T = (alpha / static_cast<float>(count)) * (outer - inner);
    }
    thr[i] = T;
    det[i] = (!isnan(T) && x[i] > T) ? 1 : 0;
}

#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// CA-CFAR with FP64 windowed sums (2-D prefix sums in LDS)
// ---------------------------------------------------------------------------
// Block tile: TR x TCW output cells; the LDS holds the tile plus a halo of
// (g+r) cells on every side as an FP64 2-D prefix-sum table.
constexpr int kWinTileR = 32;
constexpr int kWinTileC = 64;

// Separable window sums: a haloed tile of x in LDS; horizontal sums of the
// outer (2*cm+1) and inner (2*g1+1) column windows per halo row, accumulated
// in FP64 and kept as float (relative error <= 2^-24); then FP64 vertical
// sums of those over the outer (2*rm+1) and inner (2*g0+1) row windows.  Each
// thread slides its window along 8 consecutive outputs.  The FP64 threshold
// T = (alpha/N)(outer - inner) is compared with x in FP64.
constexpr int kWinRun = 8;

__global__ void __launch_bounds__(256)
cfar_window_kernel(const float* __restrict__ x, int rows, int cols, int g0, int g1, int r0, int r1,
                   double alpha_over_n, float* __restrict__ thr, unsigned char* __restrict__ det)
{
    extern __shared__ float sw[];
    const int rm = g0 + r0, cm = g1 + r1;
    const int HR = kWinTileR + 2 * rm, HC = kWinTileC + 2 * cm;
    const int XS = HC | 1;                   // odd row strides keep lanes on distinct banks
    constexpr int HS = kWinTileC + 1;
    float* xs = sw;                          // [HR][XS]
    float* ho = xs + HR * XS;                // [HR][HS] outer horizontal sums
    float* hi = ho + HR * HS;                // [HR][HS] inner horizontal sums
    const int tr0 = blockIdx.y * kWinTileR, tc0 = blockIdx.x * kWinTileC;
    const long long base = static_cast<long long>(blockIdx.z) * rows * cols;
    const float* X = x + base;
    {
        // rows of the haloed tile, 64 columns per wave pass, 8 loads in flight
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
        const int segs = (HC + 63) / 64;
        const int items = HR * segs;  // (row, 64-column segment) pairs
        for (int b0 = wave; b0 < items; b0 += 8 * nw) {
            float tmp[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int it = b0 + u * nw;
                const int hr = it / segs, hc = (it % segs) * 64 + lane;
                const int r = tr0 - rm + hr, c = tc0 - cm + hc;
                tmp[u] = (it < items && hc < HC && r >= 0 && r < rows && c >= 0 && c < cols)
                             ? X[static_cast<long long>(r) * cols + c] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int it = b0 + u * nw;
                const int hr = it / segs, hc = (it % segs) * 64 + lane;
                if (it < items && hc < HC) xs[hr * XS + hc] = tmp[u];
            }
        }
    }
    __syncthreads();
    constexpr int CPR = kWinTileC / kWinRun;  // column runs per row
    for (int task = threadIdx.x; task < HR * CPR; task += blockDim.x) {
        const int hr = task % HR, c0 = (task / HR) * kWinRun;
        const float* xr = xs + hr * XS + c0;
        double so = 0.0, si = 0.0;
        for (int d = 0; d <= 2 * cm; ++d) so += xr[d];
        for (int d = cm - g1; d <= cm + g1; ++d) si += xr[d];
        ho[hr * HS + c0] = static_cast<float>(so);
        hi[hr * HS + c0] = static_cast<float>(si);
#pragma unroll
        for (int j = 1; j < kWinRun; ++j) {
            so += static_cast<double>(xr[j + 2 * cm]) - static_cast<double>(xr[j - 1]);
            si += static_cast<double>(xr[j + cm + g1]) - static_cast<double>(xr[j - 1 + cm - g1]);
            ho[hr * HS + c0 + j] = static_cast<float>(so);
            hi[hr * HS + c0 + j] = static_cast<float>(si);
        }
    }
    __syncthreads();
    constexpr int RPC = kWinTileR / kWinRun;  // row runs per column
    for (int task = threadIdx.x; task < kWinTileC * RPC; task += blockDim.x) {
        const int lc = task % kWinTileC, lr0 = (task / kWinTileC) * kWinRun;
        const int c = tc0 + lc;
        double so = 0.0, si = 0.0;
        for (int d = 0; d <= 2 * rm; ++d) so += ho[(lr0 + d) * HS + lc];
        for (int d = rm - g0; d <= rm + g0; ++d) si += hi[(lr0 + d) * HS + lc];
#pragma unroll
        for (int j = 0; j < kWinRun; ++j) {
            if (j > 0) {
                so += static_cast<double>(ho[(lr0 + j + 2 * rm) * HS + lc]) -
                      static_cast<double>(ho[(lr0 + j - 1) * HS + lc]);
                si += static_cast<double>(hi[(lr0 + j + rm + g0) * HS + lc]) -
                      static_cast<double>(hi[(lr0 + j - 1 + rm - g0) * HS + lc]);
            }
            const int r = tr0 + lr0 + j;
            if (r >= rows || c >= cols) continue;
            double T = 0.0;
            if (r >= rm && r < rows - rm && c >= cm && c < cols - cm) T = alpha_over_n * (so - si);
            const long long o = base + static_cast<long long>(r) * cols + c;
            thr[o] = static_cast<float>(T);
            det[o] = static_cast<double>(xs[(lr0 + j + rm) * XS + lc + cm]) > T ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------------------
// Doppler FFT + |.|^2 + windowed CA-CFAR in one kernel (tile body in
// radar_tiles.hpp)
// ---------------------------------------------------------------------------
using tiles::FusedCfg;

template <int N, int CG0 = -1, int CG1 = -1, int CR0 = -1, int CR1 = -1>
__global__ void __launch_bounds__((FusedCfg<N>::THREADS))
doppler_cfar_kernel(const float2* __restrict__ x, int P, int R, int window, const float2* __restrict__ tw, int g0,
                    int g1, int r0, int r1, double alpha_over_n, float* __restrict__ thr,
                    unsigned char* __restrict__ det, float2* __restrict__ y, float* __restrict__ power)
{
    extern __shared__ float lds[];
    // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so hand
    // each XCD a contiguous run of tiles; neighbouring tiles then share their
    // overlapping halo columns in one L2.
    const int nx = gridDim.x;
    const int total = nx * gridDim.y;
    const int lin = blockIdx.y * nx + blockIdx.x;
    const int full = (total / 8) * 8;
    const int logical = lin < full ? (lin % 8) * (full / 8) + lin / 8 : lin;
    const int tile = logical % nx;
    const long long dwell = logical / nx;
    const long long obase = dwell * static_cast<long long>(N) * R;
    tiles::DcTile a;
    a.src = x + dwell * static_cast<long long>(P) * R;
    a.P = P;
    a.R = R;
    a.window = window;
    a.g0 = g0;
    a.g1 = g1;
    a.r0 = r0;
    a.r1 = r1;
    a.alpha_over_n = alpha_over_n;
    a.thr = thr + obase;
    a.det = det + obase;
    a.y = y ? y + obase : nullptr;
    a.power = power ? power + obase : nullptr;
    tiles::doppler_cfar_tile<N, CG0, CG1, CR0, CR1>(a, tile, threadIdx.x, blockDim.x, lds, tw);
}

// Strip-streaming form (dc_strip.hpp): block (strip, dwell) walks the strip's
// range bins in chunks, transforming every column once.
template <int N, int G0, int G1, int R0, int R1>
__global__ void __launch_bounds__(256, 2) doppler_cfar_strip_kernel(strip::StripArgs a)
{
    extern __shared__ float lds[];
    const int S0 = blockIdx.x * a.SW;
    const int S1 = min(S0 + a.SW, a.R);
    strip::doppler_cfar_strip<N, G0, G1, R0, R1>(a, blockIdx.y, S0, S1, lds, threadIdx.x);
}

// ---------------------------------------------------------------------------
// Whole dwell pipeline in one persistent launch (pulse compression, then the
// strip Doppler+CFAR of each dwell as soon as its pulses are compressed).
// Blocks take jobs from one queue in a fixed order: the PC jobs of dwell d
// (`rows` pulses each) and, `lead` dwells behind, the strip jobs of dwell
// d - lead.  A strip job waits until every PC job of its dwell
// has finished (a per-dwell counter); every job waits only on jobs that were
// taken from the queue before it by running blocks, so the queue always
// drains.  The compressed pulses pass from the PC jobs to the strip jobs
// through memory with system-coherent stores and loads (no cache holds a
// stale copy, whichever CU or XCD runs either side), and are read back
// within microseconds of being written, while they are still in the
// memory-side cache.
// ---------------------------------------------------------------------------
int g_fuse_chunks = 8;             // chunks per strip job
constexpr int kFuseSpinLimit = 1 << 20;  // polls of a dependency before giving up

struct FuseArgs {
    const float2* x;       // [batch][P][R]
    const float2* H;       // template spectrum (conjugated in the product)
    const float2* tw_r;    // R-point twiddles
    int batch, R, pc_jobs, strips, SW;
    int rows, lead;        // pulses per PC job; dwells by which PC runs ahead of the strips
    int job_slot;          // LDS float index of the block's current job number
    strip::StripArgs dc;   // x = the compressed cube [batch][P][R]
    int* queue;            // job queues, error flag, per-dwell PC counts (kFuseLine ints apart)
};

// Queue layout (ints): queue q (one per XCD) at q * kFuseLine; the error flag
// at 8 * kFuseLine; dwell d's count of finished PC jobs at (9 + d) * kFuseLine
// (one 256-byte line each: atomics on one line serialize at the memory side).
constexpr int kFuseLine = 64;
constexpr int kFuseQueues = 8;

template <int NR, int NP>
__global__ void __launch_bounds__(256, 2) dwell_fused_kernel(FuseArgs a)
{
    using C = RowCfg<NR>;
    static_assert(C::FPB == 1 && C::THREADS == 256, "one 256-thread row per block");
    // (no static LDS: it would sit before the dynamic region and leave that
    // region only 4-byte aligned, and the transforms' 8-byte LDS accesses need
    // 8-byte alignment)
    extern __shared__ float lds[];
    int& s_job = *reinterpret_cast<int*>(lds + a.job_slot);
    const int tid = threadIdx.x;
    // blocks are dealt to the XCDs round-robin; the blocks of XCD q take the
    // dwells d = q, q + 8, ... from queue q, so the PC jobs and strips of a
    // dwell run on one XCD and hand over through its L2 and the memory-side cache
    const int q = static_cast<int>(blockIdx.x) % kFuseQueues;
    const int nd = a.batch > q ? (a.batch - q + kFuseQueues - 1) / kFuseQueues : 0;  // dwells of this queue
    const int per = a.pc_jobs + a.strips;  // jobs per dwell slot
    const int total = (nd + a.lead) * per;
    int* const qc = a.queue + q * kFuseLine;
    int* const err = a.queue + kFuseQueues * kFuseLine;
    int* const done = a.queue + (kFuseQueues + 1) * kFuseLine;
    for (;;) {
        __syncthreads();  // the previous job's LDS readers are done
        if (tid == 0) s_job = atomicAdd(qc, 1);
        __syncthreads();
        const int j = s_job;
        if (j >= total) break;
        const int g = j / per, r = j - g * per;
        if (r < a.pc_jobs) {
            // ---- pulse compression of pulses [r*rows, (r+1)*rows) of local dwell g
            if (g >= nd) continue;
            const int dw = q + g * kFuseQueues;
            const fft::RowLds L{reinterpret_cast<float2*>(lds)};
            float2* p16 = reinterpret_cast<float2*>(lds + 2 * C::SLOT_FLOATS);
            fft::fill_pass16_twiddles<NR>(p16, a.tw_r, tid, C::THREADS);
            const fft::Pass16Twiddles TW{a.tw_r, p16};
            const auto rh = buf::rsrc(a.H, NR * 8u);
            float2 h[C::E];
#pragma unroll
            for (int k = 0; k < C::E; ++k) h[k] = buf::ld2(rh, tid * 8u, k * C::T * 8u);
            const long long row0 = static_cast<long long>(dw) * NP + r * a.rows;
            // the next row's samples load while the current row is transformed
            float2 nxt[C::E];
            {
                const auto rs = buf::rsrc(a.x + row0 * a.R, NR * 8u);
#pragma unroll
                for (int k = 0; k < C::E; ++k) nxt[k] = buf::ld2(rs, tid * 8u, k * C::T * 8u);
            }
            for (int qq = 0; qq < a.rows; ++qq) {
                const long long row = row0 + qq;
                float2 v[C::E];
#pragma unroll
                for (int k = 0; k < C::E; ++k) v[k] = nxt[k];
                if (qq + 1 < a.rows) {
                    const auto rs = buf::rsrc(a.x + (row + 1) * a.R, NR * 8u);
#pragma unroll
                    for (int k = 0; k < C::E; ++k) nxt[k] = buf::ld2(rs, tid * 8u, k * C::T * 8u);
                }
                int tt = tid;
                asm volatile("" : "+v"(tt));
                fft::transform<NR, C::E, false>(v, L, TW, tt);
#pragma unroll
                for (int k = 0; k < C::E; ++k) v[k] = cmulc(v[k], h[k]);
                fft::transform<NR, C::E, true>(v, L, TW, tt);
                const auto ry = buf::rsrc(a.dc.x + row * NR, NR * 8u);
                const float sc = 1.0f / static_cast<float>(NR);
#pragma unroll
                for (int k = 0; k < C::E; ++k) buf::st2<buf::kSystem>(ry, cscale(v[k], sc), tid * 8u, k * C::T * 8u);
            }
            // every store of this job has reached memory before the count moves
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            if (tid == 0)
                __hip_atomic_fetch_add(done + dw * kFuseLine, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            // ---- strip r - pc_jobs of local dwell g - lead
            const int gl = g - a.lead;
            if (gl < 0) continue;
            const int dw = q + gl * kFuseQueues;
            if (tid == 0) {
                int spins = 0;
                while (__hip_atomic_load(done + dw * kFuseLine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                       a.pc_jobs) {
                    __builtin_amdgcn_s_sleep(2);
                    if (++spins > kFuseSpinLimit) {
                        // bounded wait: flag the failure and go on (the kernel always ends)
                        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            __syncthreads();
            const int S0 = (r - a.pc_jobs) * a.SW;
            strip::doppler_cfar_strip<NP, 2, 2, 8, 8, 256, buf::kSystem>(a.dc, dw, S0, min(S0 + a.SW, a.R), lds,
                                                                          tid);
        }
    }
}

__global__ void zero_ints_kernel(int* p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Doppler transform of any nfft (not a power of two, or beyond the column
// kernels): corner turn into [R][nfft] rows, any-length row FFT, corner turn
// back.  FP32 throughout, like the power-of-two path.
// ---------------------------------------------------------------------------
// wtab[p] = window value of pulse p (p < used)
__global__ void doppler_window_kernel(int used, int P, int window, const float* __restrict__ explicit_window,
                                      float* __restrict__ wtab)
{
    const bool expl = window == ZKX_WINDOW_EXPLICIT_FP32 || window == ZKX_WINDOW_EXPLICIT_FP64;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < used; p += gridDim.x * blockDim.x)
        wtab[p] = expl ? explicit_window[p] : static_cast<float>(window_value(window, p, P));
}

// rows[r][f] = w[f] * x[f][r] for f < used, 0 for used <= f < nfft (32x32 tiles)
__global__ void __launch_bounds__(256) doppler_gather_kernel(const float2* __restrict__ x, int R, int used, int nfft,
                                                             const float* __restrict__ wtab, float2* __restrict__ rows)
{
    __shared__ float2 tile[32][33];
    const int r0 = blockIdx.x * 32, f0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        const int f = f0 + j, r = r0 + tx;
        float2 v = make_float2(0.f, 0.f);
        if (f < used && r < R) v = cscale(x[static_cast<long long>(f) * R + r], wtab[f]);
        tile[j][tx] = v;
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int r = r0 + j, f = f0 + tx;
        if (r < R && f < nfft) rows[static_cast<long long>(r) * nfft + f] = tile[tx][j];
    }
}

// y[f][r] = rows[r][f], power[f][r] = |.|^2 (either output may be NULL)
__global__ void __launch_bounds__(256) doppler_scatter_kernel(const float2* __restrict__ rows, int R, int nfft,
                                                              float2* __restrict__ y, float* __restrict__ power)
{
    __shared__ float2 tile[32][33];
    const int f0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        const int r = r0 + j, f = f0 + tx;
        if (r < R && f < nfft) tile[j][tx] = rows[static_cast<long long>(r) * nfft + f];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int f = f0 + j, r = r0 + tx;
        if (f < nfft && r < R) {
            const float2 v = tile[tx][j];
            const long long o = static_cast<long long>(f) * R + r;
            if (y) y[o] = v;
            if (power) power[o] = power_of(v);
        }
    }
}

thread_local ThreadWorkspace t_doppler_ws;

int doppler_any(const float2* x, int P, int R, int nfft, int window, const float* explicit_window, float2* y,
                float* power, hipStream_t st)
{
    if (!fft_any_length_ok(nfft)) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "pulse_doppler: nfft beyond the FFT range");
    // the corner turns put nfft / 32 and R / 32 tiles on a grid's y axis
    if ((R + 31) / 32 > 65535 || (nfft + 31) / 32 > 65535)
        ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "pulse_doppler: map too large for the corner turn");
    const int used = P < nfft ? P : nfft;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t wbytes = al(sizeof(float) * used);
    const size_t rbytes = al(sizeof(float2) * static_cast<size_t>(R) * nfft);
    const size_t fbytes = fft_any_workspace_bytes(R, nfft);
    int rc = fft_any_prepare(nfft);
    if (rc) return rc;
    void* ws = nullptr;
    WorkspaceLease lease_(t_doppler_ws, st);
    if ((rc = lease_.get(wbytes + rbytes + fbytes, &ws))) return rc;
    char* base = static_cast<char*>(ws);
    float* wtab = reinterpret_cast<float*>(base);
    float2* rows = reinterpret_cast<float2*>(base + wbytes);
    hipLaunchKernelGGL(doppler_window_kernel, dim3((used + 255) / 256), dim3(256), 0, st, used, P, window,
                       explicit_window, wtab);
    ZKX_LAUNCH_CHECK();
    hipLaunchKernelGGL(doppler_gather_kernel, dim3((R + 31) / 32, (nfft + 31) / 32), dim3(256), 0, st, x, R, used,
                       nfft, wtab, rows);
    ZKX_LAUNCH_CHECK();
    rc = fft_any_rows(rows, R, nfft, false, 1.0f, rows, fbytes ? base + wbytes + rbytes : nullptr, st);
    if (rc) return rc;
    hipLaunchKernelGGL(doppler_scatter_kernel, dim3((nfft + 31) / 32, (R + 31) / 32), dim3(256), 0, st, rows, R, nfft,
                       y, power);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

// ---------------------------------------------------------------------------
// dispatch helpers
// ---------------------------------------------------------------------------
constexpr int kMaxLdsFft = 8192;

// Kernels needing more than 64 KiB of dynamic LDS must opt in once.
template <class K>
int allow_lds(K kernel, size_t bytes)
{
    if (bytes > 64 * 1024)
        ZKX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    return ZKX_SUCCESS;
}

template <int N>
int launch_fft_rows(const float2* x, long long rows, bool inverse, float scale, float2* y, hipStream_t st)
{
    using C = RowCfg<N>;
    const float2* tw = nullptr;
    int rc = twiddles(N, &tw);
    if (rc) return rc;
    const long long blocks = (rows + C::FPB - 1) / C::FPB;
    if ((rc = allow_lds(fft_rows_kernel<N, true>, C::LDS_BYTES)) || (rc = allow_lds(fft_rows_kernel<N, false>, C::LDS_BYTES)))
        return rc;
    if (inverse)
        hipLaunchKernelGGL((fft_rows_kernel<N, true>), dim3(static_cast<unsigned>(blocks)), dim3(C::THREADS),
                           C::LDS_BYTES, st, x, rows, tw, scale, y);
    else
        hipLaunchKernelGGL((fft_rows_kernel<N, false>), dim3(static_cast<unsigned>(blocks)), dim3(C::THREADS),
                           C::LDS_BYTES, st, x, rows, tw, scale, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

template <int N>
int launch_pc_template(const float2* h, int taps, int keep, int normalize, int window, const float* ew, float2* H,
                       hipStream_t st)
{
    using C = RowCfg<N>;
    const float2* tw = nullptr;
    int rc = twiddles(N, &tw);
    if (rc) return rc;
    if ((rc = allow_lds(pc_template_kernel<N>, C::LDS_BYTES))) return rc;
    hipLaunchKernelGGL((pc_template_kernel<N>), dim3(1), dim3(C::THREADS), C::LDS_BYTES, st, h, taps, keep,
                       normalize, window, ew, tw, H);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

template <int N>
int launch_pc_segments(const float2* x, long long rows, int nseg, int L, int nfft, int valid, int S,
                       const float2* H, float2* y, hipStream_t st)
{
    using C = RowCfg<N>;
    const float2* tw = nullptr;
    int rc = twiddles(N, &tw);
    if (rc) return rc;
    if ((rc = allow_lds(pc_segments_kernel<N>, PcCfg<N>::LDS_BYTES))) return rc;
    if (rows >= (1LL << 31) - C::FPB) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "pulse compression: too many segments");
    const long long blocks = (rows + C::FPB - 1) / C::FPB;
    hipLaunchKernelGGL((pc_segments_kernel<N>), dim3(static_cast<unsigned>(blocks)), dim3(C::THREADS),
                       PcCfg<N>::LDS_BYTES, st, x, rows, nseg, L, nfft, valid, S, H, tw, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

// Kernel variant per operation, chosen by the autotuner (zkx_set_kernel_variant).
int g_pc_variant = 0;

// Host-forced transform-stage failure for error-path tests: while set, the
// transform stage of pulse_compression / pulse_doppler reports BACKEND_ERROR
// before any device work, as a failed FFT execution does in the reference
// (its synthesized backend status, operator_abnormal_validation.h:93-104).
std::atomic<int> g_inject_fft_failure{0};

int injected_fft_failure(const char* op)
{
    if (!g_inject_fft_failure.load(std::memory_order_relaxed)) return ZKX_SUCCESS;
    const std::string msg = std::string(op) + ": FFT execution failed (injected backend status)";
    ZKX_FAIL(ZKX_BACKEND_ERROR, msg.c_str());
}

template <int N, int MINW>
int launch_pc_persistent(const float2* x, long long rows, int S, const float2* H, const float2* tw, float2* y,
                         int per_cu, hipStream_t st)
{
    if constexpr (RowCfg<N>::FPB != 1) {
        return ZKX_INTERNAL_ERROR;
    } else {
        int rc = allow_lds(pc_persistent_kernel<N, MINW>, PcCfg<N>::LDS_BYTES);
        if (rc) return rc;
        int dev = 0, cus = 256;
        ZKX_HIP_CHECK(hipGetDevice(&dev));
        ZKX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const long long blocks = std::min<long long>(rows, static_cast<long long>(cus) * per_cu);
        hipLaunchKernelGGL((pc_persistent_kernel<N, MINW>), dim3(static_cast<unsigned>(blocks)),
                           dim3(RowCfg<N>::THREADS), PcCfg<N>::LDS_BYTES, st, x, rows, S, H, tw, y);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    }
}

template <int N>
int launch_pc_fused(const float2* x, long long rows, int S, const float2* H, float2* y, hipStream_t st)
{
    using C = RowCfg<N>;
    const float2* tw = nullptr;
    int rc = twiddles(N, &tw);
    if (rc) return rc;
    if constexpr (C::FPB == 1) {
        // persistent forms (autotuner variants) for full rows
        if (S >= N && g_pc_variant == 1) return launch_pc_persistent<N, 4>(x, rows, S, H, tw, y, 4, st);
        if (S >= N && g_pc_variant == 2) return launch_pc_persistent<N, 3>(x, rows, S, H, tw, y, 3, st);
        if (S >= N && g_pc_variant == 3) return launch_pc_persistent<N, 2>(x, rows, S, H, tw, y, 2, st);
    }
    const long long blocks = (rows + C::FPB - 1) / C::FPB;
    if ((rc = allow_lds(pc_fused_kernel<N>, PcCfg<N>::LDS_BYTES))) return rc;
    hipLaunchKernelGGL((pc_fused_kernel<N>), dim3(static_cast<unsigned>(blocks)), dim3(C::THREADS),
                       PcCfg<N>::LDS_BYTES, st, x, rows, S, H, tw, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

int g_doppler_variant = 0;

template <int N, int WIDE>
int launch_doppler_v(const float2* x, int batch, int P, int R, int window, const float* ew, float2* y,
                     float* power, float* colsum, const float2* tw, hipStream_t st)
{
    using C = ColCfg<N, WIDE>;
    dim3 grid((R + C::TC - 1) / C::TC, batch);
    int rc = allow_lds(doppler_kernel<N, WIDE>, C::LDS_BYTES);
    if (rc) return rc;
    hipLaunchKernelGGL((doppler_kernel<N, WIDE>), grid, dim3(C::THREADS), C::LDS_BYTES, st, x, P, R, window, ew,
                       tw, y, power, colsum);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

template <int N>
int launch_doppler(const float2* x, int batch, int P, int R, int window, const float* ew, float2* y,
                   float* power, float* colsum, hipStream_t st)
{
    const float2* tw = nullptr;
    int rc = twiddles(N, &tw);
    if (rc) return rc;
    if (g_doppler_variant == 1) return launch_doppler_v<N, 1>(x, batch, P, R, window, ew, y, power, colsum, tw, st);
    return launch_doppler_v<N, 0>(x, batch, P, R, window, ew, y, power, colsum, tw, st);
}

#define ZKX_POW2_CASES(F, ...)                     \
    switch (n) {                                   \
    case 2: return F<2>(__VA_ARGS__);              \
    case 4: return F<4>(__VA_ARGS__);              \
    case 8: return F<8>(__VA_ARGS__);              \
    case 16: return F<16>(__VA_ARGS__);            \
    case 32: return F<32>(__VA_ARGS__);            \
    case 64: return F<64>(__VA_ARGS__);            \
    case 128: return F<128>(__VA_ARGS__);          \
    case 256: return F<256>(__VA_ARGS__);          \
    case 512: return F<512>(__VA_ARGS__);          \
    case 1024: return F<1024>(__VA_ARGS__);        \
    case 2048: return F<2048>(__VA_ARGS__);        \
    case 4096: return F<4096>(__VA_ARGS__);        \
    case 8192: return F<8192>(__VA_ARGS__);        \
    default: break;                                \
    }

int fft_rows(int n, const float2* x, long long rows, bool inverse, float scale, float2* y, hipStream_t st)
{
    ZKX_POW2_CASES(launch_fft_rows, x, rows, inverse, scale, y, st)
    ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "FFT length must be a power of two in [2, 8192]");
}

int pc_template(int n, const float2* h, int taps, int normalize, int window, const float* ew, float2* H,
                hipStream_t st, int keep = -1)
{
    if (keep < 0) keep = taps;
    ZKX_POW2_CASES(launch_pc_template, h, taps, keep, normalize, window, ew, H, st)
    ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "pulse_compression nfft must be a power of two in [2, 8192]");
}

int pc_segments(int n, const float2* x, long long rows, int nseg, int L, int nfft, int valid, int S,
                const float2* H, float2* y, hipStream_t st)
{
    ZKX_POW2_CASES(launch_pc_segments, x, rows, nseg, L, nfft, valid, S, H, y, st)
    ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "segment FFT length must be a power of two in [16, 8192]");
}

// Segment FFT length for a circular correlation of period nfft with T taps:
// N > T, N/16 <= nfft (one-subtraction wrap), N <= 8192, cheapest by
// segments * N * (log2 N + 2).  Returns 0 if none fits.
int pc_segment_length(int nfft, int T, int* L_out, int* nseg_out)
{
    int best = 0;
    double best_cost = 0.0;
    for (int nl = 4; nl <= 13; ++nl) {
        const int n = 1 << nl;
        if (n <= T || n / 16 > nfft) continue;
        const int L = std::min(n - T + 1, nfft);
        const int ns = (nfft + L - 1) / L;
        const double cost = static_cast<double>(ns) * n * (nl + 2);
        if (!best || cost < best_cost) {
            best = n;
            best_cost = cost;
            *L_out = L;
            *nseg_out = ns;
        }
    }
    return best;
}

int pc_fused(int n, const float2* x, long long rows, int S, const float2* H, float2* y, hipStream_t st)
{
    ZKX_POW2_CASES(launch_pc_fused, x, rows, S, H, y, st)
    ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "pulse_compression nfft must be a power of two in [2, 8192]");
}

int doppler(int n, const float2* x, int batch, int P, int R, int window, const float* ew, float2* y,
            float* power, float* colsum, hipStream_t st)
{
    ZKX_POW2_CASES(launch_doppler, x, batch, P, R, window, ew, y, power, colsum, st)
    ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "pulse_doppler nfft must be a power of two in [2, 8192]");
}

bool valid_window(int w) { return w >= ZKX_WINDOW_NONE && w <= ZKX_WINDOW_EXPLICIT_FP64; }
bool explicit_kind(int w) { return w == ZKX_WINDOW_EXPLICIT_FP32 || w == ZKX_WINDOW_EXPLICIT_FP64; }

float cfar_alpha_float(int count, double pfa)
{
    const double n = static_cast<double>(count);
    return static_cast<float>(n * (std::pow(pfa, -1.0 / n) - 1.0));
}

thread_local ThreadWorkspace t_cfar_ws;

// Reference-identical 2-D CFAR.  colsum: the axis-0 FP32 running sums of x
// when the producer already made them (the fused Doppler epilogue), else NULL
// and they are built here into I.
int cfar_faithful_2d(const float* x, const float* colsum, int batch, int rows, int cols, int g0, int g1, int r0,
                     int r1, double pfa, float* thr, unsigned char* det, float* I, hipStream_t st)
{
    const int count = 2 * r0 * (2 * r1 + 2 * g1 + 1) + 2 * (2 * g0 + 1) * r1;
    const float alpha = cfar_alpha_float(count, pfa);
    if (cols % 4 == 0 && (colsum == nullptr || colsum == I)) {
        // column prefix (unless the producer made it), row prefix in place,
        // then every cell's threshold from the finished integral image
        if (colsum == nullptr) {
            hipLaunchKernelGGL(cfar_colscan_stream_kernel, dim3((cols + 255) / 256, batch), dim3(256), 0, st, x,
                               rows, cols, I);
            ZKX_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(cfar_rowscan_tile_kernel, dim3((rows + RT_TR - 1) / RT_TR, batch), dim3(RT_TH), 0, st, I,
                           rows, cols);
        ZKX_LAUNCH_CHECK();
        hipLaunchKernelGGL(cfar_threshold_rows_kernel, dim3((cols + 255) / 256, (rows + TK - 1) / TK, batch),
                           dim3(256), 0, st, x, I, rows, cols, g0, g1, r0, r1, alpha, count, thr, det);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    }
    if (!colsum) {
        hipLaunchKernelGGL(cfar_colscan_kernel, dim3((cols + 255) / 256, batch), dim3(256), 0, st, x, rows, cols, I);
        ZKX_LAUNCH_CHECK();
        colsum = I;
    }
    const int rm = g0 + r0, cm = g1 + r1;
    int ring = 32;
    while (ring < kBandCW + 2 * cm + 1) ring *= 2;
    const int halo = 2 * rm + 1;
    // band rows: whole dwells when there are enough dwells to fill the chip,
    // else narrower bands (each re-scans its halo rows) for more blocks
    auto pick = [&](int threads, int& band) {
        if (rows <= threads) {
            band = rows;
            const int want_blocks = 256;
            while (band > 32 && static_cast<long long>(batch) * ((rows + band - 1) / band) < want_blocks &&
                   (band / 2) + halo <= threads)
                band /= 2;
            return true;
        }
        band = threads - halo;
        return band >= 16;
    };
    int band = 0;
    if (ring <= 128 && pick(kBandMaxRows, band)) {
        constexpr int TH = 256;
        const size_t lds = sizeof(float) * (kBandMaxRows * (kBandCW + 1 + ring + 1) +
                                            static_cast<size_t>(band) * (ring + 1));
        int rc = allow_lds(cfar_band_kernel<TH>, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(cfar_band_kernel<TH>, dim3((rows + band - 1) / band, batch), dim3(TH), lds, st, x,
                           colsum, rows, cols, g0, g1, r0, r1, alpha, count, band, ring, thr, det);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    }
    // very wide windows: separate row scan and threshold passes
    hipLaunchKernelGGL(cfar_rowscan_kernel, dim3((rows + kScanRows - 1) / kScanRows, batch), dim3(kScanRows), 0,
                       st, I, rows, cols);
    ZKX_LAUNCH_CHECK();
    hipLaunchKernelGGL(cfar_threshold_kernel, dim3((cols + 255) / 256, rows, batch), dim3(256), 0, st, x, I, rows,
                       cols, g0, g1, r0, r1, alpha, count, thr, det);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

int cfar_window_2d(const float* x, int batch, int rows, int cols, int g0, int g1, int r0, int r1, double pfa,
                   float* thr, unsigned char* det, hipStream_t st)
{
    const int count = 2 * r0 * (2 * r1 + 2 * g1 + 1) + 2 * (2 * g0 + 1) * r1;
    const double n = static_cast<double>(count);
    const double alpha = n * (std::pow(pfa, -1.0 / n) - 1.0);
    const int rm = g0 + r0, cm = g1 + r1;
    const size_t lds = sizeof(float) * static_cast<size_t>(kWinTileR + 2 * rm) *
                       (((kWinTileC + 2 * cm) | 1) + 2 * (kWinTileC + 1));
    if (lds > 160 * 1024) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "ca_cfar window too large for the LDS tile");
    int rc = allow_lds(cfar_window_kernel, lds);
    if (rc) return rc;
    dim3 grid((cols + kWinTileC - 1) / kWinTileC, (rows + kWinTileR - 1) / kWinTileR, batch);
    hipLaunchKernelGGL(cfar_window_kernel, grid, dim3(256), lds, st, x, rows, cols, g0, g1, r0, r1, alpha / n, thr,
                       det);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

// Doppler+CFAR form: 0 strip-streaming (configured cells, 32 <= N <= 128),
// 1 one tile of 64 transformed bins per block; and the strip kernel's target
// block count per launch (strips per dwell = ceil(target / batch)): 512
// fills the 256 CUs at two blocks each once, with 8 strips per dwell at
// batch 64 (6% of each strip's columns transformed twice as halo).
int g_dc_variant = 0;
int g_dc_strip_blocks = 512;

template <int N>
int launch_doppler_cfar_strip(const float2* x, int batch, int R, int window, double alpha_over_n,
                              float* thr, unsigned char* det, float2* y, float* power, const float2* tw,
                              hipStream_t st)
{
    using C = strip::StripCfg<N, 2, 2, 8, 8>;
    int rc = allow_lds(doppler_cfar_strip_kernel<N, 2, 2, 8, 8>, C::LDS_BYTES);
    if (rc) return rc;
    const int chunks = (R + C::CW - 1) / C::CW;
    int strips = (g_dc_strip_blocks + batch - 1) / batch;
    strips = std::max(1, std::min(strips, chunks));
    const int sw = ((chunks + strips - 1) / strips) * C::CW;
    strips = (R + sw - 1) / sw;
    strip::StripArgs a;
    a.x = x;
    a.R = R;
    a.window = window;
    a.SW = sw;
    a.alpha_over_n = alpha_over_n;
    a.tw = tw;
    a.thr = thr;
    a.det = det;
    a.y = y;
    a.power = power;
    hipLaunchKernelGGL((doppler_cfar_strip_kernel<N, 2, 2, 8, 8>), dim3(strips, batch), dim3(C::THREADS),
                       C::LDS_BYTES, st, a);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

template <int N>
int launch_doppler_cfar(const float2* x, int batch, int P, int R, int window, int g0, int g1, int r0, int r1,
                        double pfa, float* thr, unsigned char* det, float2* y, float* power, hipStream_t st,
                        bool* done)
{
    using C = FusedCfg<N>;
    *done = false;
    const int cm = g1 + r1;
    if (C::NCM - 2 * cm < 8) return ZKX_SUCCESS;  // halo too wide for the tile: caller falls back
    if (static_cast<long long>(N) * R >= (1LL << 31)) return ZKX_SUCCESS;  // tile offsets are 32-bit
    const int TCO = C::NCM - 2 * cm;
    const size_t lds = sizeof(float) * static_cast<size_t>(C::floats(cm));
    if (lds > 160 * 1024) return ZKX_SUCCESS;
    const float2* tw = nullptr;
    int rc = twiddles(N, &tw);
    if (rc) return rc;
    // the configured window (guard 2x2, reference 8x8) runs a kernel with the
    // cell counts compiled in; any other window the runtime-cell kernel
    const bool fixed = g0 == 2 && g1 == 2 && r0 == 8 && r1 == 8;
    if ((rc = fixed ? allow_lds(doppler_cfar_kernel<N, 2, 2, 8, 8>, lds) : allow_lds(doppler_cfar_kernel<N>, lds)))
        return rc;
    const int count = 2 * r0 * (2 * r1 + 2 * g1 + 1) + 2 * (2 * g0 + 1) * r1;
    const double n = static_cast<double>(count);
    const double alpha = n * (std::pow(pfa, -1.0 / n) - 1.0);
    if constexpr (N >= 32) {
        // (buffer byte offsets of a dwell must stay below 2^31)
        if (fixed && g_dc_variant == 0 && P == N && static_cast<long long>(N) * R * 8 < (1LL << 31)) {
            rc = launch_doppler_cfar_strip<N>(x, batch, R, window, alpha / n, thr, det, y, power, tw, st);
            if (rc == ZKX_SUCCESS) *done = true;
            return rc;
        }
    }
    dim3 grid((R + TCO - 1) / TCO, batch);
    if (fixed)
        hipLaunchKernelGGL((doppler_cfar_kernel<N, 2, 2, 8, 8>), grid, dim3(C::THREADS), lds, st, x, P, R, window, tw,
                           g0, g1, r0, r1, alpha / n, thr, det, y, power);
    else
        hipLaunchKernelGGL((doppler_cfar_kernel<N>), grid, dim3(C::THREADS), lds, st, x, P, R, window, tw, g0, g1, r0,
                           r1, alpha / n, thr, det, y, power);
    ZKX_LAUNCH_CHECK();
    *done = true;
    return ZKX_SUCCESS;
}

int doppler_cfar(int n, const float2* x, int batch, int P, int R, int window, int g0, int g1, int r0, int r1,
                 double pfa, float* thr, unsigned char* det, float2* y, float* power, hipStream_t st, bool* done)
{
    *done = false;
    switch (n) {
    case 16: return launch_doppler_cfar<16>(x, batch, P, R, window, g0, g1, r0, r1, pfa, thr, det, y, power, st, done);
    case 32: return launch_doppler_cfar<32>(x, batch, P, R, window, g0, g1, r0, r1, pfa, thr, det, y, power, st, done);
    case 64: return launch_doppler_cfar<64>(x, batch, P, R, window, g0, g1, r0, r1, pfa, thr, det, y, power, st, done);
    case 128: return launch_doppler_cfar<128>(x, batch, P, R, window, g0, g1, r0, r1, pfa, thr, det, y, power, st, done);
    default: return ZKX_SUCCESS;
    }
}

// The one-launch pipeline (dwell_fused_kernel) for full-range dwells of
// R = 4096 samples (one 256-thread row), P = 32, 64 or 128 pulses, no range-Doppler map
// and the configured window (guard 2x2, reference 8x8); *done stays false
// otherwise and the caller runs the two-launch form.
// 0: one launch where it applies; 1 (default): two launches.  The one-launch
// form gives the same bits but measured slower on 64 dwells of 128 x 4096
// (243 us against 232 us for the two launches): its PC jobs run at the strip
// kernel's occupancy (two blocks per CU) and every job pays a queue round trip.
int g_fuse_variant = 1;
int g_fuse_rows = 4;     // pulses per PC job
int g_fuse_lead = 2;     // dwells of one XCD queue by which its PC jobs run ahead of its strip jobs

template <int NR, int NP>
int launch_dwell_fused(const float2* x, int batch, const float2* H, int window, double alpha_over_n, float* thr,
                       unsigned char* det, float2* pc, int* queue, hipStream_t st)
{
    using SC = strip::StripCfg<NP, 2, 2, 8, 8>;
    const size_t body = (std::max(PcCfg<NR>::LDS_BYTES, SC::LDS_BYTES) + 15) / 16 * 16;
    const size_t lds = body + 16;  // + the job number
    int rc = allow_lds(dwell_fused_kernel<NR, NP>, lds);
    if (rc) return rc;
    const float2 *tw_r = nullptr, *tw_p = nullptr;
    if ((rc = twiddles(NR, &tw_r)) || (rc = twiddles(NP, &tw_p))) return rc;
    int dev = 0, cus = 256;
    ZKX_HIP_CHECK(hipGetDevice(&dev));
    ZKX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    FuseArgs a;
    a.x = x;
    a.H = H;
    a.tw_r = tw_r;
    a.batch = batch;
    a.R = NR;
    a.rows = std::max(1, std::min(NP, g_fuse_rows));
    while (NP % a.rows) --a.rows;
    a.lead = std::max(0, g_fuse_lead);
    a.pc_jobs = NP / a.rows;
    a.SW = std::max(1, g_fuse_chunks) * SC::CW;
    a.strips = (NR + a.SW - 1) / a.SW;
    a.dc.x = pc;
    a.dc.R = NR;
    a.dc.window = window;
    a.dc.SW = a.SW;
    a.dc.alpha_over_n = alpha_over_n;
    a.dc.tw = tw_p;
    a.dc.thr = thr;
    a.dc.det = det;
    a.dc.y = nullptr;
    a.dc.power = nullptr;
    a.queue = queue;
    a.job_slot = static_cast<int>(body / sizeof(float));
    // the counters start at zero (a fill kernel, so the sequence can be captured)
    hipLaunchKernelGGL(zero_ints_kernel, dim3(16), dim3(256), 0, st, queue, (kFuseQueues + 1 + batch) * kFuseLine);
    ZKX_LAUNCH_CHECK();
    hipLaunchKernelGGL((dwell_fused_kernel<NR, NP>), dim3(2 * cus), dim3(256), lds, st, a);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

int dwell_fused(const float2* x, int batch, int P, int R, const float2* H, int window, int g0, int g1, int r0,
                int r1, double pfa, float* thr, unsigned char* det, float2* rd_map, float2* pc, int* queue,
                hipStream_t st, bool* done)
{
    *done = false;
    const bool fixed = g0 == 2 && g1 == 2 && r0 == 8 && r1 == 8;
    if (!fixed || rd_map || R != 4096 || (P != 32 && P != 64 && P != 128))
        return ZKX_SUCCESS;
    if (static_cast<long long>(P) * R * 8 >= (1LL << 31)) return ZKX_SUCCESS;
    const int count = 2 * r0 * (2 * r1 + 2 * g1 + 1) + 2 * (2 * g0 + 1) * r1;
    const double n = static_cast<double>(count);
    const double alpha = n * (std::pow(pfa, -1.0 / n) - 1.0);
    int rc = ZKX_SUCCESS;
#define ZKX_FUSE_CASE(NR_, NP_) \
    if (R == NR_ && P == NP_) rc = launch_dwell_fused<NR_, NP_>(x, batch, H, window, alpha / n, thr, det, pc, queue, st);
    ZKX_FUSE_CASE(4096, 32)
    ZKX_FUSE_CASE(4096, 64)
    ZKX_FUSE_CASE(4096, 128)
#undef ZKX_FUSE_CASE
    if (rc == ZKX_SUCCESS) *done = true;
    return rc;
}

int validate_cfar2d(int rows, int cols, int g0, int g1, int r0, int r1, double pfa)
{
    if (g0 < 0 || g1 < 0 || r0 < 1 || r1 < 1)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar guard/reference cells are invalid");
    if (rows < 0 || cols < 0) ZKX_FAIL(ZKX_INVALID_SHAPE, "ca_cfar shape dimensions must be nonnegative");
    if (static_cast<long long>(rows) <= 2LL * (g0 + r0) || static_cast<long long>(cols) <= 2LL * (g1 + r1))
        ZKX_FAIL(ZKX_INVALID_SHAPE, "ca_cfar input dimension is too small");
    if (!std::isfinite(pfa) || pfa <= 0.0 || pfa > 1.0) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar pfa must be in (0,1]");
    return ZKX_SUCCESS;
}

}  // namespace

// exported for other translation units of the library
int doppler_dft(const float2* x, int channels, int P, int R, int nfft, int window, const float* ew, int terms,
                float2* out, float* power, hipStream_t st);
extern int g_fir_fft;

int fft_rows_dispatch(int n, const float2* x, long long rows, bool inverse, float scale, float2* y,
                      hipStream_t st)
{
    return fft_rows(n, x, rows, inverse, scale, y, st);
}

int pc_template_dispatch(int n, const float2* h, int taps, int normalize, int window, float2* H, hipStream_t st)
{
    return pc_template(n, h, taps, normalize, window, nullptr, H, st);
}

}  // namespace zkx

using namespace zkx;

// ---------------------------------------------------------------------------
// Fixed device scratch for the allocation-free CA-CFAR entry point: the FP32
// integral image of one map of at most kNoAllocCfarCells cells (16 MiB, held
// by every device the library runs on until the process ends).  Host code
// never names `scratch` directly: its device address is looked up once per
// device and cached.
// ---------------------------------------------------------------------------
constexpr long long kNoAllocCfarCells = 1LL << 22;
static __device__ __align__(256) float scratch[kNoAllocCfarCells];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device, with that device current).
int scratch_address(float** out)
{
    int dev = 0;
    ZKX_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) ZKX_FAIL(ZKX_INTERNAL_ERROR, "ca_cfar_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error(e != hipSuccess ? hip_status_to_code(e) : ZKX_RUNTIME_ERROR,
                           "ca_cfar_noalloc: scratch address lookup failed");
            return e != hipSuccess ? hip_status_to_code(e) : ZKX_RUNTIME_ERROR;
        }
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ZKX_SUCCESS;
}
}  // namespace

extern "C" {

long long zkx_ca_cfar_noalloc_max_cells(void) { return kNoAllocCfarCells; }

int zkx_ca_cfar_f32_noalloc(const float* x, int rank, const int* shape, const int* guard, const int* reference,
                            double pfa, float* threshold, uint8_t* detections, void* stream)
{
    clear_last_error();
    if (rank == 1 || rank == 2) {
        if (!shape) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar: null pointer");
        long long count = 1;
        for (int d = 0; d < rank; ++d) count *= shape[d] > 0 ? shape[d] : 0;
        if (count > kNoAllocCfarCells)
            ZKX_FAIL(ZKX_INVALID_SHAPE, "ca_cfar_noalloc: the map exceeds the 4194304-cell scratch");
    }
    float* ws = nullptr;
    if (rank == 1 || rank == 2) {
        const int rc = scratch_address(&ws);
        if (rc) return rc;
    }
    // the integral image is written in full before it is read, so the
    // scratch needs no clearing between calls
    return zkx_ca_cfar_f32(x, rank, shape, guard, reference, pfa, threshold, detections, ws, stream);
}

int zkx_max_lds_fft_length(void) { return kMaxLdsFft; }

int zkx_set_kernel_variant(const char* op, int variant)
{
    clear_last_error();
    if (!op) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "kernel variant: null op");
    const std::string name(op);
    if (name == "pulse_compression") {
        if (variant < 0 || variant > 3) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "pulse_compression variant out of range");
        g_pc_variant = variant;
        return ZKX_SUCCESS;
    }
    if (name == "pulse_doppler") {
        if (variant < 0 || variant > 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "pulse_doppler variant out of range");
        g_doppler_variant = variant;
        return ZKX_SUCCESS;
    }
    if (name == "fir_fft") {
        if (variant < 0 || variant > 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fir_fft variant out of range");
        g_fir_fft = variant;
        return ZKX_SUCCESS;
    }
    if (name == "doppler_cfar") {
        if (variant < 0 || variant > 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "doppler_cfar variant out of range");
        g_dc_variant = variant;
        return ZKX_SUCCESS;
    }
    if (name == "dwell_fused") {
        if (variant < 0 || variant > 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "dwell_fused variant must be 0 or 1");
        g_fuse_variant = variant;
        return ZKX_SUCCESS;
    }
    if (name == "dwell_fused_rows") {
        if (variant < 1 || variant > 128) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "dwell_fused_rows must be in [1, 128]");
        g_fuse_rows = variant;
        return ZKX_SUCCESS;
    }
    if (name == "dwell_fused_chunks") {
        if (variant < 1 || variant > 256) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "dwell_fused_chunks must be in [1, 256]");
        g_fuse_chunks = variant;
        return ZKX_SUCCESS;
    }
    if (name == "dwell_fused_lead") {
        if (variant < 0 || variant > 1024) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "dwell_fused_lead must be in [0, 1024]");
        g_fuse_lead = variant;
        return ZKX_SUCCESS;
    }
    if (name == "doppler_cfar_strip_blocks") {
        if (variant < 1 || variant > (1 << 20)) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "strip block target out of range");
        g_dc_strip_blocks = variant;
        return ZKX_SUCCESS;
    }
    if (name == "inject_fft_failure") {
        if (variant < 0 || variant > 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "inject_fft_failure must be 0 or 1");
        g_inject_fft_failure.store(variant);
        return ZKX_SUCCESS;
    }

    ZKX_FAIL(ZKX_INVALID_ARGUMENT, "unknown kernel variant op");
}

int zkx_get_kernel_variant(const char* op)
{
    if (!op) return -1;
    const std::string name(op);
    if (name == "pulse_compression") return g_pc_variant;
    if (name == "pulse_doppler") return g_doppler_variant;
    if (name == "fir_fft") return g_fir_fft;
    if (name == "doppler_cfar") return g_dc_variant;
    if (name == "doppler_cfar_strip_blocks") return g_dc_strip_blocks;
    if (name == "dwell_fused") return g_fuse_variant;
    if (name == "dwell_fused_rows") return g_fuse_rows;
    if (name == "dwell_fused_lead") return g_fuse_lead;
    if (name == "inject_fft_failure") return g_inject_fft_failure.load();
    return -1;
}

int zkx_pc_template_prepare(const zkx_complex64* tmpl, int template_len, int nfft, int normalize, int window,
                            const float* explicit_window, zkx_complex64* spectrum, void* stream)
{
    clear_last_error();
    if (nfft < 1 || template_len < 1 || !tmpl || !spectrum)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "pulse_compression shape is invalid");
    if (!valid_window(window)) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "unsupported pulse_compression window");
    if (explicit_kind(window) != (explicit_window != nullptr))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "pulse_compression window is invalid");
    if (nfft == 1) {
        // a 1-point transform is the (windowed, normalised) first tap itself
        ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "pulse_compression nfft must be a power of two in [2, 8192]");
    }
    return pc_template(nfft, reinterpret_cast<const float2*>(tmpl), template_len, normalize, window,
                       explicit_window, reinterpret_cast<float2*>(spectrum), (hipStream_t)stream);
}

int zkx_pulse_compression_prepared_c64(const zkx_complex64* x, int num_pulses, int samples_per_pulse,
                                       const zkx_complex64* spectrum, int nfft, zkx_complex64* out,
                                       void* stream)
{
    clear_last_error();
    if (num_pulses < 1 || samples_per_pulse < 1 || nfft < 1 || !x || !spectrum || !out)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "pulse_compression shape is invalid");
    if (int rc = injected_fft_failure("pulse_compression")) return rc;
    return pc_fused(nfft, reinterpret_cast<const float2*>(x), num_pulses, samples_per_pulse,
                    reinterpret_cast<const float2*>(spectrum), reinterpret_cast<float2*>(out), (hipStream_t)stream);
}

int zkx_pulse_compression_c64(const zkx_complex64* x, int num_pulses, int samples_per_pulse,
                              const zkx_complex64* tmpl, int template_len, int nfft, int normalize, int window,
                              const float* explicit_window, zkx_complex64* out, void* stream)
{
    clear_last_error();
    if (nfft == -1) nfft = samples_per_pulse;
    const long long in_count = static_cast<long long>(num_pulses) * samples_per_pulse;
    const long long out_count = static_cast<long long>(num_pulses) * nfft;
    if (num_pulses < 1 || samples_per_pulse < 1 || nfft < 1 || template_len < 1 ||
        in_count > std::numeric_limits<int>::max() || out_count > std::numeric_limits<int>::max() || !x ||
        !tmpl || !out)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "pulse_compression shape is invalid");
    if (!valid_window(window)) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "unsupported pulse_compression window");
    if (explicit_kind(window) != (explicit_window != nullptr))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "pulse_compression window is invalid");
    if (int rc = injected_fft_failure("pulse_compression")) return rc;
    // per-call spectrum buffer, kept per host thread like the reference's plan pool
    static thread_local ThreadWorkspace ws;
    if (!is_pow2(nfft) || nfft < 2 || nfft > kMaxLdsFft) {
        // any other nfft: overlap-save segments of a power-of-two FFT over
        // the period nfft (the reference uses Bluestein for these lengths)
        const int keep = std::min(template_len, nfft);
        int L = 0, nseg = 0;
        const int N = pc_segment_length(nfft, keep, &L, &nseg);
        if (!N) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "pulse_compression: template too long for this nfft");
        void* H = nullptr;
        WorkspaceLease lease_(ws, (hipStream_t)stream);
        int rc = lease_.get(sizeof(float2) * static_cast<size_t>(N), &H);
        if (rc) return rc;
        rc = pc_template(N, reinterpret_cast<const float2*>(tmpl), template_len, normalize, window, explicit_window,
                         static_cast<float2*>(H), (hipStream_t)stream, keep);
        if (rc) return rc;
        return pc_segments(N, reinterpret_cast<const float2*>(x), static_cast<long long>(num_pulses) * nseg, nseg, L,
                           nfft, std::min(samples_per_pulse, nfft), samples_per_pulse,
                           static_cast<const float2*>(H), reinterpret_cast<float2*>(out), (hipStream_t)stream);
    }
    void* H = nullptr;
    WorkspaceLease lease_(ws, (hipStream_t)stream);
    int rc = lease_.get(sizeof(float2) * static_cast<size_t>(nfft), &H);
    if (rc) return rc;
    rc = zkx_pc_template_prepare(tmpl, template_len, nfft, normalize, window, explicit_window,
                                 static_cast<zkx_complex64*>(H), stream);
    if (rc) return rc;
    return zkx_pulse_compression_prepared_c64(x, num_pulses, samples_per_pulse, static_cast<zkx_complex64*>(H),
                                              nfft, out, stream);
}

int zkx_pulse_doppler_c64(const zkx_complex64* x, int num_pulses, int samples_per_pulse, int nfft, int window,
                          const float* explicit_window, zkx_complex64* out, float* power, void* stream)
{
    clear_last_error();
    if (nfft == -1) nfft = num_pulses;
    const long long in_count = static_cast<long long>(num_pulses) * samples_per_pulse;
    const long long out_count = static_cast<long long>(nfft) * samples_per_pulse;
    if (num_pulses < 1 || samples_per_pulse < 1 || nfft < 1 || in_count > std::numeric_limits<int>::max() ||
        out_count > std::numeric_limits<int>::max() || !x || (!out && !power))
        ZKX_FAIL(ZKX_INVALID_SHAPE, "pulse_doppler shape is invalid");
    if (!valid_window(window)) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "unsupported pulse_doppler window");
    if (explicit_kind(window) != (explicit_window != nullptr))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "pulse_doppler window is invalid");
    if (int rc = injected_fft_failure("pulse_doppler")) return rc;
    if (!is_pow2(nfft) || nfft < 2 || nfft > kMaxLdsFft) {
        // other lengths: FP32 row transforms of any length (Bluestein beyond
        // the powers of two, as the reference's backend does)
        return doppler_any(reinterpret_cast<const float2*>(x), num_pulses, samples_per_pulse, nfft, window,
                           explicit_window, reinterpret_cast<float2*>(out), power, (hipStream_t)stream);
    }
    return doppler(nfft, reinterpret_cast<const float2*>(x), 1, num_pulses, samples_per_pulse, window,
                   explicit_window, reinterpret_cast<float2*>(out), power, nullptr, (hipStream_t)stream);
}

int zkx_cfar_alpha(double pfa, int reference, double* alpha)
{
    clear_last_error();
    if (!alpha) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "cfar_alpha: null output");
    if (reference < 1 || !std::isfinite(pfa) || pfa <= 0.0 || pfa > 1.0)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "cfar_alpha requires reference >= 1 and pfa in (0,1]");
    *alpha = static_cast<double>(reference) * (std::pow(pfa, -1.0 / static_cast<double>(reference)) - 1.0);
    return ZKX_SUCCESS;
}

int zkx_ca_cfar_f32(const float* x, int rank, const int* shape, const int* guard, const int* reference, double pfa,
                    float* threshold, uint8_t* detections, float* workspace, void* stream)
{
    clear_last_error();
    hipStream_t st = (hipStream_t)stream;
    if (rank > 2 || rank < 0) ZKX_FAIL(ZKX_INVALID_SHAPE, "ca_cfar supports only scalar, 1D, or 2D input");
    if (rank == 0) {
        // scalar: threshold 0, detection x > 0 (no reference window exists)
        if (!x || !threshold || !detections) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar: null pointer");
        ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "ca_cfar scalar input is handled by the host wrapper");
    }
    if (!shape || !guard || !reference || !x || !threshold || !detections)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar: null pointer");
    for (int d = 0; d < rank; ++d) {
        if (shape[d] < 0) ZKX_FAIL(ZKX_INVALID_SHAPE, "ca_cfar shape dimensions must be nonnegative");
        if (guard[d] < 0 || reference[d] < 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar guard/reference cells are invalid");
        if (static_cast<long long>(shape[d]) <= 2LL * (guard[d] + reference[d]))
            ZKX_FAIL(ZKX_INVALID_SHAPE, "ca_cfar input dimension is too small");
    }
    if (!std::isfinite(pfa) || pfa <= 0.0 || pfa > 1.0) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar pfa must be in (0,1]");
    const long long count = rank == 2 ? static_cast<long long>(shape[0]) * shape[1] : shape[0];
    if (count > std::numeric_limits<int>::max()) ZKX_FAIL(ZKX_INVALID_SHAPE, "ca_cfar input size or capacity is invalid");
    float* I = workspace;
    WorkspaceLease lease_(t_cfar_ws, st);  // held until the kernels below are enqueued
    if (!I) {
        void* p = nullptr;
        int rc = lease_.get(sizeof(float) * static_cast<size_t>(count), &p);
        if (rc) return rc;
        I = static_cast<float*>(p);
    }
    if (rank == 1) {
        const int n = shape[0], g = guard[0], r = reference[0];
        const int cnt = 2 * r;
        const float alpha = cfar_alpha_float(cnt, pfa);
        hipLaunchKernelGGL(cfar1d_scan_kernel, dim3(1), dim3(64), 0, st, x, n, I);
        ZKX_LAUNCH_CHECK();
        hipLaunchKernelGGL(cfar1d_threshold_kernel, dim3((n + 255) / 256), dim3(256), 0, st, x, I, n, g, r, alpha,
                           cnt, threshold, detections);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    }
    return cfar_faithful_2d(x, nullptr, 1, shape[0], shape[1], guard[0], guard[1], reference[0], reference[1], pfa,
                            threshold, detections, I, st);
}

int zkx_ca_cfar_window_f32(const float* x, int rows, int cols, int guard_rows, int guard_cols, int ref_rows,
                           int ref_cols, double pfa, float* threshold, uint8_t* detections, void* stream)
{
    clear_last_error();
    if (!x || !threshold || !detections) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar: null pointer");
    int rc = validate_cfar2d(rows, cols, guard_rows, guard_cols, ref_rows, ref_cols, pfa);
    if (rc) return rc;
    return cfar_window_2d(x, 1, rows, cols, guard_rows, guard_cols, ref_rows, ref_cols, pfa, threshold, detections,
                          (hipStream_t)stream);
}

int zkx_doppler_cfar_c64(const zkx_complex64* x, int batch, int num_pulses, int range_bins, int doppler_window,
                         int guard_doppler, int guard_range, int ref_doppler, int ref_range, double pfa,
                         float* threshold, uint8_t* detections, zkx_complex64* rd_map, float* power, void* stream)
{
    clear_last_error();
    const int P = num_pulses, R = range_bins;
    if (batch < 1 || P < 2 || R < 1 || !x || !threshold || !detections)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "doppler_cfar shape is invalid");
    if (!is_pow2(P) || P > 128) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "doppler_cfar needs power-of-two pulses <= 128");
    if (!valid_window(doppler_window) || explicit_kind(doppler_window))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "doppler_cfar window must be none, hann or hamming");
    int rc = validate_cfar2d(P, R, guard_doppler, guard_range, ref_doppler, ref_range, pfa);
    if (rc) return rc;
    bool done = false;
    rc = doppler_cfar(P, reinterpret_cast<const float2*>(x), batch, P, R, doppler_window, guard_doppler, guard_range,
                      ref_doppler, ref_range, pfa, threshold, detections, reinterpret_cast<float2*>(rd_map), power,
                      (hipStream_t)stream, &done);
    if (rc) return rc;
    if (!done) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "doppler_cfar range window too wide for the fused tile");
    return ZKX_SUCCESS;
}

size_t zkx_dwell_workspace_bytes(int batch, int num_pulses, int range_bins)
{
    const size_t cells = static_cast<size_t>(batch) * num_pulses * range_bins;
    // range-compressed cube (c64) + power map (f32) + integral image (f32) +
    // the fused pipeline's job counters (8 queues, an error flag and one
    // counter per dwell, a 256-byte line each), 256-byte aligned
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    return al(cells * sizeof(float2)) + al(cells * sizeof(float)) + al(cells * sizeof(float)) +
           (static_cast<size_t>(batch) + 9) * 256;
}

int zkx_dwell_batch_c64(const zkx_complex64* x, int batch, int num_pulses, int range_bins,
                        const zkx_complex64* spectrum, int doppler_window, int guard_doppler, int guard_range,
                        int ref_doppler, int ref_range, double pfa, int cfar_mode, float* threshold,
                        uint8_t* detections, zkx_complex64* rd_map, void* workspace, void* stream)
{
    clear_last_error();
    hipStream_t st = (hipStream_t)stream;
    const int P = num_pulses, R = range_bins;
    if (batch < 1 || P < 2 || R < 2 || !x || !spectrum || !threshold || !detections || !workspace)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "dwell batch shape is invalid");
    if (!is_pow2(P) || !is_pow2(R) || P > kMaxLdsFft || R > kMaxLdsFft)
        ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "dwell batch needs power-of-two pulses and range bins <= 8192");
    if (!valid_window(doppler_window) || explicit_kind(doppler_window))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "dwell batch doppler window must be none, hann or hamming");
    if (cfar_mode != 0 && cfar_mode != 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "cfar_mode must be 0 or 1");
    int rc = validate_cfar2d(P, R, guard_doppler, guard_range, ref_doppler, ref_range, pfa);
    if (rc) return rc;
    const size_t cells = static_cast<size_t>(batch) * P * R;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    char* ws = static_cast<char*>(workspace);
    float2* pc = reinterpret_cast<float2*>(ws);
    float* pw = reinterpret_cast<float*>(ws + al(cells * sizeof(float2)));
    float* I = reinterpret_cast<float*>(ws + al(cells * sizeof(float2)) + al(cells * sizeof(float)));
    int* queue = reinterpret_cast<int*>(ws + al(cells * sizeof(float2)) + 2 * al(cells * sizeof(float)));
    if (cfar_mode == 1 && g_fuse_variant == 0) {
        bool done = false;
        rc = dwell_fused(reinterpret_cast<const float2*>(x), batch, P, R, reinterpret_cast<const float2*>(spectrum),
                         doppler_window, guard_doppler, guard_range, ref_doppler, ref_range, pfa, threshold,
                         detections, reinterpret_cast<float2*>(rd_map), pc, queue, st, &done);
        if (rc || done) return rc;
    }
    rc = pc_fused(R, reinterpret_cast<const float2*>(x), static_cast<long long>(batch) * P, R,
                  reinterpret_cast<const float2*>(spectrum), pc, st);
    if (rc) return rc;
    if (cfar_mode == 1) {
        bool done = false;
        rc = doppler_cfar(P, pc, batch, P, R, doppler_window, guard_doppler, guard_range, ref_doppler, ref_range, pfa,
                          threshold, detections, reinterpret_cast<float2*>(rd_map), nullptr, st, &done);
        if (rc || done) return rc;
    }
    rc = doppler(P, pc, batch, P, R, doppler_window, nullptr, reinterpret_cast<float2*>(rd_map), pw,
                 cfar_mode == 0 ? I : nullptr, st);
    if (rc) return rc;
    if (cfar_mode == 0)
        return cfar_faithful_2d(pw, I, batch, P, R, guard_doppler, guard_range, ref_doppler, ref_range, pfa, threshold,
                                detections, I, st);
    return cfar_window_2d(pw, batch, P, R, guard_doppler, guard_range, ref_doppler, ref_range, pfa, threshold,
                          detections, st);
}

}  // extern "C"
