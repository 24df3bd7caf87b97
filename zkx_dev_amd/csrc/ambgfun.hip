// Normalized ambiguity function of a pulse waveform for gfx950.
//
// Follows ambgfun_device / ambgfun_fp32_compute_device
// (ZKX/cusignal_cpp/src/operators/radartools/radartools_typed.cpp:411-462,
//  radartools_typed.cu:30-126, kernels radartools_kernels.cuh:257-460):
//
//   xn = x / ||x||, yn = y / ||y|| (y = x when not given), rows = nx + ny - 1,
//   nfreq = next power of two >= rows.
//   2-D cut [rows, nfreq]:
//     A[r, (f + nfreq/2) % nfreq] = | sum_{c < ny} yn[c] conj(xn[c - (nx-1) + r]) e^{+2 pi i f c / nfreq} |
//     (nfreq * |IFFT| of the delay products, fftshifted along Doppler).
//   delay cut [nfreq]: xs = IFFT(FFT(xn) * e^{2 pi i Fd tau}), Fd = -fs/2 + fs k/nfreq;
//     out[i] = nfreq * |IFFT(yn * conj(xs))[(i + nfreq/2) % nfreq]|.
//   Doppler cut [rows]: out[i] = |sum_a yn[a] conj(xm[(a - m) mod rows])|,
//     m = (i - rows/2) mod rows, xm = xn e^{2 pi i nu n / fs}
//     (the reference's circular correlation of length rows by FFTs; with
//     rows = nx + ny - 1 it has no wrap-around, so it is evaluated directly).
//
// Beyond one LDS pass (nfreq > 8192, up to 2^26) the same steps run through
// the global any-length FFT: the 2-D cut in chunks of rows (delay products
// written to a workspace, unscaled inverse transforms, |.| stored with the
// Doppler shift), the delay cut as three whole-length transforms.
//
// MI355X layout: the 2-D cut is one fused kernel -- each block builds its rows'
// delay products in registers from the two waveforms (L2-resident), runs the
// inverse transform in LDS (Stockham, radix 16) and writes |.| with the
// Doppler shift folded into the store index, so HBM sees only the output
// (4 or 8 bytes per cell).  The delay cut's three transforms run in one block.
// Norms are FP64 sums (the reference uses an FP32 scaled sum); every other
// step is FP32, as in the reference.
#include "fft_any.hpp"
#include "fft_lds.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace {

using zkx::fft::RowCfg;
using zkx::fft::RowLds;

constexpr int kMaxLen = 8192;
constexpr int kNormThreads = 1024;

// norm[b] = ||src_b|| over count_b complex samples, FP64 accumulation.
__global__ void __launch_bounds__(kNormThreads)
amb_norm_kernel(const float2* __restrict__ x, long long x_count, const float2* __restrict__ y,
                long long y_count, float* __restrict__ norm)
{
    __shared__ double part[kNormThreads / 64];
    const float2* src = blockIdx.x == 0 ? x : y;
    const long long count = blockIdx.x == 0 ? x_count : y_count;
    double s = 0.0;
    for (long long i = threadIdx.x; i < count; i += kNormThreads) {
        const float2 v = src[i];
        s += static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kNormThreads / 64; ++w) t += part[w];
        norm[blockIdx.x] = static_cast<float>(sqrt(t));
    }
}

template <class O>
__device__ __forceinline__ O out_value(float v) { return static_cast<O>(v); }

// |v| without hypot's rescaling and with the hardware square root (1 ulp):
// the normalised products keep |v| <= ny, far from overflow, and values below
// 1e-19 only lose digits they do not need.
__device__ __forceinline__ float magnitude(float2 v) { return __builtin_amdgcn_sqrtf(fmaf(v.x, v.x, v.y * v.y)); }

// 2-D cut.  Each LDS slot owns one delay row r; thread t of the slot holds
// columns c = t + k*T.  The unscaled inverse transform gives
// sum_c d[c] e^{+2 pi i f c / N} = N * IFFT(d)[f], whose magnitude is stored
// at column (f + N/2) % N.
template <int N, class O>
__global__ void __launch_bounds__(RowCfg<N>::THREADS)
amb_2d_kernel(const float2* __restrict__ x, int nx, const float2* __restrict__ y, int ny, int rows,
              const float* __restrict__ norms, const float2* __restrict__ tw, O* __restrict__ out)
{
    using C = RowCfg<N>;
    extern __shared__ float lds[];
    const int slot = threadIdx.x / C::T, t = threadIdx.x % C::T;
    const int row = blockIdx.x * C::FPB + slot;
    const bool active = row < rows;
    // one reciprocal per thread instead of a division per product
    const float inv_den = 1.0f / (norms[0] * norms[1]);
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int c = t + k * C::T;
        const int xc = c - (nx - 1) + row;
        float2 d = make_float2(0.f, 0.f);
        if (active && c < ny && xc >= 0 && xc < nx) d = zkx::cscale(zkx::cmulc(y[c], x[xc]), inv_den);
        v[k] = d;
    }
    if constexpr (N > 1) {
        RowLds L{reinterpret_cast<float2*>(lds + slot * 2 * C::SLOT_FLOATS)};
        zkx::fft::transform<N, C::E, true>(v, L, tw, t);
    }
    if (active) {
        O* dst = out + static_cast<long long>(row) * N;
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            const int f = t + k * C::T;
            dst[(f + N / 2) & (N - 1)] = out_value<O>(magnitude(v[k]));
        }
    }
}

// Delay cut in one block of exactly RowCfg<N>::T threads (one signal).
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::T)
amb_delay_kernel(const float2* __restrict__ x, long long x_count, const float2* __restrict__ y, int ny,
                 const float* __restrict__ norms, double fs, double tau, const float2* __restrict__ tw,
                 double* __restrict__ out)
{
    using C = RowCfg<N>;
    extern __shared__ float lds[];
    const int t = threadIdx.x;
    RowLds L{reinterpret_cast<float2*>(lds)};
    const float xn = norms[0], yn = norms[1];
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int i = t + k * C::T;
        float2 a = make_float2(0.f, 0.f);
        if (i < x_count) a = make_float2(x[i].x / xn, x[i].y / xn);
        v[k] = a;
    }
    if constexpr (N > 1) zkx::fft::transform<N, C::E, false>(v, L, tw, t);
    // spectrum x e^{2 pi i Fd tau}, phase reduced in FP64
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int i = t + k * C::T;
        const double fd = -0.5 * fs + fs * static_cast<double>(i) / N;
        double turns = fd * tau;
        turns -= rint(turns);
        double s, c;
        sincospi(2.0 * turns, &s, &c);
        v[k] = zkx::cmul(v[k], make_float2(static_cast<float>(c), static_cast<float>(s)));
    }
    if constexpr (N > 1) zkx::fft::transform<N, C::E, true>(v, L, tw, t);
    const float inv_n = 1.0f / N;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int i = t + k * C::T;
        float2 p = make_float2(0.f, 0.f);
        if (i < ny) {
            const float2 xs = zkx::cscale(v[k], inv_n);
            const float2 b = make_float2(y[i].x / yn, y[i].y / yn);
            p = zkx::cmulc(b, xs);
        }
        v[k] = p;
    }
    if constexpr (N > 1) zkx::fft::transform<N, C::E, true>(v, L, tw, t);
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int f = t + k * C::T;
        out[(f + N / 2) & (N - 1)] = static_cast<double>(hypotf(v[k].x, v[k].y));
    }
}

// Doppler cut, step 1: ws[0, nx) = xn * e^{2 pi i nu n / fs}, ws[nx, nx+ny) = yn.
__global__ void amb_doppler_prep_kernel(const float2* __restrict__ x, int nx, const float2* __restrict__ y,
                                        int ny, const float* __restrict__ norms, double fs, double nu,
                                        float2* __restrict__ ws)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nx) {
        double turns = nu * (static_cast<double>(i) / fs);
        turns -= rint(turns);
        double s, c;
        sincospi(2.0 * turns, &s, &c);
        const float2 a = make_float2(x[i].x / norms[0], x[i].y / norms[0]);
        ws[i] = zkx::cmul(a, make_float2(static_cast<float>(c), static_cast<float>(s)));
    } else if (i < nx + ny) {
        const int j = i - nx;
        ws[i] = make_float2(y[j].x / norms[1], y[j].y / norms[1]);
    }
}

// Doppler cut, step 2: one output lag per wavefront, its lanes striding the
// sum so that the loads of a wave are contiguous.  For lag m the live terms
// are a in [max(0, m'), min(ny, nx + m')) with a - m' the x index, m' the
// signed lag (m for m < ny, m - rows otherwise, so that a - m' stays in
// [0, nx)).
constexpr int kDopplerWaves = 4;
__global__ void __launch_bounds__(64 * kDopplerWaves)
amb_doppler_kernel(const float2* __restrict__ ws, int nx, int ny, int rows, double* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * kDopplerWaves + (threadIdx.x >> 6);
    if (i >= rows) return;  // whole waves leave together; no barrier follows
    const int m = (i - rows / 2 + rows) % rows;
    const int lag = m < ny ? m : m - rows;
    const int a0 = lag > 0 ? lag : 0;
    const int a1 = (nx + lag) < ny ? (nx + lag) : ny;
    const float2* xm = ws;
    const float2* yn = ws + nx;
    float re = 0.f, im = 0.f;
    for (int a = a0 + lane; a < a1; a += 64) {
        const float2 p = yn[a], q = xm[a - lag];
        re = fmaf(p.x, q.x, fmaf(p.y, q.y, re));
        im = fmaf(p.y, q.x, fmaf(-p.x, q.y, im));
    }
    for (int o = 32; o > 0; o >>= 1) {
        re += __shfl_xor(re, o);
        im += __shfl_xor(im, o);
    }
    if (lane == 0) out[i] = static_cast<double>(hypotf(re, im));
}

template <int N, class O>
int launch_2d(const float2* x, int nx, const float2* y, int ny, int rows, const float* norms, O* out,
              hipStream_t st)
{
    using C = RowCfg<N>;
    const float2* tw = nullptr;
    if (N > 1) {
        int rc = zkx::twiddles(N, &tw);
        if (rc) return rc;
    }
    const size_t lds = C::LDS_BYTES;
    if (lds > 64 * 1024)
        ZKX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(amb_2d_kernel<N, O>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    const unsigned blocks = static_cast<unsigned>((rows + C::FPB - 1) / C::FPB);
    hipLaunchKernelGGL((amb_2d_kernel<N, O>), dim3(blocks), dim3(C::THREADS), lds, st, x, nx, y, ny, rows,
                       norms, tw, out);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

template <int N>
int launch_delay(const float2* x, long long x_count, const float2* y, int ny, const float* norms, double fs,
                 double tau, double* out, hipStream_t st)
{
    using C = RowCfg<N>;
    const float2* tw = nullptr;
    if (N > 1) {
        int rc = zkx::twiddles(N, &tw);
        if (rc) return rc;
    }
    const size_t lds = sizeof(float) * 2 * C::SLOT_FLOATS;
    if (lds > 64 * 1024)
        ZKX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(amb_delay_kernel<N>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL((amb_delay_kernel<N>), dim3(1), dim3(C::T), lds, st, x, x_count, y, ny, norms, fs, tau,
                       tw, out);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

template <class O>
int dispatch_2d(int nfreq, const float2* x, int nx, const float2* y, int ny, int rows, const float* norms, O* out,
                hipStream_t st)
{
    switch (nfreq) {
    case 1: return launch_2d<1, O>(x, nx, y, ny, rows, norms, out, st);
    case 2: return launch_2d<2, O>(x, nx, y, ny, rows, norms, out, st);
    case 4: return launch_2d<4, O>(x, nx, y, ny, rows, norms, out, st);
    case 8: return launch_2d<8, O>(x, nx, y, ny, rows, norms, out, st);
    case 16: return launch_2d<16, O>(x, nx, y, ny, rows, norms, out, st);
    case 32: return launch_2d<32, O>(x, nx, y, ny, rows, norms, out, st);
    case 64: return launch_2d<64, O>(x, nx, y, ny, rows, norms, out, st);
    case 128: return launch_2d<128, O>(x, nx, y, ny, rows, norms, out, st);
    case 256: return launch_2d<256, O>(x, nx, y, ny, rows, norms, out, st);
    case 512: return launch_2d<512, O>(x, nx, y, ny, rows, norms, out, st);
    case 1024: return launch_2d<1024, O>(x, nx, y, ny, rows, norms, out, st);
    case 2048: return launch_2d<2048, O>(x, nx, y, ny, rows, norms, out, st);
    case 4096: return launch_2d<4096, O>(x, nx, y, ny, rows, norms, out, st);
    case 8192: return launch_2d<8192, O>(x, nx, y, ny, rows, norms, out, st);
    default: ZKX_FAIL(ZKX_INTERNAL_ERROR, "ambgfun length");
    }
}

int dispatch_delay(int nfreq, const float2* x, long long x_count, const float2* y, int ny, const float* norms,
                   double fs, double tau, double* out, hipStream_t st)
{
    switch (nfreq) {
    case 1: return launch_delay<1>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 2: return launch_delay<2>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 4: return launch_delay<4>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 8: return launch_delay<8>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 16: return launch_delay<16>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 32: return launch_delay<32>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 64: return launch_delay<64>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 128: return launch_delay<128>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 256: return launch_delay<256>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 512: return launch_delay<512>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 1024: return launch_delay<1024>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 2048: return launch_delay<2048>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 4096: return launch_delay<4096>(x, x_count, y, ny, norms, fs, tau, out, st);
    case 8192: return launch_delay<8192>(x, x_count, y, ny, norms, fs, tau, out, st);
    default: ZKX_FAIL(ZKX_INTERNAL_ERROR, "ambgfun length");
    }
}

thread_local zkx::ThreadWorkspace t_amb_ws;

// ---------------------------------------------------------------------------
// nfreq > 8192
// ---------------------------------------------------------------------------
// D[rr][c] = yn[c] conj(xn[c - (nx-1) + row0 + rr]) for c < N (0 outside).
__global__ void __launch_bounds__(256)
amb_products_kernel(const float2* __restrict__ x, int nx, const float2* __restrict__ y, int ny, long long row0,
                    long long chunk, int N, const float* __restrict__ norms, float2* __restrict__ D)
{
    const float inv_den = 1.0f / (norms[0] * norms[1]);
    const long long total = chunk * N;
    const int lg = __builtin_ctz(static_cast<unsigned>(N));  // N is a power of two
    for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += static_cast<long long>(gridDim.x) * 256) {
        const long long rr = e >> lg;
        const int c = static_cast<int>(e & (N - 1));
        const long long xc = c - (nx - 1) + row0 + rr;
        float2 d = make_float2(0.f, 0.f);
        if (c < ny && xc >= 0 && xc < nx) d = zkx::cscale(zkx::cmulc(y[c], x[xc]), inv_den);
        D[e] = d;
    }
}

template <class O>
__global__ void __launch_bounds__(256)
amb_store_kernel(const float2* __restrict__ D, long long chunk, int N, long long row0, O* __restrict__ out)
{
    const long long total = chunk * N;
    const int lg = __builtin_ctz(static_cast<unsigned>(N));
    for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += static_cast<long long>(gridDim.x) * 256) {
        const long long rr = e >> lg;
        const int f = static_cast<int>(e & (N - 1));
        out[(row0 + rr) * N + ((f + N / 2) & (N - 1))] = out_value<O>(magnitude(D[e]));
    }
}

// Delay cut steps on one signal of N samples.
__global__ void __launch_bounds__(256)
amb_delay_load_kernel(const float2* __restrict__ x, long long x_count, int N, const float* __restrict__ norms,
                      float2* __restrict__ v)
{
    const float xn = norms[0];
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < N; i += static_cast<long long>(gridDim.x) * 256)
        v[i] = i < x_count ? make_float2(x[i].x / xn, x[i].y / xn) : make_float2(0.f, 0.f);
}

__global__ void __launch_bounds__(256)
amb_delay_phase_kernel(int N, double fs, double tau, float2* __restrict__ v)
{
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < N; i += static_cast<long long>(gridDim.x) * 256) {
        const double fd = -0.5 * fs + fs * static_cast<double>(i) / N;
        double turns = fd * tau;
        turns -= rint(turns);
        double s, c;
        sincospi(2.0 * turns, &s, &c);
        v[i] = zkx::cmul(v[i], make_float2(static_cast<float>(c), static_cast<float>(s)));
    }
}

__global__ void __launch_bounds__(256)
amb_delay_product_kernel(const float2* __restrict__ y, int ny, int N, const float* __restrict__ norms,
                         float2* __restrict__ v)
{
    const float yn = norms[1];
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < N; i += static_cast<long long>(gridDim.x) * 256) {
        float2 p = make_float2(0.f, 0.f);
        if (i < ny) p = zkx::cmulc(make_float2(y[i].x / yn, y[i].y / yn), v[i]);
        v[i] = p;
    }
}

__global__ void __launch_bounds__(256)
amb_delay_store_kernel(const float2* __restrict__ v, int N, double* __restrict__ out)
{
    for (long long f = blockIdx.x * 256LL + threadIdx.x; f < N; f += static_cast<long long>(gridDim.x) * 256)
        out[(f + N / 2) & (N - 1)] = static_cast<double>(hypotf(v[f].x, v[f].y));
}

unsigned amb_grid(long long n)
{
    const long long b = (n + 255) / 256;
    return static_cast<unsigned>(b < 16384 ? (b > 0 ? b : 1) : 16384);
}

// Rows per chunk of the long 2-D cut: about 256 MiB of complex products.
long long amb_chunk_rows(long long rows, int N)
{
    const long long c = std::max<long long>(1, (256LL << 20) / (8LL * N));
    return std::min(rows, c);
}

size_t amb_long_bytes(int cut, long long rows, int N)
{
    const long long r = cut == 0 ? amb_chunk_rows(rows, N) : 1;
    return (sizeof(float2) * static_cast<size_t>(r) * N + 255) / 256 * 256 + zkx::fft_any_workspace_bytes(r, N);
}

template <class O>
int amb_2d_long(const float2* x, int nx, const float2* y, int ny, long long rows, int N, const float* norms,
                char* ws, O* out, hipStream_t st)
{
    const long long chunk = amb_chunk_rows(rows, N);
    float2* D = reinterpret_cast<float2*>(ws);
    char* any = ws + (sizeof(float2) * static_cast<size_t>(chunk) * N + 255) / 256 * 256;
    for (long long row0 = 0; row0 < rows; row0 += chunk) {
        const long long c = std::min(chunk, rows - row0);
        hipLaunchKernelGGL(amb_products_kernel, dim3(amb_grid(c * N)), dim3(256), 0, st, x, nx, y, ny, row0, c, N,
                           norms, D);
        ZKX_LAUNCH_CHECK();
        int rc = zkx::fft_any_rows(D, c, N, true, 1.0f, D, any, st);
        if (rc) return rc;
        hipLaunchKernelGGL((amb_store_kernel<O>), dim3(amb_grid(c * N)), dim3(256), 0, st, D, c, N, row0, out);
        ZKX_LAUNCH_CHECK();
    }
    return ZKX_SUCCESS;
}

int amb_delay_long(const float2* x, long long x_count, const float2* y, int ny, int N, const float* norms, double fs,
                   double tau, char* ws, double* out, hipStream_t st)
{
    float2* v = reinterpret_cast<float2*>(ws);
    char* any = ws + (sizeof(float2) * static_cast<size_t>(N) + 255) / 256 * 256;
    const unsigned g = amb_grid(N);
    hipLaunchKernelGGL(amb_delay_load_kernel, dim3(g), dim3(256), 0, st, x, x_count, N, norms, v);
    ZKX_LAUNCH_CHECK();
    int rc = zkx::fft_any_rows(v, 1, N, false, 1.0f, v, any, st);
    if (rc) return rc;
    hipLaunchKernelGGL(amb_delay_phase_kernel, dim3(g), dim3(256), 0, st, N, fs, tau, v);
    ZKX_LAUNCH_CHECK();
    if ((rc = zkx::fft_any_rows(v, 1, N, true, 1.0f / static_cast<float>(N), v, any, st))) return rc;
    hipLaunchKernelGGL(amb_delay_product_kernel, dim3(g), dim3(256), 0, st, y, ny, N, norms, v);
    ZKX_LAUNCH_CHECK();
    if ((rc = zkx::fft_any_rows(v, 1, N, true, 1.0f, v, any, st))) return rc;
    hipLaunchKernelGGL(amb_delay_store_kernel, dim3(g), dim3(256), 0, st, v, N, out);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // namespace

extern "C" {

long long zkx_ambgfun_out_len(int nx, int ny, int cut)
{
    if (nx < 1 || ny < 1 || cut < 0 || cut > 2) return -1;
    const long long rows = static_cast<long long>(nx) + ny - 1;
    if (rows > std::numeric_limits<int>::max()) return -1;
    long long nfreq = 1;
    while (nfreq < rows) nfreq <<= 1;
    if (nfreq > std::numeric_limits<int>::max()) return -1;
    return cut == 0 ? rows * nfreq : (cut == 1 ? nfreq : rows);
}

int zkx_ambgfun_c64(const zkx_complex64* x, size_t x_count, int nx, const zkx_complex64* y, size_t y_count,
                    int ny, int cut, double fs, double cut_value, int out_fp64, void* out, void* stream)
{
    zkx::clear_last_error();
    if (cut < 0 || cut > 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ambgfun cut is invalid");
    if (cut != 0 && fs == 0.0) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ambgfun fs must be nonzero for one-dimensional cuts");
    if (x == nullptr || out == nullptr) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ambgfun pointers must be non-null");
    if (y == nullptr) {
        y = x;
        y_count = x_count;
        ny = nx;
    }
    if (x_count == 0 || y_count == 0) ZKX_FAIL(ZKX_INVALID_SHAPE, "ambgfun inputs must be nonempty");
    if (nx < 1 || ny < 1 || static_cast<size_t>(nx) > x_count || static_cast<size_t>(ny) > y_count)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "ambgfun lengths exceed the inputs");
    if (cut != 0 && (static_cast<size_t>(nx) != x_count || static_cast<size_t>(ny) != y_count))
        ZKX_FAIL(ZKX_INVALID_SHAPE, "ambgfun high-rank input is only defined for cut='2d'");
    if (x_count > static_cast<size_t>(std::numeric_limits<int>::max()) ||
        y_count > static_cast<size_t>(std::numeric_limits<int>::max()))
        ZKX_FAIL(ZKX_INVALID_SHAPE, "ambgfun input length exceeds int capacity");
    const long long rows = static_cast<long long>(nx) + ny - 1;
    long long nfreq = 1;
    while (nfreq < rows) nfreq <<= 1;
    if (cut != 2 && nfreq > zkx::kMaxAnyFftPow2)
        ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "ambgfun 2-D and delay cuts need nx + ny - 1 <= 2^26");
    const bool long_form = cut != 2 && nfreq > kMaxLen;
    if (cut != 1 && rows > std::numeric_limits<int>::max())
        ZKX_FAIL(ZKX_INVALID_SHAPE, "ambgfun input length exceeds int capacity");
    if (cut != 0 && !out_fp64) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ambgfun one-dimensional cuts return FP64");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const float2* xd = reinterpret_cast<const float2*>(x);
    const float2* yd = reinterpret_cast<const float2*>(y);

    const size_t norm_bytes = 256;
    const size_t ws_bytes = norm_bytes + (cut == 2 ? sizeof(float2) * static_cast<size_t>(nx + ny) : 0) +
                            (long_form ? amb_long_bytes(cut, rows, static_cast<int>(nfreq)) : 0);
    void* ws = nullptr;
    zkx::WorkspaceLease lease_(t_amb_ws, st);
    int rc = lease_.get(ws_bytes, &ws);
    if (rc) return rc;
    float* norms = static_cast<float*>(ws);
    hipLaunchKernelGGL(amb_norm_kernel, dim3(2), dim3(kNormThreads), 0, st, xd, static_cast<long long>(x_count),
                       yd, static_cast<long long>(y_count), norms);
    ZKX_LAUNCH_CHECK();

    if (long_form) {
        char* lw = static_cast<char*>(ws) + norm_bytes;
        if ((rc = zkx::fft_any_prepare(static_cast<int>(nfreq)))) return rc;
        if (cut == 1)
            return amb_delay_long(xd, static_cast<long long>(x_count), yd, ny, static_cast<int>(nfreq), norms, fs,
                                  cut_value, lw, static_cast<double*>(out), st);
        if (out_fp64)
            return amb_2d_long<double>(xd, nx, yd, ny, rows, static_cast<int>(nfreq), norms, lw,
                                       static_cast<double*>(out), st);
        return amb_2d_long<float>(xd, nx, yd, ny, rows, static_cast<int>(nfreq), norms, lw, static_cast<float*>(out),
                                  st);
    }
    if (cut == 0) {
        if (out_fp64)
            return dispatch_2d<double>(static_cast<int>(nfreq), xd, nx, yd, ny, static_cast<int>(rows), norms,
                                       static_cast<double*>(out), st);
        return dispatch_2d<float>(static_cast<int>(nfreq), xd, nx, yd, ny, static_cast<int>(rows), norms,
                                  static_cast<float*>(out), st);
    }
    if (cut == 1)
        return dispatch_delay(static_cast<int>(nfreq), xd, static_cast<long long>(x_count), yd, ny, norms, fs,
                              cut_value, static_cast<double*>(out), st);
    float2* prep = reinterpret_cast<float2*>(static_cast<char*>(ws) + norm_bytes);
    const int n_prep = nx + ny;
    hipLaunchKernelGGL(amb_doppler_prep_kernel, dim3((n_prep + 255) / 256), dim3(256), 0, st, xd, nx, yd, ny, norms,
                       fs, cut_value, prep);
    ZKX_LAUNCH_CHECK();
    hipLaunchKernelGGL(amb_doppler_kernel, dim3(static_cast<unsigned>((rows + kDopplerWaves - 1) / kDopplerWaves)),
                       dim3(64 * kDopplerWaves), 0, st, prep,
                       nx, ny, static_cast<int>(rows), static_cast<double*>(out));
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // extern "C"
