// The reference's FFT backend C ABI (fft_interface.h:189-447) on the gfx950
// LDS Stockham FFT, so this library can replace libzkx_fft_thrust.so at link
// time.  Semantics follow the reference backend
// (ZKX/cusignal_cpp/src/backends/fft/fft_thrust.cu):
//   * C2C of any length: powers of two up to 8192 in one LDS pass per row,
//     other lengths up to 4096 by Bluestein's chirp-z method (one fused
//     kernel per row: chirp, FFT_M, x B, IFFT_M, chirp; M = 2^ceil(log2(2n))),
//     longer ones (powers of two to 2^26, others to 2^25) by the global-pass
//     forms of fft_any.hip; inverse scaled by 1/n (fft_thrust.cu:700-800);
//   * R2C: forward transform of real rows, first n/2+1 bins
//     (fft_thrust.cu:1096-1120);
//   * C2R: the half spectrum is mirrored as the reference does
//     (conjugate_symmetry_kernel, fft_thrust.cu:566-581: bins 1 .. n/2+1-2
//     mirrored, everything else zero), inverse transformed, scaled by 1/n and
//     the real part kept; for lengths that are not powers of two the
//     reference's Bluestein inverse scales by 1/n once more
//     (fft_thrust.cu:1122-1160 with :750-760), and so does this one;
//   * 2-D plans transform complex [ny][nx] data (rows, then columns, for any
//     plan type, as fft_thrust.cu:1014-1058 does), inverse scaled by 1/(nx ny);
//   * shift, frequency and error-string helpers as fft_thrust.cu:520-550,
//     1228-1361, including fft_shift == fft_inverse_shift for 1-D arrays.
#include "../../include/zkx_fft.h"
#include "fft_any.hpp"
#include "fft_lds.hpp"
#include "runtime.hpp"

#include <cmath>
#include <new>
#include <vector>

namespace zkx {
int fft_rows_dispatch(int n, const float2* x, long long rows, bool inverse, float scale, float2* y,
                      hipStream_t st);
}

namespace {

using zkx::fft::RowCfg;

constexpr int kMaxPow2 = 8192;
constexpr int kMaxBluestein = 4096;

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
int next_pow2(int n)
{
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}

// Bluestein tables of one length n (both directions share |chirp|; the
// inverse uses conjugates): chirp[k] = e^{-i pi k^2 / n}, k < n, and the
// spectrum of the circular kernel b (b[k] = b[M-k] = conj(chirp[k])).
struct Chirp {
    int n = 0, M = 0;
    float2* chirp = nullptr;      // [n]
    float2* spec_fwd = nullptr;   // [M]
    float2* spec_inv = nullptr;   // [M]
};

struct Plan {
    int nx = 0, ny = 0, batch = 1;
    FFTType type = FFT_TYPE_C2C;
    FFTDirection dir = FFT_FORWARD;
    Chirp cx, cy;
    float2* work = nullptr;  // scratch: full-length complex rows / transposes
    size_t work_elems = 0;
    void* any_ws = nullptr;  // scratch of the global-pass (long) transforms
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int M, bool INV>
__global__ void __launch_bounds__(RowCfg<M>::THREADS)
bluestein_kernel(const float2* __restrict__ x, long long rows, int n, const float2* __restrict__ chirp,
                 const float2* __restrict__ spec, const float2* __restrict__ tw, float scale, float2* __restrict__ y)
{
    using C = RowCfg<M>;
    extern __shared__ float lds[];
    const int slot = threadIdx.x / C::T, t = threadIdx.x % C::T;
    const long long row = static_cast<long long>(blockIdx.x) * C::FPB + slot;
    const bool active = row < rows;
    zkx::fft::RowLds L{reinterpret_cast<float2*>(lds + slot * 2 * C::SLOT_FLOATS)};
    const float2* src = x + (active ? row : 0) * static_cast<long long>(n);
    float2 v[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int i = t + k * C::T;
        float2 a = make_float2(0.f, 0.f);
        if (active && i < n) {
            // forward: x * chirp; inverse: x * conj(chirp)
            a = INV ? zkx::cmulc(src[i], chirp[i]) : zkx::cmul(src[i], chirp[i]);
        }
        v[k] = a;
    }
    zkx::fft::transform<M, C::E, false>(v, L, tw, t);
#pragma unroll
    for (int k = 0; k < C::E; ++k) v[k] = zkx::cmul(v[k], spec[t + k * C::T]);
    zkx::fft::transform<M, C::E, true>(v, L, tw, t);
    if (active) {
        float2* dst = y + row * static_cast<long long>(n);
        const float s = scale / static_cast<float>(M);
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            const int i = t + k * C::T;
            if (i < n) dst[i] = zkx::cscale(INV ? zkx::cmulc(v[k], chirp[i]) : zkx::cmul(v[k], chirp[i]), s);
        }
    }
}

__global__ void real_to_complex_kernel(const float* __restrict__ in, float2* __restrict__ out, long long total)
{
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x)
        out[g] = make_float2(in[g], 0.f);
}

// first `half` bins of each full row of n
__global__ void half_rows_kernel(const float2* __restrict__ full, int n, int half, long long rows,
                                 float2* __restrict__ out)
{
    const long long total = static_cast<long long>(half) * rows;
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long r = g / half;
        const int k = static_cast<int>(g - r * half);
        out[g] = full[r * n + k];
    }
}

// full row of n from a half spectrum, mirroring bins 1 .. half-2 only
__global__ void mirror_rows_kernel(const float2* __restrict__ in, int n, int half, long long rows,
                                   float2* __restrict__ full)
{
    const long long total = static_cast<long long>(n) * rows;
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long r = g / n;
        const int k = static_cast<int>(g - r * n);
        float2 v = make_float2(0.f, 0.f);
        if (k < half) {
            v = in[r * half + k];
        } else {
            const int m = n - k;  // mirror source
            if (m >= 1 && m < half - 1) {
                const float2 s = in[r * half + m];
                v = make_float2(s.x, -s.y);
            }
        }
        full[g] = v;
    }
}

__global__ void real_part_kernel(const float2* __restrict__ in, long long total, float scale, float* __restrict__ out)
{
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x)
        out[g] = in[g].x * scale;
}

// out[c][r] = in[r][c] for in [rows][cols], through 32x32 LDS tiles
__global__ void transpose_kernel(const float2* __restrict__ in, int rows, int cols, float2* __restrict__ out)
{
    __shared__ float2 tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 8 rows per pass
    for (int j = ty; j < 32; j += 8) {
        const int r = r0 + j, c = c0 + tx;
        if (r < rows && c < cols) tile[j][tx] = in[static_cast<long long>(r) * cols + c];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int c = c0 + j, r = r0 + tx;
        if (c < cols && r < rows) out[static_cast<long long>(c) * rows + r] = tile[tx][j];
    }
}

__global__ void shift1d_kernel(const float2* __restrict__ in, float2* __restrict__ out, int n, int half, int batch)
{
    const long long total = static_cast<long long>(n) * batch;
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long b = g / n;
        const int i = static_cast<int>(g % n);
        out[b * n + i] = in[b * n + (i + half) % n];
    }
}

__global__ void shift2d_kernel(const float2* __restrict__ in, float2* __restrict__ out, int nx, int ny, int sx, int sy)
{
    const long long total = static_cast<long long>(nx) * ny;
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int x = static_cast<int>(g % nx), y = static_cast<int>(g / nx);
        out[static_cast<long long>(y) * nx + x] =
            in[static_cast<long long>((y + sy) % ny) * nx + (x + sx) % nx];
    }
}

unsigned blocks_for(long long n)
{
    const long long b = (n + 255) / 256;
    return static_cast<unsigned>(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
template <int M>
int launch_bluestein(const float2* x, long long rows, int n, const Chirp& c, bool inverse, float scale, float2* y,
                     hipStream_t st)
{
    using C = RowCfg<M>;
    const float2* tw = nullptr;
    int rc = zkx::twiddles(M, &tw);
    if (rc) return rc;
    const long long blocks = (rows + C::FPB - 1) / C::FPB;
    if (C::LDS_BYTES > 64 * 1024) {
        ZKX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(bluestein_kernel<M, true>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(C::LDS_BYTES)));
        ZKX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(bluestein_kernel<M, false>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(C::LDS_BYTES)));
    }
    if (inverse)
        hipLaunchKernelGGL((bluestein_kernel<M, true>), dim3(static_cast<unsigned>(blocks)), dim3(C::THREADS),
                           C::LDS_BYTES, st, x, rows, n, c.chirp, c.spec_inv, tw, scale, y);
    else
        hipLaunchKernelGGL((bluestein_kernel<M, false>), dim3(static_cast<unsigned>(blocks)), dim3(C::THREADS),
                           C::LDS_BYTES, st, x, rows, n, c.chirp, c.spec_fwd, tw, scale, y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

int bluestein(const float2* x, long long rows, int n, const Chirp& c, bool inverse, float scale, float2* y,
              hipStream_t st)
{
    switch (c.M) {
    case 8: return launch_bluestein<8>(x, rows, n, c, inverse, scale, y, st);
    case 16: return launch_bluestein<16>(x, rows, n, c, inverse, scale, y, st);
    case 32: return launch_bluestein<32>(x, rows, n, c, inverse, scale, y, st);
    case 64: return launch_bluestein<64>(x, rows, n, c, inverse, scale, y, st);
    case 128: return launch_bluestein<128>(x, rows, n, c, inverse, scale, y, st);
    case 256: return launch_bluestein<256>(x, rows, n, c, inverse, scale, y, st);
    case 512: return launch_bluestein<512>(x, rows, n, c, inverse, scale, y, st);
    case 1024: return launch_bluestein<1024>(x, rows, n, c, inverse, scale, y, st);
    case 2048: return launch_bluestein<2048>(x, rows, n, c, inverse, scale, y, st);
    case 4096: return launch_bluestein<4096>(x, rows, n, c, inverse, scale, y, st);
    case 8192: return launch_bluestein<8192>(x, rows, n, c, inverse, scale, y, st);
    default: ZKX_FAIL(ZKX_INTERNAL_ERROR, "bluestein length");
    }
}

// Transform rows of length n (any supported length) with the 1/n-style
// scale factor `scale` applied to the result.
int rows_fft(const float2* x, long long rows, int n, const Chirp& c, bool inverse, float scale, float2* y,
             void* any_ws, hipStream_t st)
{
    if (n == 1) {
        // a 1-point transform is the identity (every scale here is 1/1)
        if (x != y) ZKX_HIP_CHECK(hipMemcpyAsync(y, x, sizeof(float2) * rows, hipMemcpyDeviceToDevice, st));
        return ZKX_SUCCESS;
    }
    if (pow2(n) && n <= kMaxPow2) return zkx::fft_rows_dispatch(n, x, rows, inverse, scale, y, st);
    if (!pow2(n) && n <= kMaxBluestein) return bluestein(x, rows, n, c, inverse, scale, y, st);
    return zkx::fft_any_rows(x, rows, n, inverse, scale, y, any_ws, st);
}

void free_chirp(Chirp& c)
{
    if (c.chirp) (void)hipFree(c.chirp);
    if (c.spec_fwd) (void)hipFree(c.spec_fwd);
    if (c.spec_inv) (void)hipFree(c.spec_inv);
    c = Chirp();
}

// Chirp tables in FP64 on the host (k^2 reduced mod 2n exactly), spectra of
// the kernels by the device FFT.
FFTResult make_chirp(int n, Chirp& c)
{
    if (pow2(n) || n == 1) return FFT_SUCCESS;
    if (n > kMaxBluestein)  // the global-pass form keeps its own tables
        return zkx::fft_any_prepare(n) == ZKX_SUCCESS ? FFT_SUCCESS : FFT_ALLOC_FAILED;
    c.n = n;
    c.M = next_pow2(2 * n - 1);
    std::vector<float2> ch(n), bf(c.M, make_float2(0.f, 0.f)), bi(c.M, make_float2(0.f, 0.f));
    for (int k = 0; k < n; ++k) {
        const long long q = (static_cast<long long>(k) * k) % (2LL * n);
        const double th = M_PI * static_cast<double>(q) / n;
        ch[k] = make_float2(static_cast<float>(std::cos(th)), static_cast<float>(-std::sin(th)));
        // forward kernel conj(chirp) = e^{+i th}; inverse kernel chirp = e^{-i th}
        const float2 kf = make_float2(static_cast<float>(std::cos(th)), static_cast<float>(std::sin(th)));
        const float2 ki = make_float2(kf.x, -kf.y);
        bf[k] = kf;
        bi[k] = ki;
        if (k > 0) {
            bf[c.M - k] = kf;
            bi[c.M - k] = ki;
        }
    }
    if (hipMalloc(reinterpret_cast<void**>(&c.chirp), sizeof(float2) * n) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&c.spec_fwd), sizeof(float2) * c.M) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&c.spec_inv), sizeof(float2) * c.M) != hipSuccess) {
        free_chirp(c);
        return FFT_ALLOC_FAILED;
    }
    bool ok = hipMemcpy(c.chirp, ch.data(), sizeof(float2) * n, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c.spec_fwd, bf.data(), sizeof(float2) * c.M, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c.spec_inv, bi.data(), sizeof(float2) * c.M, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && zkx::fft_rows_dispatch(c.M, c.spec_fwd, 1, false, 1.0f, c.spec_fwd, nullptr) == ZKX_SUCCESS &&
         zkx::fft_rows_dispatch(c.M, c.spec_inv, 1, false, 1.0f, c.spec_inv, nullptr) == ZKX_SUCCESS &&
         hipStreamSynchronize(nullptr) == hipSuccess;
    if (!ok) {
        free_chirp(c);
        return FFT_EXEC_FAILED;
    }
    return FFT_SUCCESS;
}

bool length_ok(int n) { return zkx::fft_any_length_ok(n); }

// Workspace of the global-pass transforms a plan may run, allocated once.
FFTResult alloc_any_ws(Plan* p, size_t bytes)
{
    if (!bytes) return FFT_SUCCESS;
    if (hipMalloc(&p->any_ws, bytes) != hipSuccess) {
        (void)hipGetLastError();
        p->any_ws = nullptr;
        return FFT_ALLOC_FAILED;
    }
    return FFT_SUCCESS;
}

FFTResult to_result(int rc)
{
    return rc == ZKX_SUCCESS ? FFT_SUCCESS : (rc == ZKX_OUT_OF_MEMORY ? FFT_ALLOC_FAILED : FFT_EXEC_FAILED);
}

FFTResult run_plan(FFTPlanHandle handle, void* out, const void* in, hipStream_t st)
{
    if (!handle || !out || !in) return FFT_INVALID_VALUE;
    Plan* p = static_cast<Plan*>(handle);
    const bool inv = p->dir == FFT_INVERSE;
    const int nx = p->nx;
    int rc = ZKX_SUCCESS;
    if (p->ny > 0) {
        // 2-D: complex [ny][nx] in any plan type (as the reference backend)
        const int ny = p->ny;
        const long long total = static_cast<long long>(nx) * ny;
        float2* w = p->work;  // [total] scratch for the transposed array
        const float2* src = static_cast<const float2*>(in);
        float2* dst = static_cast<float2*>(out);
        rc = rows_fft(src, ny, nx, p->cx, inv, inv ? 1.0f / nx : 1.0f, dst, p->any_ws, st);
        if (!rc) {
            hipLaunchKernelGGL(transpose_kernel, dim3((nx + 31) / 32, (ny + 31) / 32), dim3(256), 0, st, dst, ny, nx, w);
            rc = hipGetLastError() == hipSuccess ? ZKX_SUCCESS : ZKX_RUNTIME_ERROR;
        }
        if (!rc) rc = rows_fft(w, nx, ny, p->cy, inv, inv ? 1.0f / ny : 1.0f, w, p->any_ws, st);
        if (!rc) {
            hipLaunchKernelGGL(transpose_kernel, dim3((ny + 31) / 32, (nx + 31) / 32), dim3(256), 0, st, w, nx, ny, dst);
            rc = hipGetLastError() == hipSuccess ? ZKX_SUCCESS : ZKX_RUNTIME_ERROR;
        }
        (void)total;
        return to_result(rc);
    }
    const long long rows = p->batch;
    const int half = nx / 2 + 1;
    switch (p->type) {
    case FFT_TYPE_C2C:
        rc = rows_fft(static_cast<const float2*>(in), rows, nx, p->cx, inv, inv ? 1.0f / nx : 1.0f,
                      static_cast<float2*>(out), p->any_ws, st);
        break;
    case FFT_TYPE_R2C: {
        const long long total = rows * nx;
        hipLaunchKernelGGL(real_to_complex_kernel, dim3(blocks_for(total)), dim3(256), 0, st,
                           static_cast<const float*>(in), p->work, total);
        rc = rows_fft(p->work, rows, nx, p->cx, false, 1.0f, p->work, p->any_ws, st);
        if (!rc)
            hipLaunchKernelGGL(half_rows_kernel, dim3(blocks_for(rows * half)), dim3(256), 0, st, p->work, nx, half,
                               rows, static_cast<float2*>(out));
        break;
    }
    case FFT_TYPE_C2R: {
        const long long total = rows * nx;
        hipLaunchKernelGGL(mirror_rows_kernel, dim3(blocks_for(total)), dim3(256), 0, st,
                           static_cast<const float2*>(in), nx, half, rows, p->work);
        // unnormalised inverse, then 1/n (a non-power-of-two inverse carries
        // its own 1/n as well, as in the reference)
        const float inner = pow2(nx) || nx == 1 ? 1.0f : 1.0f / nx;
        rc = rows_fft(p->work, rows, nx, p->cx, true, inner, p->work, p->any_ws, st);
        if (!rc)
            hipLaunchKernelGGL(real_part_kernel, dim3(blocks_for(total)), dim3(256), 0, st, p->work, total,
                               1.0f / nx, static_cast<float*>(out));
        break;
    }
    default: return FFT_INVALID_VALUE;
    }
    if (!rc && hipGetLastError() != hipSuccess) rc = ZKX_RUNTIME_ERROR;
    return to_result(rc);
}

FFTResult store_freqs(float* dst, const std::vector<float>& v)
{
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, dst) == hipSuccess && attr.type == hipMemoryTypeDevice) {
        if (hipMemcpy(dst, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
            return FFT_EXEC_FAILED;
        return FFT_SUCCESS;
    }
    (void)hipGetLastError();
    for (size_t i = 0; i < v.size(); ++i) dst[i] = v[i];
    return FFT_SUCCESS;
}

void destroy(Plan* p)
{
    if (!p) return;
    free_chirp(p->cx);
    free_chirp(p->cy);
    if (p->work) (void)hipFree(p->work);
    if (p->any_ws) (void)hipFree(p->any_ws);
    delete p;
}

}  // namespace

extern "C" {

FFTResult fft_plan_create_1d(FFTPlanHandle* plan, int nx, int batch, FFTType type, FFTDirection direction)
{
    if (!plan) return FFT_INVALID_VALUE;
    *plan = nullptr;
    if (batch < 1 || !length_ok(nx)) return FFT_INVALID_VALUE;
    if (type != FFT_TYPE_C2C && type != FFT_TYPE_R2C && type != FFT_TYPE_C2R) return FFT_INVALID_VALUE;
    if (direction != FFT_FORWARD && direction != FFT_INVERSE) return FFT_INVALID_VALUE;
    Plan* p = new (std::nothrow) Plan();
    if (!p) return FFT_ALLOC_FAILED;
    p->nx = nx;
    p->batch = batch;
    p->type = type;
    p->dir = direction;
    FFTResult r = make_chirp(nx, p->cx);
    if (r == FFT_SUCCESS && type != FFT_TYPE_C2C) {
        p->work_elems = static_cast<size_t>(nx) * batch;
        if (hipMalloc(reinterpret_cast<void**>(&p->work), sizeof(float2) * p->work_elems) != hipSuccess)
            r = FFT_ALLOC_FAILED;
    }
    // build the twiddle tables and scratch now so execution never allocates
    const float2* tw = nullptr;
    if (r == FFT_SUCCESS && pow2(nx) && nx > 1 && zkx::fft_any_prepare(nx) != ZKX_SUCCESS) r = FFT_ALLOC_FAILED;
    if (r == FFT_SUCCESS && p->cx.M && zkx::twiddles(p->cx.M, &tw) != ZKX_SUCCESS) r = FFT_ALLOC_FAILED;
    if (r == FFT_SUCCESS) r = alloc_any_ws(p, zkx::fft_any_workspace_bytes(batch, nx));
    if (r != FFT_SUCCESS) {
        destroy(p);
        return r;
    }
    *plan = p;
    return FFT_SUCCESS;
}

FFTResult fft_plan_create_2d(FFTPlanHandle* plan, int nx, int ny, FFTType type, FFTDirection direction)
{
    if (!plan) return FFT_INVALID_VALUE;
    *plan = nullptr;
    if (!length_ok(nx) || !length_ok(ny)) return FFT_INVALID_VALUE;
    if (type != FFT_TYPE_C2C && type != FFT_TYPE_R2C && type != FFT_TYPE_C2R) return FFT_INVALID_VALUE;
    if (direction != FFT_FORWARD && direction != FFT_INVERSE) return FFT_INVALID_VALUE;
    Plan* p = new (std::nothrow) Plan();
    if (!p) return FFT_ALLOC_FAILED;
    p->nx = nx;
    p->ny = ny;
    p->type = type;
    p->dir = direction;
    FFTResult r = make_chirp(nx, p->cx);
    if (r == FFT_SUCCESS) r = make_chirp(ny, p->cy);
    p->work_elems = static_cast<size_t>(nx) * ny;
    if (r == FFT_SUCCESS && hipMalloc(reinterpret_cast<void**>(&p->work), sizeof(float2) * p->work_elems) != hipSuccess)
        r = FFT_ALLOC_FAILED;
    for (int n : {nx, ny, p->cx.M, p->cy.M})
        if (r == FFT_SUCCESS && n > 1 && pow2(n) && zkx::fft_any_prepare(n) != ZKX_SUCCESS) r = FFT_ALLOC_FAILED;
    if (r == FFT_SUCCESS) {
        const size_t a = zkx::fft_any_workspace_bytes(ny, nx), b = zkx::fft_any_workspace_bytes(nx, ny);
        r = alloc_any_ws(p, a > b ? a : b);
    }
    if (r != FFT_SUCCESS) {
        destroy(p);
        return r;
    }
    *plan = p;
    return FFT_SUCCESS;
}

FFTResult fft_execute(FFTPlanHandle plan, void* output, const void* input)
{
    return run_plan(plan, output, input, nullptr);
}

FFTResult fft_execute_on_stream(FFTPlanHandle plan, void* output, const void* input, void* stream)
{
    return run_plan(plan, output, input, static_cast<hipStream_t>(stream));
}

FFTResult fft_execute_inplace(FFTPlanHandle plan, void* data)
{
    if (!plan) return FFT_INVALID_VALUE;
    const Plan* p = static_cast<const Plan*>(plan);
    // R2C/C2R change the element type, so they only run out of place
    if (p->ny == 0 && p->type != FFT_TYPE_C2C) return FFT_INVALID_VALUE;
    return run_plan(plan, data, data, nullptr);
}

FFTResult fft_plan_destroy(FFTPlanHandle plan)
{
    if (!plan) return FFT_INVALID_VALUE;
    destroy(static_cast<Plan*>(plan));
    return FFT_SUCCESS;
}

FFTResult fft_shift(void* output, const void* input, int nx, int batch)
{
    if (!output || !input || nx <= 0 || batch <= 0) return FFT_INVALID_VALUE;
    hipLaunchKernelGGL(shift1d_kernel, dim3(blocks_for(static_cast<long long>(nx) * batch)), dim3(256), 0, nullptr,
                       static_cast<const float2*>(input), static_cast<float2*>(output), nx, nx / 2, batch);
    return hipGetLastError() == hipSuccess ? FFT_SUCCESS : FFT_EXEC_FAILED;
}

FFTResult fft_inverse_shift(void* output, const void* input, int nx, int batch)
{
    return fft_shift(output, input, nx, batch);
}

FFTResult fft_shift_2d(void* output, const void* input, int nx, int ny)
{
    if (!output || !input || nx <= 0 || ny <= 0) return FFT_INVALID_VALUE;
    hipLaunchKernelGGL(shift2d_kernel, dim3(blocks_for(static_cast<long long>(nx) * ny)), dim3(256), 0, nullptr,
                       static_cast<const float2*>(input), static_cast<float2*>(output), nx, ny, nx / 2, ny / 2);
    return hipGetLastError() == hipSuccess ? FFT_SUCCESS : FFT_EXEC_FAILED;
}

FFTResult fft_inverse_shift_2d(void* output, const void* input, int nx, int ny)
{
    if (!output || !input || nx <= 0 || ny <= 0) return FFT_INVALID_VALUE;
    hipLaunchKernelGGL(shift2d_kernel, dim3(blocks_for(static_cast<long long>(nx) * ny)), dim3(256), 0, nullptr,
                       static_cast<const float2*>(input), static_cast<float2*>(output), nx, ny, (nx + 1) / 2,
                       (ny + 1) / 2);
    return hipGetLastError() == hipSuccess ? FFT_SUCCESS : FFT_EXEC_FAILED;
}

const char* fft_error_string(FFTResult result)
{
    switch (result) {
    case FFT_SUCCESS: return "Success";
    case FFT_INVALID_VALUE: return "Invalid value";
    case FFT_ALLOC_FAILED: return "Allocation failed";
    case FFT_EXEC_FAILED: return "Execution failed";
    case FFT_INTERNAL_ERROR: return "Internal error";
    default: return "Unknown error";
    }
}

int fft_version(void) { return 1; }

FFTResult fft_freq(float* freqs, int n, float d)
{
    if (!freqs || n <= 0) return FFT_INVALID_VALUE;
    std::vector<float> v(static_cast<size_t>(n));
    const float inv_n_d = 1.0f / (static_cast<float>(n) * d);
    const int num_pos = (n + 1) / 2;
    for (int i = 0; i < num_pos; ++i) v[i] = static_cast<float>(i) * inv_n_d;
    for (int i = 0; i < n - num_pos; ++i) v[num_pos + i] = -static_cast<float>(n - num_pos - i) * inv_n_d;
    return store_freqs(freqs, v);
}

FFTResult fft_rfftfreq(float* freqs, int n, float d)
{
    if (!freqs || n <= 0) return FFT_INVALID_VALUE;
    const int len = n / 2 + 1;
    std::vector<float> v(static_cast<size_t>(len));
    const float inv_n_d = 1.0f / (static_cast<float>(n) * d);
    for (int i = 0; i < len; ++i) v[i] = static_cast<float>(i) * inv_n_d;
    return store_freqs(freqs, v);
}

}  // extern "C"
