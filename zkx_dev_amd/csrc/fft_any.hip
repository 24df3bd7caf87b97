// Any-length batched complex transforms (see fft_any.hpp).  The reference's
// backend takes any length (ZKX/cusignal_cpp/src/backends/fft/fft_thrust.cu:
// radix-4 stages for powers of two, Bluestein with m = next_pow2(2n) for the
// rest, :14,27,722-764); here the large lengths run as global passes over the
// LDS row kernels.
#include "fft_any.hpp"
#include "runtime.hpp"

#include <cmath>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace zkx {
int fft_rows_dispatch(int n, const float2* x, long long rows, bool inverse, float scale, float2* y,
                      hipStream_t st);

namespace {

constexpr int kLdsMax = 8192;

bool is_pow2_ll(long long n) { return n > 0 && (n & (n - 1)) == 0; }
long long next_pow2_ll(long long n)
{
    long long m = 1;
    while (m < n) m <<= 1;
    return m;
}
size_t al256(size_t b) { return (b + 255) / 256 * 256; }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// out[b][c][r] = in[b][r][c] * (twiddle ? W^(r*c) : 1) for in [batch][rows][cols],
// W = exp(-+2 pi i / n), through 32x32 LDS tiles; (r*c) mod n is exact in
// 64-bit and the angle is evaluated in FP64.
template <bool TWIDDLE>
__global__ void __launch_bounds__(256) transpose_batched_kernel(const float2* __restrict__ in, int rows, int cols,
                                                                long long n, bool inverse, float2* __restrict__ out)
{
    __shared__ float2 tile[32][33];
    const long long mat = static_cast<long long>(rows) * cols;
    const float2* src = in + blockIdx.z * mat;
    float2* dst = out + blockIdx.z * mat;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        const int r = r0 + j, c = c0 + tx;
        if (r < rows && c < cols) {
            float2 v = src[static_cast<long long>(r) * cols + c];
            if constexpr (TWIDDLE) {
                const long long m = (static_cast<long long>(r) * c) % n;
                double s, co;
                sincospi(2.0 * static_cast<double>(m) / static_cast<double>(n), &s, &co);
                const float2 w = make_float2(static_cast<float>(co), static_cast<float>(inverse ? s : -s));
                v = cmul(v, w);
            }
            tile[j][tx] = v;
        }
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int c = c0 + j, r = r0 + tx;
        if (c < cols && r < rows) dst[static_cast<long long>(c) * rows + r] = tile[tx][j];
    }
}

// a[r][k] = x[r][k] * chirp[k] (conj for the inverse) for k < n, 0 up to M
__global__ void __launch_bounds__(256) chirp_pad_kernel(const float2* __restrict__ x, long long rows, int n, int M,
                                                        const float2* __restrict__ chirp, bool inverse,
                                                        float2* __restrict__ a)
{
    const long long total = rows * M;
    for (long long g = blockIdx.x * 256LL + threadIdx.x; g < total; g += static_cast<long long>(gridDim.x) * 256) {
        const long long r = g / M;
        const int k = static_cast<int>(g - r * M);
        float2 v = make_float2(0.f, 0.f);
        if (k < n) {
            const float2 xv = x[r * n + k];
            v = inverse ? cmulc(xv, chirp[k]) : cmul(xv, chirp[k]);
        }
        a[g] = v;
    }
}

// a[r][k] *= spec[k]
__global__ void __launch_bounds__(256) spec_mul_kernel(float2* __restrict__ a, long long rows, int M,
                                                       const float2* __restrict__ spec)
{
    const long long total = rows * M;
    for (long long g = blockIdx.x * 256LL + threadIdx.x; g < total; g += static_cast<long long>(gridDim.x) * 256)
        a[g] = cmul(a[g], spec[g % M]);
}

// y[r][k] = scale * a[r][k] * chirp[k] (conj for the inverse), k < n
__global__ void __launch_bounds__(256) chirp_out_kernel(const float2* __restrict__ a, long long rows, int n, int M,
                                                        const float2* __restrict__ chirp, bool inverse, float scale,
                                                        float2* __restrict__ y)
{
    const long long total = rows * n;
    for (long long g = blockIdx.x * 256LL + threadIdx.x; g < total; g += static_cast<long long>(gridDim.x) * 256) {
        const long long r = g / n;
        const int k = static_cast<int>(g - r * n);
        const float2 v = a[r * M + k];
        y[g] = cscale(inverse ? cmulc(v, chirp[k]) : cmul(v, chirp[k]), scale);
    }
}

// y = scale * x (1-point transforms)
__global__ void __launch_bounds__(256) scale_kernel(const float2* __restrict__ x, long long total, float scale,
                                                    float2* __restrict__ y)
{
    for (long long g = blockIdx.x * 256LL + threadIdx.x; g < total; g += static_cast<long long>(gridDim.x) * 256)
        y[g] = cscale(x[g], scale);
}

// p[0][k] = a[k] (k < la), p[1][k] = b[k] (k < lb), zeros up to M
__global__ void __launch_bounds__(256) pad_pair_kernel(const float2* __restrict__ a, long long la,
                                                       const float2* __restrict__ b, long long lb, long long M,
                                                       float2* __restrict__ p)
{
    for (long long g = blockIdx.x * 256LL + threadIdx.x; g < 2 * M; g += static_cast<long long>(gridDim.x) * 256) {
        const long long k = g < M ? g : g - M;
        const float2 z = make_float2(0.f, 0.f);
        p[g] = g < M ? (k < la ? a[k] : z) : (k < lb ? b[k] : z);
    }
}

// p[0][k] *= p[1][k]
__global__ void __launch_bounds__(256) pair_mul_kernel(float2* __restrict__ p, long long M)
{
    for (long long g = blockIdx.x * 256LL + threadIdx.x; g < M; g += static_cast<long long>(gridDim.x) * 256)
        p[g] = cmul(p[g], p[M + g]);
}

unsigned grid_for(long long total)
{
    const long long b = (total + 255) / 256;
    return static_cast<unsigned>(b < 1 ? 1 : (b > 1048576 ? 1048576 : b));
}

// ---------------------------------------------------------------------------
// Bluestein tables, per (device, n)
// ---------------------------------------------------------------------------
struct ChirpTables {
    int M = 0;
    float2* chirp = nullptr;     // [n]  e^{-i pi k^2 / n}
    float2* spec_fwd = nullptr;  // [M]  FFT of the forward kernel conj(chirp), wrapped
    float2* spec_inv = nullptr;  // [M]  FFT of the inverse kernel chirp, wrapped
};
std::mutex g_chirp_mu;
std::map<std::pair<int, int>, ChirpTables> g_chirps;

int pow2_rows(const float2* x, long long rows, long long n, bool inverse, float scale, float2* y, char* ws,
              hipStream_t st);

int chirp_tables(int n, const ChirpTables** out)
{
    int dev = 0;
    ZKX_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_chirp_mu);
    auto it = g_chirps.find({dev, n});
    if (it != g_chirps.end()) {
        *out = &it->second;
        return ZKX_SUCCESS;
    }
    ChirpTables c;
    c.M = static_cast<int>(next_pow2_ll(2LL * n - 1));
    std::vector<float2> ch(n), bf(c.M, make_float2(0.f, 0.f)), bi(c.M, make_float2(0.f, 0.f));
    for (int k = 0; k < n; ++k) {
        // k^2 mod 2n exactly, so the angle pi*q/n stays small
        const long long q = (static_cast<long long>(k) * k) % (2LL * n);
        const double th = M_PI * static_cast<double>(q) / static_cast<double>(n);
        const float cs = static_cast<float>(std::cos(th)), sn = static_cast<float>(std::sin(th));
        ch[k] = make_float2(cs, -sn);
        bf[k] = make_float2(cs, sn);
        bi[k] = make_float2(cs, -sn);
        if (k > 0) {
            bf[c.M - k] = bf[k];
            bi[c.M - k] = bi[k];
        }
    }
    const size_t tab = al256(sizeof(float2) * n) + 2 * al256(sizeof(float2) * c.M);
    char* base = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&base), tab) != hipSuccess) {
        (void)hipGetLastError();
        ZKX_FAIL(ZKX_OUT_OF_MEMORY, "fft: chirp table allocation failed");
    }
    c.chirp = reinterpret_cast<float2*>(base);
    c.spec_fwd = reinterpret_cast<float2*>(base + al256(sizeof(float2) * n));
    c.spec_inv = reinterpret_cast<float2*>(base + al256(sizeof(float2) * n) + al256(sizeof(float2) * c.M));
    char* ws = nullptr;
    const size_t wsb = fft_any_workspace_bytes(1, c.M);
    int rc = ZKX_SUCCESS;
    if (wsb && hipMalloc(reinterpret_cast<void**>(&ws), wsb) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(base);
        ZKX_FAIL(ZKX_OUT_OF_MEMORY, "fft: chirp workspace allocation failed");
    }
    if (hipMemcpy(c.chirp, ch.data(), sizeof(float2) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c.spec_fwd, bf.data(), sizeof(float2) * c.M, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c.spec_inv, bi.data(), sizeof(float2) * c.M, hipMemcpyHostToDevice) != hipSuccess)
        rc = ZKX_RUNTIME_ERROR;
    if (!rc) rc = pow2_rows(c.spec_fwd, 1, c.M, false, 1.0f, c.spec_fwd, ws, nullptr);
    if (!rc) rc = pow2_rows(c.spec_inv, 1, c.M, false, 1.0f, c.spec_inv, ws, nullptr);
    if (!rc && hipStreamSynchronize(nullptr) != hipSuccess) rc = ZKX_RUNTIME_ERROR;
    if (ws) (void)hipFree(ws);
    if (rc) {
        (void)hipFree(base);
        ZKX_FAIL(rc, "fft: chirp spectra failed");
    }
    auto ins = g_chirps.emplace(std::make_pair(dev, n), c);
    *out = &ins.first->second;
    return ZKX_SUCCESS;
}

// ---------------------------------------------------------------------------
// passes
// ---------------------------------------------------------------------------
// split n = n1 * n2, both powers of two <= 8192, n1 >= n2
void four_step_split(long long n, int* n1, int* n2)
{
    int k = 0;
    while ((1LL << k) < n) ++k;
    const int k2 = k / 2;
    *n2 = 1 << k2;
    *n1 = 1 << (k - k2);
}

int transpose(const float2* in, long long batch, int rows, int cols, long long n, bool twiddle, bool inverse,
              float2* out, hipStream_t st)
{
    // grid.z carries the batch (<= 65535 per launch)
    for (long long b0 = 0; b0 < batch; b0 += 65535) {
        const unsigned nb = static_cast<unsigned>(batch - b0 < 65535 ? batch - b0 : 65535);
        const long long off = b0 * static_cast<long long>(rows) * cols;
        dim3 grid((cols + 31) / 32, (rows + 31) / 32, nb);
        if (twiddle)
            hipLaunchKernelGGL(transpose_batched_kernel<true>, grid, dim3(256), 0, st, in + off, rows, cols, n,
                               inverse, out + off);
        else
            hipLaunchKernelGGL(transpose_batched_kernel<false>, grid, dim3(256), 0, st, in + off, rows, cols, n,
                               inverse, out + off);
        ZKX_LAUNCH_CHECK();
    }
    return ZKX_SUCCESS;
}

// power-of-two rows: one LDS pass up to 8192, else four-step through ws
// (2 * rows * n complex)
int pow2_rows(const float2* x, long long rows, long long n, bool inverse, float scale, float2* y, char* ws,
              hipStream_t st)
{
    if (n <= kLdsMax) return fft_rows_dispatch(static_cast<int>(n), x, rows, inverse, scale, y, st);
    int n1 = 0, n2 = 0;
    four_step_split(n, &n1, &n2);
    float2* t1 = reinterpret_cast<float2*>(ws);
    float2* t2 = reinterpret_cast<float2*>(ws + al256(sizeof(float2) * rows * n));
    // x as [rows][n1][n2]: columns j2 become rows of n1 points
    int rc = transpose(x, rows, n1, n2, n, false, inverse, t1, st);
    if (!rc) rc = fft_rows_dispatch(n1, t1, rows * n2, inverse, 1.0f, t1, st);
    // A[j2][k1] * W_n^(j2 k1), back to [k1][j2]
    if (!rc) rc = transpose(t1, rows, n2, n1, n, true, inverse, t2, st);
    if (!rc) rc = fft_rows_dispatch(n2, t2, rows * n1, inverse, scale, t2, st);
    // B[k1][k2] holds X[k1 + n1 k2]
    if (!rc) rc = transpose(t2, rows, n1, n2, n, false, inverse, y, st);
    return rc;
}

}  // namespace

bool fft_any_length_ok(long long n)
{
    if (n < 1) return false;
    return is_pow2_ll(n) ? n <= kMaxAnyFftPow2 : n <= kMaxAnyFftOther;
}

size_t fft_any_workspace_bytes(long long rows, int n)
{
    if (n <= 1) return 0;
    if (is_pow2_ll(n)) return n <= kLdsMax ? 0 : 2 * al256(sizeof(float2) * rows * n);
    const long long M = next_pow2_ll(2LL * n - 1);
    return al256(sizeof(float2) * rows * M) + (M > kLdsMax ? 2 * al256(sizeof(float2) * rows * M) : 0);
}

int fft_any_prepare(int n)
{
    if (!fft_any_length_ok(n)) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft: unsupported length");
    const float2* tw = nullptr;
    if (n <= 1) return ZKX_SUCCESS;
    auto tables = [&](long long m) -> int {
        if (m <= kLdsMax) return twiddles(static_cast<int>(m), &tw);
        int n1 = 0, n2 = 0;
        four_step_split(m, &n1, &n2);
        int rc = twiddles(n1, &tw);
        return rc ? rc : twiddles(n2, &tw);
    };
    if (is_pow2_ll(n)) return tables(n);
    const ChirpTables* c = nullptr;
    int rc = tables(next_pow2_ll(2LL * n - 1));
    return rc ? rc : chirp_tables(n, &c);
}

size_t fft_convolve_workspace_bytes(long long la, long long lb)
{
    const long long M = next_pow2_ll(la + lb - 1);
    return al256(sizeof(float2) * 2 * M) + fft_any_workspace_bytes(2, static_cast<int>(M));
}

int fft_convolve_full(const float2* a, long long la, const float2* b, long long lb, float2* y, void* ws,
                      hipStream_t st)
{
    if (la < 1 || lb < 1) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: empty operand");
    const long long full = la + lb - 1;
    const long long M = next_pow2_ll(full);
    if (M > kMaxAnyFftPow2) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "fft convolve: operands too long");
    int rc = fft_any_prepare(static_cast<int>(M));
    if (rc) return rc;
    float2* p = static_cast<float2*>(ws);
    char* inner = static_cast<char*>(ws) + al256(sizeof(float2) * 2 * M);
    hipLaunchKernelGGL(pad_pair_kernel, dim3(grid_for(2 * M)), dim3(256), 0, st, a, la, b, lb, M, p);
    ZKX_LAUNCH_CHECK();
    if ((rc = pow2_rows(p, 2, M, false, 1.0f, p, inner, st))) return rc;
    hipLaunchKernelGGL(pair_mul_kernel, dim3(grid_for(M)), dim3(256), 0, st, p, M);
    ZKX_LAUNCH_CHECK();
    if ((rc = pow2_rows(p, 1, M, true, 1.0f / static_cast<float>(M), p, inner, st))) return rc;
    ZKX_HIP_CHECK(hipMemcpyAsync(y, p, sizeof(float2) * full, hipMemcpyDeviceToDevice, st));
    return ZKX_SUCCESS;
}

int fft_any_rows(const float2* x, long long rows, int n, bool inverse, float scale, float2* y, void* ws,
                 hipStream_t st)
{
    if (!fft_any_length_ok(n) || rows < 0) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft: unsupported length");
    if (rows == 0) return ZKX_SUCCESS;
    if (n == 1) {
        hipLaunchKernelGGL(scale_kernel, dim3(grid_for(rows)), dim3(256), 0, st, x, rows, scale, y);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    }
    char* w = static_cast<char*>(ws);
    if (fft_any_workspace_bytes(rows, n) && !w) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fft: workspace required");
    if (is_pow2_ll(n)) return pow2_rows(x, rows, n, inverse, scale, y, w, st);
    const ChirpTables* c = nullptr;
    int rc = chirp_tables(n, &c);
    if (rc) return rc;
    const int M = c->M;
    float2* a = reinterpret_cast<float2*>(w);
    char* inner = w + al256(sizeof(float2) * rows * M);
    hipLaunchKernelGGL(chirp_pad_kernel, dim3(grid_for(rows * M)), dim3(256), 0, st, x, rows, n, M, c->chirp,
                       inverse, a);
    ZKX_LAUNCH_CHECK();
    rc = pow2_rows(a, rows, M, false, 1.0f, a, inner, st);
    if (rc) return rc;
    hipLaunchKernelGGL(spec_mul_kernel, dim3(grid_for(rows * M)), dim3(256), 0, st, a, rows, M,
                       inverse ? c->spec_inv : c->spec_fwd);
    ZKX_LAUNCH_CHECK();
    rc = pow2_rows(a, rows, M, true, 1.0f, a, inner, st);
    if (rc) return rc;
    hipLaunchKernelGGL(chirp_out_kernel, dim3(grid_for(rows * n)), dim3(256), 0, st, a, rows, n, M, c->chirp,
                       inverse, scale / static_cast<float>(M), y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // namespace zkx
