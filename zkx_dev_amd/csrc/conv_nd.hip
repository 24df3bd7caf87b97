// N-dimensional FFT convolution engine (declared in conv_nd.hpp) and the
// fftconvolve entry points over N-D axes.
//
// fftconvolve follows cuSignal (ZKX/cusignal-23.08.00/python/cusignal/
// convolution/convolve.py:159-320, convolution_utils.py:40-100):
//   * in1 and in2 have one rank; the convolved axes default to all; on the
//     other axes the extents must be equal or 1 (broadcast);
//   * full extent: s1 + s2 - 1 on a convolved axis, max(s1, s2) elsewhere;
//   * 'valid' swaps the operands when in1 is not at least as large as in2
//     in every dimension (and fails when neither is); its extent is
//     s1 - s2 + 1 on a convolved axis, the full extent elsewhere;
//   * 'same' keeps in1's extents in every dimension;
//   * the kept block is centred: start = (full - kept) / 2 per dimension;
//   * a real result for two real operands, complex otherwise.
// cuSignal pads each axis to a 5-smooth length for FFTPACK; this engine pads
// to a power of two, which changes only the rounding of the result.
#include "conv_nd.hpp"
#include "fft_any.hpp"
#include "runtime.hpp"

#include <algorithm>

namespace zkx {
namespace {

long long next_pow2_ll(long long n)
{
    long long p = 1;
    while (p < n) p <<= 1;
    return p;
}

size_t al256(size_t b) { return (b + 255) / 256 * 256; }

long long count_of(const int* d, int rank)
{
    long long c = 1;
    for (int k = 0; k < rank; ++k) c *= d[k];
    return c;
}

// out[b][c][r] = in[b][r][c] for b < B, r < R, c < C; 32 x 32 tiles through
// LDS (one padding column keeps the column reads on distinct banks).  The
// tile index runs over blockIdx.x (either extent may be large), the batch
// over blockIdx.y.
__global__ void __launch_bounds__(256)
nd_transpose_kernel(const float2* __restrict__ in, long long B, int R, int C, float2* __restrict__ out)
{
    __shared__ float2 tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int tiles_c = (C + 31) / 32;
    const int r0 = static_cast<int>(blockIdx.x / tiles_c) * 32, c0 = static_cast<int>(blockIdx.x % tiles_c) * 32;
    const long long plane = static_cast<long long>(R) * C;
    for (long long b = blockIdx.y; b < B; b += gridDim.y) {
        const float2* src = in + b * plane;
        float2* dst = out + b * plane;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int r = r0 + ty + 8 * k, c = c0 + tx;
            if (r < R && c < C) tile[ty + 8 * k][tx] = src[static_cast<long long>(r) * C + c];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = c0 + ty + 8 * k, r = r0 + tx;
            if (r < R && c < C) dst[static_cast<long long>(c) * R + r] = tile[tx][ty + 8 * k];
        }
        __syncthreads();
    }
}

int transpose(const float2* in, long long B, int R, int C, float2* out, hipStream_t st)
{
    const long long tiles = static_cast<long long>((R + 31) / 32) * ((C + 31) / 32);
    if (tiles > 0x7fffffffLL) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: transpose extent too large");
    const dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(std::min<long long>(B, 65535)));
    hipLaunchKernelGGL(nd_transpose_kernel, grid, dim3(256), 0, st, in, B, R, C, out);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

// Y (extents g) = X1 (g1) * X2 (g2), broadcasting extents of 1.  Y may be X1
// when g1 == g.
struct NdMul {
    int rank;
    int g[kMaxNdRank];
    long long s1[kMaxNdRank], s2[kMaxNdRank];  // strides of X1 / X2 per dim of g (0 when broadcast)
};

__global__ void __launch_bounds__(256)
nd_mul_kernel(const float2* X1, const float2* __restrict__ X2, NdMul p, float2* Y)
{
    const int last = p.rank - 1;
    const int gl = p.g[last];
    const int j = blockIdx.x * 256 + threadIdx.x;
    const long long rows = nd_rows(p.g, p.rank);
    for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
        long long q = r, o1 = 0, o2 = 0;
        for (int d = last - 1; d >= 0; --d) {
            const long long c = q % p.g[d];
            q /= p.g[d];
            o1 += c * p.s1[d];
            o2 += c * p.s2[d];
        }
        if (j < gl) Y[r * gl + j] = cmul(X1[o1 + j * p.s1[last]], X2[o2 + j * p.s2[last]]);
    }
}

// Transform X (extents g) along dimension d; T: scratch of X's size.
int fft_along(float2* X, const int* g, int rank, int d, bool inverse, float2* T, char* any_ws, hipStream_t st)
{
    const int n = g[d];
    if (n == 1) return ZKX_SUCCESS;
    long long outer = 1, inner = 1;
    for (int k = 0; k < d; ++k) outer *= g[k];
    for (int k = d + 1; k < rank; ++k) inner *= g[k];
    const float scale = inverse ? 1.0f / static_cast<float>(n) : 1.0f;
    if (inner == 1) return fft_any_rows(X, outer, n, inverse, scale, X, any_ws, st);
    if (inner > 0x7fffffffLL) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: extent too large");
    int rc = transpose(X, outer, n, static_cast<int>(inner), T, st);
    if (rc) return rc;
    rc = fft_any_rows(T, outer * inner, n, inverse, scale, T, any_ws, st);
    if (rc) return rc;
    return transpose(T, outer, static_cast<int>(inner), n, X, st);
}

int fft_axes(float2* X, const int* g, int rank, unsigned axes, bool inverse, float2* T, char* any_ws,
             hipStream_t st)
{
    for (int d = rank - 1; d >= 0; --d) {
        if (!(axes >> d & 1u)) continue;
        int rc = fft_along(X, g, rank, d, inverse, T, any_ws, st);
        if (rc) return rc;
    }
    return ZKX_SUCCESS;
}

}  // namespace

int nd_conv_plan(const int* s1, const int* s2, int rank, unsigned axes, NdConvPlan* p)
{
    if (rank < 1 || rank > kMaxNdRank) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: rank must be 1..8");
    NdConvPlan q;
    q.rank = rank;
    q.axes = axes;
    size_t any = 0;
    for (int d = 0; d < rank; ++d) {
        if (s1[d] < 1 || s2[d] < 1) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: empty extent");
        if (axes >> d & 1u) {
            const long long f = next_pow2_ll(static_cast<long long>(s1[d]) + s2[d] - 1);
            if (f > kMaxAnyFftPow2) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: axis too long");
            q.g1[d] = q.g2[d] = q.g[d] = static_cast<int>(f);
        } else {
            if (s1[d] != s2[d] && s1[d] != 1 && s2[d] != 1)
                ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: incompatible shapes for in1 and in2");
            q.g1[d] = s1[d];
            q.g2[d] = s2[d];
            q.g[d] = std::max(s1[d], s2[d]);
        }
    }
    q.n1 = count_of(q.g1, rank);
    q.n2 = count_of(q.g2, rank);
    q.n = count_of(q.g, rank);
    if (q.n1 > kMaxNdElements || q.n2 > kMaxNdElements || q.n > kMaxNdElements)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: padded operands exceed 2^31 elements");
    q.y_is_x1 = q.n1 == q.n;
    for (int d = 0; d < rank; ++d) {
        if (!(axes >> d & 1u)) continue;
        any = std::max(any, fft_any_workspace_bytes(std::max(q.n1, std::max(q.n2, q.n)) / q.g[d], q.g[d]));
    }
    const long long big = std::max(q.n1, std::max(q.n2, q.n));
    size_t off = 0;
    q.off_x1 = off;
    off += al256(sizeof(float2) * q.n1);
    q.off_x2 = off;
    off += al256(sizeof(float2) * q.n2);
    q.off_y = q.y_is_x1 ? q.off_x1 : off;
    if (!q.y_is_x1) off += al256(sizeof(float2) * q.n);
    q.off_t = off;
    off += al256(sizeof(float2) * big);
    q.off_any = off;
    off += any;
    q.bytes = off;
    *p = q;
    return ZKX_SUCCESS;
}

int nd_conv_run(const NdConvPlan& p, char* ws, hipStream_t st)
{
    for (int d = 0; d < p.rank; ++d) {
        if (!(p.axes >> d & 1u)) continue;
        int rc = fft_any_prepare(p.g[d]);
        if (rc) return rc;
    }
    float2* X1 = reinterpret_cast<float2*>(ws + p.off_x1);
    float2* X2 = reinterpret_cast<float2*>(ws + p.off_x2);
    float2* Y = reinterpret_cast<float2*>(ws + p.off_y);
    float2* T = reinterpret_cast<float2*>(ws + p.off_t);
    char* any = ws + p.off_any;
    int rc = fft_axes(X1, p.g1, p.rank, p.axes, false, T, any, st);
    if (rc) return rc;
    rc = fft_axes(X2, p.g2, p.rank, p.axes, false, T, any, st);
    if (rc) return rc;
    NdMul m{};
    m.rank = p.rank;
    long long st1[kMaxNdRank], st2[kMaxNdRank];
    nd_strides(p.g1, p.rank, st1);
    nd_strides(p.g2, p.rank, st2);
    for (int d = 0; d < p.rank; ++d) {
        m.g[d] = p.g[d];
        m.s1[d] = p.g1[d] == 1 && p.g[d] > 1 ? 0 : st1[d];
        m.s2[d] = p.g2[d] == 1 && p.g[d] > 1 ? 0 : st2[d];
    }
    const long long rows = p.n / p.g[p.rank - 1];
    hipLaunchKernelGGL(nd_mul_kernel, nd_grid(rows, p.g[p.rank - 1]), dim3(256), 0, st, X1, X2, m, Y);
    ZKX_LAUNCH_CHECK();
    return fft_axes(Y, p.g, p.rank, p.axes, true, T, any, st);
}

}  // namespace zkx

using namespace zkx;

namespace {

thread_local ThreadWorkspace t_convnd_ws;

struct FftconvShape {
    int rank = 0;
    int s1[kMaxNdRank] = {}, s2[kMaxNdRank] = {};
    int full[kMaxNdRank] = {}, out[kMaxNdRank] = {};
    bool swap = false;
};

// Extents of fftconvolve's result (convolve.py:254-320 of cuSignal).
int fftconv_shape(const long long* shape1, const long long* shape2, int rank, unsigned axes, int mode,
                  FftconvShape* r)
{
    if (rank < 1 || rank > kMaxNdRank) ZKX_FAIL(ZKX_INVALID_SHAPE, "fftconvolve: rank must be 1..8");
    if (!shape1 || !shape2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve: null shape");
    if (mode < 0 || mode > 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "acceptable mode flags are 'valid', 'same', or 'full'");
    if (axes == 0 || (rank < 32 && (axes >> rank) != 0))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve: axes must name dimensions of the inputs");
    FftconvShape q;
    q.rank = rank;
    for (int d = 0; d < rank; ++d) {
        if (shape1[d] < 1 || shape2[d] < 1 || shape1[d] > 0x7fffffffLL || shape2[d] > 0x7fffffffLL)
            ZKX_FAIL(ZKX_INVALID_SHAPE, "fftconvolve: extents must be 1 .. 2^31-1");
        q.s1[d] = static_cast<int>(shape1[d]);
        q.s2[d] = static_cast<int>(shape2[d]);
        if (!(axes >> d & 1u) && q.s1[d] != q.s2[d] && q.s1[d] != 1 && q.s2[d] != 1)
            ZKX_FAIL(ZKX_INVALID_SHAPE, "incompatible shapes for in1 and in2");
        const long long f = (axes >> d & 1u) ? shape1[d] + shape2[d] - 1 : std::max(shape1[d], shape2[d]);
        if (f > 0x7fffffffLL) ZKX_FAIL(ZKX_INVALID_SHAPE, "fftconvolve: extent too large");
        q.full[d] = static_cast<int>(f);
    }
    if (mode == 2) {
        bool ok1 = true, ok2 = true;
        for (int d = 0; d < rank; ++d) {
            ok1 = ok1 && q.s1[d] >= q.s2[d];
            ok2 = ok2 && q.s2[d] >= q.s1[d];
        }
        if (!ok1 && !ok2)
            ZKX_FAIL(ZKX_INVALID_SHAPE, "For 'valid' mode, one must be at least as large as the other in every dimension");
        q.swap = !ok1;
    }
    for (int d = 0; d < rank; ++d) {
        const int a = q.swap ? q.s2[d] : q.s1[d], b = q.swap ? q.s1[d] : q.s2[d];
        if (mode == 0) q.out[d] = q.full[d];
        else if (mode == 1) q.out[d] = q.s1[d];
        else q.out[d] = (axes >> d & 1u) ? a - b + 1 : q.full[d];
    }
    *r = q;
    return ZKX_SUCCESS;
}

}  // namespace

extern "C" {

int zkx_fftconvolve_nd_shape(const long long* shape1, const long long* shape2, int rank, unsigned axes_mask, int mode,
                             long long* out_shape)
{
    clear_last_error();
    if (!out_shape) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve: null output shape");
    FftconvShape q;
    int rc = fftconv_shape(shape1, shape2, rank, axes_mask, mode, &q);
    if (rc) return rc;
    for (int d = 0; d < rank; ++d) out_shape[d] = q.out[d];
    return ZKX_SUCCESS;
}

int zkx_fftconvolve_nd(const void* in1, const long long* shape1, int complex1, const void* in2,
                       const long long* shape2, int complex2, int rank, unsigned axes_mask, int mode, void* out,
                       void* stream)
{
    clear_last_error();
    hipStream_t st = (hipStream_t)stream;
    FftconvShape q;
    int rc = fftconv_shape(shape1, shape2, rank, axes_mask, mode, &q);
    if (rc) return rc;
    if (!in1 || !in2 || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve: null pointer");
    // one complex kernel row (leading extents 1, at most 4096 taps) against
    // every row of a complex in1 along the last axis: overlap-save rows,
    // without padding each row to a power of two
    bool rows_form = complex1 && complex2 && axes_mask == (1u << (rank - 1)) && !q.swap && q.s2[rank - 1] <= 4096;
    long long lead = 1;
    for (int d = 0; d + 1 < rank; ++d) {
        rows_form = rows_form && q.s2[d] == 1;
        lead *= q.s1[d];
    }
    if (rows_form && lead <= 0x7fffffffLL)
        return zkx_fftconvolve_c64(static_cast<const zkx_complex64*>(in1), static_cast<int>(lead), q.s1[rank - 1],
                                   static_cast<const zkx_complex64*>(in2), q.s2[rank - 1], mode,
                                   static_cast<zkx_complex64*>(out), stream);
    // convolution commutes: 'valid' takes the larger operand first
    const void* a = q.swap ? in2 : in1;
    const void* b = q.swap ? in1 : in2;
    const int ca = q.swap ? complex2 : complex1, cb = q.swap ? complex1 : complex2;
    int sa[kMaxNdRank], sb[kMaxNdRank];
    for (int d = 0; d < rank; ++d) {
        sa[d] = q.swap ? q.s2[d] : q.s1[d];
        sb[d] = q.swap ? q.s1[d] : q.s2[d];
    }
    NdConvPlan p;
    rc = nd_conv_plan(sa, sb, rank, axes_mask, &p);
    if (rc) return rc;
    void* wsv = nullptr;
    WorkspaceLease lease_(t_convnd_ws, st);
    rc = lease_.get(p.bytes, &wsv);
    if (rc) return rc;
    char* ws = static_cast<char*>(wsv);
    auto pad = [&](const void* src, int cplx, const int* s, const int* g, float2* X) -> int {
        NdPad pp{};
        pp.rank = rank;
        for (int d = 0; d < rank; ++d) {
            pp.g[d] = g[d];
            pp.s[d] = s[d];
        }
        nd_strides(s, rank, pp.sstride);
        pp.reverse = 0;
        const dim3 grid = nd_grid(count_of(g, rank) / g[rank - 1], g[rank - 1]);
        if (cplx)
            hipLaunchKernelGGL(nd_pad_kernel<NdLoadC64>, grid, dim3(256), 0, st,
                               NdLoadC64{static_cast<const float2*>(src)}, pp, X);
        else
            hipLaunchKernelGGL(nd_pad_kernel<NdLoadF32>, grid, dim3(256), 0, st,
                               NdLoadF32{static_cast<const float*>(src)}, pp, X);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    };
    if ((rc = pad(a, ca, sa, p.g1, reinterpret_cast<float2*>(ws + p.off_x1)))) return rc;
    if ((rc = pad(b, cb, sb, p.g2, reinterpret_cast<float2*>(ws + p.off_x2)))) return rc;
    if ((rc = nd_conv_run(p, ws, st))) return rc;
    // centred crop of the full result, which sits at the origin of Y
    NdCrop c{};
    c.rank = rank;
    nd_strides(p.g, rank, c.ystride);
    for (int d = 0; d < rank; ++d) {
        c.o[d] = q.out[d];
        c.start[d] = (q.full[d] - q.out[d]) / 2;
    }
    const float2* Y = reinterpret_cast<const float2*>(ws + p.off_y);
    const dim3 grid = nd_grid(count_of(q.out, rank) / q.out[rank - 1], q.out[rank - 1]);
    if (complex1 || complex2)
        hipLaunchKernelGGL(nd_crop_kernel<NdStoreC64>, grid, dim3(256), 0, st, Y, c,
                           NdStoreC64{static_cast<float2*>(out)});
    else
        hipLaunchKernelGGL(nd_crop_kernel<NdStoreRealF32>, grid, dim3(256), 0, st, Y, c,
                           NdStoreRealF32{static_cast<float*>(out)});
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // extern "C"
