// N-dimensional linear convolution by power-of-two FFTs on gfx950, shared by
// fftconvolve over N-D axes and correlate over N-D inputs.
//
// Arrays are row-major complex64.  A plan fixes, per dimension d, the
// spectral extent of in1 (g1), in2 (g2) and of their product (g): on a
// convolved axis all three are F[d] = next power of two >= s1[d] + s2[d] - 1,
// so the circular convolution of the zero-padded operands is the linear one;
// on any other axis g1/g2 are the operands' own extents (equal, or one of
// them 1 and broadcast) and g their maximum.  The caller fills X1 and X2 with
// nd_pad_kernel, nd_conv_run transforms both along the convolved axes,
// multiplies them (broadcasting) and inverse transforms the product into Y
// (scaled by 1/F per axis), and the caller crops Y with nd_crop_kernel.
//
// An axis whose elements are not contiguous is brought to the rows of a
// scratch array by a tiled transpose, transformed by the row FFT (one LDS
// pass per row up to 8192 points, four-step beyond), and transposed back.
// Kernels walk arrays by rows (all dimensions but the last): a block decodes
// its row index once, with wave-uniform integer arithmetic, and its threads
// cover the contiguous last dimension.
#pragma once

#include "zkx_hip_common.hpp"

#include <cstddef>

namespace zkx {

constexpr int kMaxNdRank = 8;
constexpr long long kMaxNdElements = (1LL << 31) - 1;

// Extents and element strides of one row-major array (strides of the
// array read through it; 0 for a broadcast dimension).
struct NdDims {
    int rank = 0;
    int d[kMaxNdRank] = {};
    long long stride[kMaxNdRank] = {};
};

struct NdConvPlan {
    int rank = 0;
    unsigned axes = 0;  // bit d set: dimension d is convolved
    int g1[kMaxNdRank] = {}, g2[kMaxNdRank] = {}, g[kMaxNdRank] = {};
    long long n1 = 0, n2 = 0, n = 0;  // elements of X1, X2, Y
    bool y_is_x1 = false;             // the product overwrites X1 (no broadcast into it)
    size_t off_x1 = 0, off_x2 = 0, off_y = 0, off_t = 0, off_any = 0, bytes = 0;
};

// Plan for operands of extents s1, s2 (rank 1..8, every extent >= 1, the
// non-convolved extents equal or 1).  Fails with INVALID_SHAPE when an
// array or transform would be too large for the device indexing used here.
int nd_conv_plan(const int* s1, const int* s2, int rank, unsigned axes, NdConvPlan* p);

// Spectral part on a workspace of p.bytes bytes whose X1 / X2 regions the
// caller has filled: Y (at ws + p.off_y) = the linear convolution, extents
// p.g, over the convolved axes.
int nd_conv_run(const NdConvPlan& p, char* ws, hipStream_t st);

// Loader for nd_pad_kernel: element i of the source as complex.
struct NdLoadC64 {
    const float2* p;
    __device__ float2 operator()(long long i) const { return p[i]; }
};
struct NdLoadF32 {
    const float* p;
    __device__ float2 operator()(long long i) const { return make_float2(p[i], 0.0f); }
};

// X (extents g) = src (extents s, row-major) zero padded; with `reverse`
// every dimension of the source is read backwards (index s-1-c), which turns
// a convolution into a correlation.
struct NdPad {
    int rank;
    int g[kMaxNdRank];
    int s[kMaxNdRank];
    long long sstride[kMaxNdRank];
    int reverse;
};

__device__ __forceinline__ long long nd_rows(const int* ext, int rank)
{
    long long r = 1;
    for (int d = 0; d + 1 < rank; ++d) r *= ext[d];
    return r;
}

template <class Load>
__global__ void __launch_bounds__(256) nd_pad_kernel(Load ld, NdPad p, float2* __restrict__ X)
{
    const int last = p.rank - 1;
    const int gl = p.g[last], sl = p.s[last];
    const int j = blockIdx.x * 256 + threadIdx.x;
    const long long rows = nd_rows(p.g, p.rank);
    for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
        // leading coordinates of row r (uniform across the block)
        long long q = r, src = 0;
        bool inside = true;
        for (int d = last - 1; d >= 0; --d) {
            const int c = static_cast<int>(q % p.g[d]);
            q /= p.g[d];
            inside = inside && c < p.s[d];
            src += static_cast<long long>(p.reverse ? p.s[d] - 1 - c : c) * p.sstride[d];
        }
        if (j >= gl) continue;
        float2 v = make_float2(0.f, 0.f);
        if (inside && j < sl) v = ld(src + (p.reverse ? sl - 1 - j : j));
        X[r * gl + j] = v;
    }
}

// out (extents o) = Y (extents g) from corner `start`; Store(i, value).
struct NdCrop {
    int rank;
    int o[kMaxNdRank];
    long long ystride[kMaxNdRank];
    int start[kMaxNdRank];
};

struct NdStoreC64 {
    float2* p;
    __device__ void operator()(long long i, float2 v) const { p[i] = v; }
};
struct NdStoreRealF32 {
    float* p;
    __device__ void operator()(long long i, float2 v) const { p[i] = v.x; }
};

template <class Store>
__global__ void __launch_bounds__(256) nd_crop_kernel(const float2* __restrict__ Y, NdCrop p, Store st)
{
    const int last = p.rank - 1;
    const int ol = p.o[last];
    const int j = blockIdx.x * 256 + threadIdx.x;
    const long long rows = nd_rows(p.o, p.rank);
    for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
        long long q = r, src = static_cast<long long>(p.start[last]);
        for (int d = last - 1; d >= 0; --d) {
            const int c = static_cast<int>(q % p.o[d]);
            q /= p.o[d];
            src += static_cast<long long>(p.start[d] + c) * p.ystride[d];
        }
        if (j < ol) st(r * ol + j, Y[src + j]);
    }
}

// Launch shape of the row kernels over `rows` rows of `last` elements.
inline dim3 nd_grid(long long rows, int last)
{
    const unsigned gx = static_cast<unsigned>((last + 255) / 256);
    const unsigned gy = static_cast<unsigned>(rows < 65535 ? (rows > 0 ? rows : 1) : 65535);
    return dim3(gx, gy);
}

// Row-major strides of extents d[0..rank).
inline void nd_strides(const int* d, int rank, long long* stride)
{
    long long s = 1;
    for (int k = rank - 1; k >= 0; --k) {
        stride[k] = s;
        s *= d[k];
    }
}

}  // namespace zkx
