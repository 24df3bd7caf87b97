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
// Every transform runs on rows (one LDS pass per row up to 8192 points,
// four-step beyond): the forward part transforms the last dimension, then
// rotates it to the front (a tiled transpose, [outer][last] -> [last][outer])
// until the lowest convolved dimension has been last; the product is taken
// in that rotated order and the inverse undoes the rotations, so a 2-D
// convolution costs three transposes in all.
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
    int s1[kMaxNdRank] = {}, s2[kMaxNdRank] = {};  // the operands' own extents
    int g1[kMaxNdRank] = {}, g2[kMaxNdRank] = {}, g[kMaxNdRank] = {};
    long long n1 = 0, n2 = 0, n = 0;  // elements of X1, X2, Y
    bool y_is_x1 = false;             // the product overwrites X1 (no broadcast into it)
    // rows of dimension 0 the caller reads from Y, [keep_lo0, keep_hi0) (both
    // 0: all); with rank >= 2 the last inverse rotation and transform skip
    // the others, so Y holds only those rows
    int keep_lo0 = 0, keep_hi0 = 0;
    size_t off_x1 = 0, off_x2 = 0, off_y = 0, off_t = 0, off_any = 0, off_scal = 0, bytes = 0;  // four equal buffers, FFT scratch, scalars (2 words + block maxima)
};

// Rows (of the last dimension, row-major over the others) of a padded
// operand of extents g that can hold data: its dimension-0 coordinate is
// below s[0], the operand's own extent (rank 1: the one row).
inline long long nd_live_rows(const int* s, const int* g, int rank)
{
    if (rank < 2) return 1;
    long long r = s[0];
    for (int d = 1; d + 1 < rank; ++d) r *= g[d];
    return r;
}

// Rows of the padded operand (extents g) the pad kernels must write: the
// forward part reads only the live rows when a rotation follows the first
// transform (rows past them are read as zeros), else all rows.
inline long long nd_pad_rows(const NdConvPlan& p, const int* s, const int* g)
{
    int lo = 0;
    while (!(p.axes >> lo & 1u)) ++lo;
    long long all = 1;
    for (int d = 0; d + 1 < p.rank; ++d) all *= g[d];
    return p.rank - 1 - lo >= 1 ? nd_live_rows(s, g, p.rank) : all;
}

// Plan for operands of extents s1, s2 (rank 1..8, every extent >= 1, the
// non-convolved extents equal or 1).  Fails with INVALID_SHAPE when an
// array or transform would be too large for the device indexing used here.
int nd_conv_plan(const int* s1, const int* s2, int rank, unsigned axes, NdConvPlan* p);

// Spectral part on a workspace of p.bytes bytes whose X1 / X2 regions the
// caller has filled: *y (one of the workspace's buffers) = the linear
// convolution over the convolved axes, extents p.g in the original order.
int nd_conv_run(const NdConvPlan& p, char* ws, float2** y, hipStream_t st);

// The same for two real operands packed into X1 as Z = A + i s B (requires
// g1 == g2; nd_pack_real_pair does the packing): one forward transform, then
// A^ B^ = (Z[k]^2 - conj(Z[-k])^2) / (4 i s), k negated along the convolved
// axes only; *y holds the real convolution in its real parts.
int nd_conv_run_real_pair(const NdConvPlan& p, char* ws, float2** y, hipStream_t st);

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
    long long rows;  // leading rows of X written (nd_pad_rows)
};

__device__ __forceinline__ long long nd_rows(const int* ext, int rank)
{
    long long r = 1;
    for (int d = 0; d + 1 < rank; ++d) r *= ext[d];
    return r;
}

// A block covers kNdSpan elements of a row (kNdPer per thread, 256 apart).
constexpr int kNdPer = 4;
constexpr int kNdSpan = 256 * kNdPer;

template <class Load>
__global__ void __launch_bounds__(256) nd_pad_kernel(Load ld, NdPad p, float2* __restrict__ X)
{
    const int last = p.rank - 1;
    const int gl = p.g[last], sl = p.s[last];
    const int j0 = blockIdx.x * kNdSpan + threadIdx.x;
    const long long rows = p.rows;
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
        float2* row = X + r * gl;
#pragma unroll
        for (int k = 0; k < kNdPer; ++k) {
            const int j = j0 + k * 256;
            if (j >= gl) break;
            float2 v = make_float2(0.f, 0.f);
            if (inside && j < sl) v = ld(src + (p.reverse ? sl - 1 - j : j));
            row[j] = v;
        }
    }
}

// Two real operands of one padded extent g packed into one complex array,
// Z = A + i s B (B optionally read backwards in every dimension), so that one
// forward transform serves both.  s is the power of two that brings max|B|
// to the binade of max|A|: the product A^ B^ is recovered from Z by a
// difference of squares, whose rounding error scales with |A^|^2 + |s B^|^2,
// so balanced operands keep it at the level of a plain product; a power of
// two scales exactly.  absmax[0..1] hold max|A|, max|B| as float bits.
struct NdPad2 {
    int rank;
    int g[kMaxNdRank];
    int sa[kMaxNdRank], sb[kMaxNdRank];
    long long sastride[kMaxNdRank], sbstride[kMaxNdRank];
    int reverse_b;
    const unsigned* absmax;
    long long rows;  // leading rows of Z written (nd_pad_rows)
};

// log2 of the balancing scale s from the two maxima (0 when either is 0).
__device__ __forceinline__ int nd_pair_shift(const unsigned* absmax)
{
    const float ma = __uint_as_float(absmax[0]), mb = __uint_as_float(absmax[1]);
    if (!(ma > 0.0f) || !(mb > 0.0f)) return 0;
    int ea, eb;
    (void)frexpf(ma, &ea);
    (void)frexpf(mb, &eb);
    const int d = ea - eb;  // kept within +-60 so that 1/(4s) stays a normal float
    return d > 60 ? 60 : (d < -60 ? -60 : d);
}

// Block maxima of |re| of two real sources: part[2 b + k] for block b and
// operand k (A: 0, B: 1).  Same-address atomics serialize at tens of ns
// each, so the blocks store partial maxima and nd_absmax2_reduce_kernel
// combines them.
constexpr int kNdAbsBlocks = 512;
template <class LoadA, class LoadB>
__global__ void __launch_bounds__(256)
nd_absmax2_kernel(LoadA la, long long na, LoadB lb, long long nb, float* __restrict__ part)
{
    float m[2] = {0.0f, 0.0f};
    // kAbsU independent loads per thread and step keep enough reads in flight
    constexpr int kAbsU = 8;
    const long long stride = static_cast<long long>(gridDim.x) * 256 * kAbsU;
    for (long long i0 = blockIdx.x * 256LL * kAbsU + threadIdx.x; i0 < na; i0 += stride) {
        float v[kAbsU];
#pragma unroll
        for (int u = 0; u < kAbsU; ++u) {
            const long long i = i0 + u * 256LL;
            v[u] = i < na ? fabsf(la(i).x) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kAbsU; ++u) m[0] = fmaxf(m[0], v[u]);
    }
    for (long long i0 = blockIdx.x * 256LL * kAbsU + threadIdx.x; i0 < nb; i0 += stride) {
        float v[kAbsU];
#pragma unroll
        for (int u = 0; u < kAbsU; ++u) {
            const long long i = i0 + u * 256LL;
            v[u] = i < nb ? fabsf(lb(i).x) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kAbsU; ++u) m[1] = fmaxf(m[1], v[u]);
    }
    __shared__ float wmax[2][4];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        float v = m[k];
        for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
        if ((threadIdx.x & 63) == 0) wmax[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 2)
        part[2 * blockIdx.x + threadIdx.x] = fmaxf(fmaxf(wmax[threadIdx.x][0], wmax[threadIdx.x][1]),
                                                   fmaxf(wmax[threadIdx.x][2], wmax[threadIdx.x][3]));
}

// absmax[k] = float bits of the maximum of part[2 b + k] over `blocks` blocks
// (one block of 256 threads).
static __global__ void __launch_bounds__(256)
nd_absmax2_reduce_kernel(const float* __restrict__ part, int blocks, unsigned* __restrict__ absmax)
{
    float m[2] = {0.0f, 0.0f};
    for (int b = threadIdx.x; b < blocks; b += 256) {
        m[0] = fmaxf(m[0], part[2 * b]);
        m[1] = fmaxf(m[1], part[2 * b + 1]);
    }
    __shared__ float wmax[2][4];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        float v = m[k];
        for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
        if ((threadIdx.x & 63) == 0) wmax[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 2)
        absmax[threadIdx.x] = __float_as_uint(fmaxf(fmaxf(wmax[threadIdx.x][0], wmax[threadIdx.x][1]),
                                                    fmaxf(wmax[threadIdx.x][2], wmax[threadIdx.x][3])));
}

template <class LoadA, class LoadB>
__global__ void __launch_bounds__(256) nd_pad2_kernel(LoadA la, LoadB lb, NdPad2 p, float2* __restrict__ Z)
{
    const int last = p.rank - 1;
    const int gl = p.g[last], sla = p.sa[last], slb = p.sb[last];
    const int shift = nd_pair_shift(p.absmax);
    const int j0 = blockIdx.x * kNdSpan + threadIdx.x;
    const long long rows = p.rows;
    for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
        long long q = r, oa = 0, ob = 0;
        bool ina = true, inb = true;
        for (int d = last - 1; d >= 0; --d) {
            const int c = static_cast<int>(q % p.g[d]);
            q /= p.g[d];
            ina = ina && c < p.sa[d];
            inb = inb && c < p.sb[d];
            oa += static_cast<long long>(c) * p.sastride[d];
            ob += static_cast<long long>(p.reverse_b ? p.sb[d] - 1 - c : c) * p.sbstride[d];
        }
        float2* row = Z + r * gl;
#pragma unroll
        for (int k = 0; k < kNdPer; ++k) {
            const int j = j0 + k * 256;
            if (j >= gl) break;
            const float a = (ina && j < sla) ? la(oa + j).x : 0.0f;
            const float b = (inb && j < slb) ? lb(ob + (p.reverse_b ? slb - 1 - j : j)).x : 0.0f;
            row[j] = make_float2(a, ldexpf(b, shift));
        }
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
    const int j0 = blockIdx.x * kNdSpan + threadIdx.x;
    const long long rows = nd_rows(p.o, p.rank);
    for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
        long long q = r, src = static_cast<long long>(p.start[last]);
        for (int d = last - 1; d >= 0; --d) {
            const int c = static_cast<int>(q % p.o[d]);
            q /= p.o[d];
            src += static_cast<long long>(p.start[d] + c) * p.ystride[d];
        }
#pragma unroll
        for (int k = 0; k < kNdPer; ++k) {
            const int j = j0 + k * 256;
            if (j >= ol) break;
            st(r * ol + j, Y[src + j]);
        }
    }
}

// Launch shape of the pad and crop kernels over `rows` rows of `last`
// elements (kNdSpan per block along a row).
inline dim3 nd_grid(long long rows, int last)
{
    const unsigned gx = static_cast<unsigned>((last + kNdSpan - 1) / kNdSpan);
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

// Pack A (extents sa) and B (extents sb, read backwards when reverse_b) into
// X1 of the plan's workspace, with the balancing scale.
template <class LoadA, class LoadB>
int nd_pack_real_pair(const NdConvPlan& p, char* ws, LoadA la, const int* sa, LoadB lb, const int* sb,
                      int reverse_b, hipStream_t st)
{
    const int rank = p.rank;
    unsigned* absmax = reinterpret_cast<unsigned*>(ws + p.off_scal);
    long long na = 1, nb = 1;
    NdPad2 pp{};
    pp.rank = rank;
    int smax[kMaxNdRank];
    for (int d = 0; d < rank; ++d) {
        na *= sa[d];
        nb *= sb[d];
        pp.g[d] = p.g1[d];
        pp.sa[d] = sa[d];
        pp.sb[d] = sb[d];
        smax[d] = sa[d] > sb[d] ? sa[d] : sb[d];
    }
    const long long rows = nd_pad_rows(p, smax, p.g1);
    pp.rows = rows;
    nd_strides(sa, rank, pp.sastride);
    nd_strides(sb, rank, pp.sbstride);
    pp.reverse_b = reverse_b;
    pp.absmax = absmax;
    float* part = reinterpret_cast<float*>(ws + p.off_scal + 256);
    const long long nmax = na > nb ? na : nb;
    // two blocks per CU at most, striding over the operands 2048 elements at a time
    const long long want = (nmax + 2047) / 2048;
    const int blocks = static_cast<int>(want < kNdAbsBlocks ? want : kNdAbsBlocks);
    hipLaunchKernelGGL((nd_absmax2_kernel<LoadA, LoadB>), dim3(blocks), dim3(256), 0, st, la, na, lb, nb, part);
    ZKX_LAUNCH_CHECK();
    hipLaunchKernelGGL(nd_absmax2_reduce_kernel, dim3(1), dim3(256), 0, st, part, blocks, absmax);
    ZKX_LAUNCH_CHECK();
    hipLaunchKernelGGL((nd_pad2_kernel<LoadA, LoadB>), nd_grid(rows, p.g1[rank - 1]), dim3(256), 0, st, la, lb, pp,
                       reinterpret_cast<float2*>(ws + p.off_x1));
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // namespace zkx
